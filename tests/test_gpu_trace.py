"""GPU parity of closest-hit traversal against the oracle.

The bar: hit ids equal on >= 99.99% of rays and t within 1e-5 relative on
agreeing hits (both are IEEE binary32 restatements of the same operations,
so in practice they agree exactly; the exact agreement is reported).
"""
import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu

ID_AGREEMENT = 0.9999
T_RTOL = 1e-5


def _compare(prim_g, t_g, prim_o, t_o):
    agree = prim_g == prim_o
    frac = float(agree.mean()) if agree.size else 1.0
    hit = agree & (prim_o != sa.MISS)
    rel = np.abs(t_g[hit].astype(np.float64) - t_o[hit]) / np.maximum(np.abs(t_o[hit].astype(np.float64)), 1e-30)
    assert frac >= ID_AGREEMENT, f"hit-id agreement {frac:.6f}"
    assert rel.size == 0 or rel.max() <= T_RTOL, f"max rel t error {rel.max():.3g}"
    return frac, float(np.mean(t_g[agree].view(np.uint32) == t_o[agree].view(np.uint32))) if agree.any() else 1.0


def _build(ctx, w):
    scene = sa.FaceSoup(ctx, w.scene)
    B = sa.FastTwoLevelGridBuilder if w.kind == 1 else sa.FastSimpleGridBuilder
    b = B(ctx, **w.build)
    b.init(scene)
    return scene, b.build()


def test_uniform_soup10k_primary(ctx, oracle):
    w = W.soup10k()
    og = oracle.build_uniform(w.scene, **w.oracle)
    _, g = _build(ctx, w)
    prim, t, st = g.trace_host(w.rays, with_stats=True)
    po, to, so = og.trace(w.rays, with_stats=True)
    frac, exact_t = _compare(prim, t, po, to)
    assert frac == 1.0 and exact_t == 1.0
    np.testing.assert_array_equal(st, so)


def test_twolevel_heightfield2m_primary_and_bounce(ctx, oracle):
    w = W.heightfield2m()
    og = oracle.build_twolevel(w.scene, **w.oracle)
    scene, g = _build(ctx, w)
    rb = sa.RayBatch(ctx, w.rays)
    stats = sa.DeviceBuffer(ctx, 16 * rb.n)
    prim_d, t_d = g.trace(rb, stats=stats)
    prim = prim_d.download(np.uint32, rb.n)
    t = t_d.download(np.float32, rb.n)
    st = stats.download(np.uint32, 4 * rb.n).reshape(rb.n, 4)
    sub = np.arange(0, rb.n, 7)
    po, to, so = og.trace(np.ascontiguousarray(w.rays[:, sub]), with_stats=True)
    frac, exact_t = _compare(prim[sub], t[sub], po, to)
    assert frac == 1.0 and exact_t == 1.0
    np.testing.assert_array_equal(st[sub], so)

    # bounce rays generated on the device == oracle bounce rays, bit for bit
    out = sa.RayBatch(ctx, capacity=rb.n)
    src = sa.DeviceBuffer(ctx, 4 * rb.n)
    count = sa.bounce(ctx, scene, rb, prim_d, t_d, w.bounce_seed, out, src=src)
    m = int(count.download(np.uint32, 1)[0])
    assert m == int((prim != sa.MISS).sum())
    b_g = out.download(m)
    src_g = src.download(np.uint32, m)
    b_o, src_o = oracle.bounce(w.scene, w.rays, prim, t, w.bounce_seed)
    np.testing.assert_array_equal(src_g, src_o)
    np.testing.assert_array_equal(b_g.view(np.uint32), b_o.view(np.uint32))
    # and traced on the device with the ray count read from device memory
    out.n = m
    prim2_d = sa.DeviceBuffer(ctx, 4 * rb.n)
    t2_d = sa.DeviceBuffer(ctx, 4 * rb.n)
    g.trace(out, prim=prim2_d, t=t2_d, d_count=count.ptr, n=rb.n)
    prim2 = prim2_d.download(np.uint32, m)
    t2 = t2_d.download(np.float32, m)
    sub2 = np.arange(0, m, 5)
    po2, to2 = og.trace(np.ascontiguousarray(b_o[:, sub2]))
    frac2, _ = _compare(prim2[sub2], t2[sub2], po2, to2)
    assert frac2 == 1.0


def test_fp64_referee_small_scene(ctx, oracle):
    """FP64 brute force over all triangles is the referee on a small scene."""
    soa = sa.gen_soup(11, 2000, 0.03)
    rays = sa.gen_random_rays(12, 4000, 0.0, 1.0)
    for kind in (0, 1):
        scene = sa.FaceSoup(ctx, soa)
        B = sa.FastTwoLevelGridBuilder if kind else sa.FastSimpleGridBuilder
        b = B(ctx, lam=5.0, lambda_leaf=2.0)
        b.init(scene)
        g = b.build()
        prim, t = g.trace_host(rays)
        p64, t64 = oracle.brute_f64(soa, rays)
        agree = np.mean(prim == p64)
        assert agree >= 0.995, agree
        hit = (prim == p64) & (p64 != sa.MISS)
        assert np.max(np.abs(t[hit] - t64[hit]) / t64[hit]) < 1e-4


def test_tie_scene_first_candidate_wins(ctx, oracle):
    """IntersectorTest's scene (src/Test/IntersectorTest.h:179-258): triangle 0
    in plane z=-2, triangles 1..399 identical in plane y=-4; a camera at the
    origin looking down -y sees triangle 0 where it covers the view and
    triangle 1 (first of the equal-t duplicates) elsewhere, or a miss."""
    n = 400
    soa = np.zeros((9, n), np.float32)
    # triangle 0: plane z = -2, large
    soa[:, 0] = [-3, -10, -2, 3, -10, -2, 0, 10, -2]
    # 1..399: plane y = -4
    for i in range(1, n):
        soa[:, i] = [-6, -4, -6, 6, -4, -6, 0, -4, 6]
    rays = sa.gen_camera_rays((0, 0, 0), (0, -1, 0), (0, 0, 1), 66.0, 64, 64)
    for kind in (0, 1):
        scene = sa.FaceSoup(ctx, soa)
        B = sa.FastTwoLevelGridBuilder if kind else sa.FastSimpleGridBuilder
        b = B(ctx)
        b.init(scene)
        g = b.build()
        prim, t = g.trace_host(rays)
        assert set(np.unique(prim)).issubset({0, 1, sa.MISS})
        assert (prim == 1).any()
        og = (oracle.build_twolevel if kind else oracle.build_uniform)(soa, top_divisor=6 if kind else 1)
        po, to = og.trace(rays)
        np.testing.assert_array_equal(prim, po)


def test_special_rays(ctx, oracle):
    """axis-parallel directions (rcp = +-inf), origins inside the grid, tmax
    shorter than the hit and tmax <= 0"""
    soa = sa.gen_soup(21, 5000, 0.02)
    rng = np.random.default_rng(5)
    n = 3000
    rays = np.zeros((7, n), np.float32)
    rays[0:3] = rng.uniform(-0.2, 1.2, (3, n))
    axis = rng.integers(0, 3, n)
    sign = rng.choice([-1.0, 1.0], n)
    for i in range(n):
        rays[3 + axis[i], i] = sign[i]
    rays[6] = np.float32(3.402823466e38)
    rays[6, : n // 10] = rng.uniform(0.0, 0.3, n // 10)
    rays[6, n // 10: n // 10 + 20] = 0.0
    rays[6, n // 10 + 20: n // 10 + 40] = -1.0
    for kind in (0, 1):
        scene = sa.FaceSoup(ctx, soa)
        B = sa.FastTwoLevelGridBuilder if kind else sa.FastSimpleGridBuilder
        b = B(ctx, lambda_leaf=2.0)
        b.init(scene)
        g = b.build()
        prim, t = g.trace_host(rays)
        og = (oracle.build_twolevel if kind else oracle.build_uniform)(soa, top_divisor=6 if kind else 1,
                                                                      lambda_leaf=2.0)
        po, to = og.trace(rays)
        np.testing.assert_array_equal(prim, po)
        np.testing.assert_array_equal(t.view(np.uint32), to.view(np.uint32))
        assert (prim[n // 10: n // 10 + 40] == sa.MISS).all()


def _shadow_segments(prim, t, rays, light):
    hit = prim != sa.MISS
    o = rays[:3, hit] + rays[3:6, hit] * (t[hit] - np.float32(1e-4))
    seg = np.empty((7, o.shape[1]), np.float32)
    seg[:3] = o
    seg[3:6] = np.asarray(light, np.float32)[:, None] - o
    seg[6] = 1.0
    return seg


@pytest.mark.parametrize("which", ["soup10k", "heightfield"])
def test_shadow_any_hit(ctx, oracle, which):
    """sg_trace_shadow == oracle any-hit traversal, bit for bit, on segments
    from the primary hits to a point light; occlusion agrees with the
    closest hit along each segment."""
    if which == "soup10k":
        w = W.soup10k()
        og = oracle.build_uniform(w.scene, **w.oracle)
        light = (0.5, 0.5, -2.0)
    else:
        w = W.heightfield2m(grid=256, res_x=480, res_y=270)
        og = oracle.build_twolevel(w.scene, **w.oracle)
        light = (-0.5, 0.5, 0.35)
    _, g = _build(ctx, w)
    prim, t = g.trace_host(w.rays)
    seg = _shadow_segments(prim, t, w.rays, light)
    assert seg.shape[1] > 1000
    ps, ts, ss = g.trace_host(seg, with_stats=True, shadow=True)
    po, to, so = og.trace(seg, with_stats=True, shadow=True)
    np.testing.assert_array_equal(ps, po)
    np.testing.assert_array_equal(ts.view(np.uint32), to.view(np.uint32))
    np.testing.assert_array_equal(ss, so)
    pc, tc = g.trace_host(seg)
    np.testing.assert_array_equal(ts < 0.9999, tc < 0.9999)
    occ = float((ts < 0.9999).mean())
    assert 0.01 < occ < 0.99, occ


def test_cluster_paths_depth4(ctx, oracle):
    """dense-cluster scene (500k tiny + 2k large triangles), diffuse paths of
    four segments chained on the device (bounce counts never read back):
    every segment's rays, ids and t equal the oracle's trace / bounce
    composition, bit for bit"""
    w = W.cluster(res_x=160, res_y=90)
    og = oracle.build_twolevel(w.scene, **w.oracle)
    scene, g = _build(ctx, w)
    rb = sa.RayBatch(ctx, w.rays)
    counts, segs = sa.trace_paths(ctx, scene, g, rb, depth=4, seed=w.bounce_seed, keep=True)
    ctx.sync()
    rays_o = w.rays
    for d, (rays_d, prim_d, t_d, _) in enumerate(segs):
        m = rays_o.shape[1]
        if d:
            assert int(counts[d - 1].download(np.uint32, 1)[0]) == m
            np.testing.assert_array_equal(rays_d.download(m).view(np.uint32), rays_o.view(np.uint32))
        po, to = og.trace(rays_o)
        np.testing.assert_array_equal(prim_d.download(np.uint32, m), po, err_msg=f"segment {d}")
        np.testing.assert_array_equal(t_d.download(np.float32, m).view(np.uint32), to.view(np.uint32))
        if d < 3:
            rays_o, _ = oracle.bounce(w.scene, rays_o, po, to, w.bounce_seed + d)
            if rays_o.shape[1] == 0:
                break


@pytest.mark.parametrize("kind", [0, 1])
@pytest.mark.parametrize("shadow", [False, True])
@pytest.mark.parametrize("count", [1, 63, 777, 5_000, 40_000, 150_000, 250_000])
def test_device_count_small_batches(ctx, oracle, kind, shadow, count):
    """a device-side ray count far below the capacity: the batch is traced
    with fewer rays per wave (8 to 32; 64 for the largest), lanes without a
    ray taking part in the wave-shared tests; ids, t and the untouched
    outputs beyond the count are checked"""
    scene = sa.gen_soup(90 + kind, 20_000, 0.02)
    cap = 400_000
    r7 = sa.gen_random_rays(91, cap, 0.0, 1.0)
    if shadow:
        r7[6] = 1.0
    soup = sa.FaceSoup(ctx, scene)
    B = sa.FastTwoLevelGridBuilder if kind else sa.FastSimpleGridBuilder
    b = B(ctx, lam=5.0) if kind == 0 else B(ctx, lam=5.0, lambda_leaf=2.0)
    b.init(soup)
    g = b.build()
    rb = sa.RayBatch(ctx, r7)
    d_n = sa.DeviceBuffer(ctx, 4)
    d_n.upload(np.array([count], np.uint32))
    prim = sa.DeviceBuffer(ctx, 4 * cap)
    t = sa.DeviceBuffer(ctx, 4 * cap)
    prim.upload(np.full(cap, 0xABCDEF01, np.uint32))
    g.trace(rb, prim=prim, t=t, d_count=d_n.ptr, n=cap, shadow=shadow)
    og = oracle.build_twolevel(scene, lam=5.0, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, lam=5.0,
                                                                                                   top_divisor=1)
    po, to = og.trace(np.ascontiguousarray(r7[:, :count]), shadow=shadow)
    pg = prim.download(np.uint32, cap)
    np.testing.assert_array_equal(pg[:count], po)
    np.testing.assert_array_equal(t.download(np.float32, count).view(np.uint32), to.view(np.uint32))
    assert (pg[count:count + 4096] == 0xABCDEF01).all()


ORDER_CASES = [
    ("heightfield_exit", lambda: W.heightfield2m(grid=256, res_x=320, res_y=200), "exit"),
    ("heightfield_axis", lambda: W.heightfield2m(grid=256, res_x=320, res_y=200), "axis"),
    ("cluster_exit", lambda: W.cluster(n_small=60_000, n_large=400, res_x=320, res_y=180), "exit"),
]


@pytest.mark.parametrize("name,make,order", ORDER_CASES, ids=[c[0] for c in ORDER_CASES])
def test_bounce_ordered_equals_oracle(ctx, oracle, name, make, order):
    """sg_bounce_ordered: the rays, their sources and their count equal the
    oracle's ordered bounce bit for bit (exit distance from the grid's
    bounds, longest first; or |d . up| smallest first)"""
    w = make()
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    rb = sa.RayBatch(ctx, w.rays)
    prim_d, t_d = g.trace(rb)
    ctx.sync()
    prim, t = prim_d.download(np.uint32, rb.n), t_d.download(np.float32, rb.n)
    info = g.info()
    box = (tuple(info.bmin), tuple(info.bmax))
    q = sa.bounce_order(order, grid=g, axis=(0.0, 0.0, 1.0))
    out = sa.RayBatch(ctx, capacity=rb.n)
    src = sa.DeviceBuffer(ctx, 4 * rb.n)
    count = sa.bounce(ctx, scene, rb, prim_d, t_d, w.bounce_seed, out, src=src, order=q)
    ctx.sync()
    m = int(count.download(np.uint32, 1)[0])
    b_o, src_o = oracle.bounce(w.scene, w.rays, prim, t, w.bounce_seed, order=order, box=box, axis=(0.0, 0.0, 1.0))
    assert m == b_o.shape[1] > 1000
    np.testing.assert_array_equal(out.download(m).view(np.uint32), b_o.view(np.uint32))
    np.testing.assert_array_equal(src.download(np.uint32, m), src_o)
    g.close()
    scene.close()


def test_paths_with_ordered_bounces_equal_oracle(ctx, oracle):
    """four path segments chained on the device with exit-ordered bounces
    (counts never read back): every segment's rays, ids and t equal the
    oracle's"""
    w = W.cluster(n_small=60_000, n_large=400, res_x=128, res_y=72)
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    info = g.info()
    box = (tuple(info.bmin), tuple(info.bmax))
    rb = sa.RayBatch(ctx, w.rays)
    counts, segs = sa.trace_paths(ctx, scene, g, rb, depth=4, seed=w.bounce_seed, keep=True,
                                  order=sa.bounce_order("exit", box=box))
    ctx.sync()
    og = oracle.build_twolevel(w.scene, **w.oracle)
    rays_o = w.rays
    for d, (rays_d, p, t, cnt) in enumerate(segs):
        n = rays_o.shape[1]
        if d:
            assert int(counts[d - 1].download(np.uint32, 1)[0]) == n
        np.testing.assert_array_equal(rays_d.download(n).view(np.uint32), rays_o.view(np.uint32), err_msg=f"seg {d}")
        po, to = og.trace(rays_o)
        np.testing.assert_array_equal(p.download(np.uint32, n), po, err_msg=f"seg {d}")
        np.testing.assert_array_equal(t.download(np.float32, n).view(np.uint32), to.view(np.uint32))
        if d + 1 < len(segs):
            rays_o, _ = oracle.bounce(w.scene, rays_o, po, to, w.bounce_seed + d, order="exit", box=box)
    g.close()
    scene.close()
