"""Pins for the oracle's traversal + Moller-Trumbore closest hit
(src/Control/Algorithm/SimpleGridTraverser.h:274-401,
src/Control/Algorithm/TwoLevelGridTraverser.h:445-671,
src/Intersector/SimpleIntersector.hpp:50-89): FP64 brute force over every
triangle is the referee, FP32 brute force with the same formula must find the
same hit wherever the grid visits the hit triangle, and the reference's
IntersectorTest scene pins tie-breaking."""
import numpy as np
import pytest

import shinobi_amd as sa

MISS = 0xFFFFFFFF


@pytest.fixture(scope="module")
def scene():
    return sa.gen_soup(31, 3000, 0.04)


@pytest.fixture(scope="module")
def rays():
    return sa.gen_random_rays(32, 6000, 0.0, 1.0)


@pytest.mark.parametrize("kind", [0, 1])
def test_fp64_referee(oracle, scene, rays, kind):
    g = (oracle.build_twolevel(scene, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, top_divisor=1))
    prim, t = g.trace(rays)
    p64, t64 = oracle.brute_f64(scene, rays)
    # the hit-id bar of the grid path (99.99 %): on these 6000 rays every
    # id must match the referee
    agree = np.mean(prim == p64)
    assert agree >= 0.9999, agree
    hit = (prim == p64) & (p64 != MISS)
    assert np.max(np.abs(t[hit] - t64[hit]) / t64[hit]) < 1e-4


@pytest.mark.parametrize("kind", [0, 1])
def test_fp32_brute_force_same_formula(oracle, scene, rays, kind):
    """same MT arithmetic: where the grid and brute force agree on the
    triangle, t is bit-identical; disagreements are rare (cells the plane
    test drops, exact ties)"""
    g = (oracle.build_twolevel(scene, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, top_divisor=1))
    prim, t = g.trace(rays)
    pb, tb = oracle.brute_f32(scene, rays)
    same = prim == pb
    assert same.mean() >= 0.995
    assert np.array_equal(t[same].view(np.uint32), tb[same].view(np.uint32))
    # never a closer hit than brute force (the grid tests a subset)
    hit = prim != MISS
    assert (t[hit] >= tb[hit]).all()


def test_tie_scene(oracle):
    """IntersectorTest (src/Test/IntersectorTest.h:179-258): triangle 0 in the
    plane z=-2, triangles 1..399 identical in the plane y=-4; camera at the
    origin looking along -y.  Strict '<' keeps the first of the equal-t
    duplicates, so the answer is 0, 1 or a miss -- exactly the FP64 referee."""
    n = 400
    soa = np.zeros((9, n), np.float32)
    soa[:, 0] = [-3, -10, -2, 3, -10, -2, 0, 10, -2]
    for i in range(1, n):
        soa[:, i] = [-6, -4, -6, 6, -4, -6, 0, -4, 6]
    r = sa.gen_camera_rays((0, 0, 0), (0, -1, 0), (0, 0, 1), 66.0, 64, 64)
    p64, _ = oracle.brute_f64(soa, r)
    for kind in (0, 1):
        g = oracle.build_twolevel(soa) if kind else oracle.build_uniform(soa, top_divisor=1)
        prim, _ = g.trace(r)
        assert set(np.unique(prim)) == {0, 1, MISS}
        assert np.mean(prim == p64) >= 0.999


def test_axis_parallel_inside_and_short_rays(oracle, scene):
    rng = np.random.default_rng(4)
    n = 4000
    r = np.zeros((7, n), np.float32)
    r[0:3] = rng.uniform(-0.1, 1.1, (3, n))
    ax = rng.integers(0, 3, n)
    r[3 + ax, np.arange(n)] = rng.choice([-1.0, 1.0], n)
    r[6] = np.float32(3.402823466e38)
    r[6, :400] = 0.05
    r[6, 400:420] = 0.0
    pb, tb = oracle.brute_f32(scene, r)
    for kind in (0, 1):
        g = oracle.build_twolevel(scene, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, top_divisor=1)
        prim, t = g.trace(r)
        assert (prim[400:420] == MISS).all()
        assert (t[:400][prim[:400] != MISS] < 0.05).all()
        assert np.mean(prim == pb) >= 0.995


def test_non_finite_rays_terminate(oracle, scene):
    """NaN / inf components end as misses instead of looping (step cap)"""
    r = np.zeros((7, 6), np.float32)
    r[0:3] = 0.5
    r[3:6, 0] = np.nan
    r[3:6, 1] = [np.inf, 0, 0]
    r[0:3, 2] = [1e30, 0.5, 0.5]
    r[3:6, 2] = [-1, 0, 0]
    r[3:6, 3] = [0, 0, 0]
    r[3:6, 4] = [1, 1e-30, 0]
    r[3:6, 5] = [0.3, 0.4, 0.5]
    r[6] = np.float32(3.402823466e38)
    for kind in (0, 1):
        g = oracle.build_twolevel(scene) if kind else oracle.build_uniform(scene, top_divisor=1)
        prim, t = g.trace(r)      # returns
        assert prim.shape == (6,)


def _dir_class(d):
    """4 bins per component of the direction, class = bx | by<<2 | bz<<4:
    bin floor(2 (c + 1)) clamped to 0..3, in binary32 as the generator
    computes it (-1e-9 + 1 rounds to 1, so tiny negatives share bin 2)"""
    f = (d.astype(np.float32) + np.float32(1)) * np.float32(2)
    bins = np.digitize(f, np.array([1, 2, 3], np.float32))
    return bins[0] | bins[1] << 2 | bins[2] << 4


def _bounce_order_ok(cls, src, round_=256):
    """the output is sorted by (rank // round_, class, src), rank the ray's
    position within its class in input (src) order"""
    rank = np.empty(len(src), np.int64)
    for c in np.unique(cls):
        idx = np.flatnonzero(cls == c)
        rank[idx[np.argsort(src[idx])]] = np.arange(len(idx))
    key = np.stack([rank // round_, cls.astype(np.int64), src.astype(np.int64)])
    ordered = np.lexsort(key[::-1])
    return np.array_equal(ordered, np.arange(len(src)))


def test_bounce_rays(oracle, scene):
    """one cosine-weighted bounce per hit: every hit exactly once, ordered by
    direction class in rounds of 256 rays (primary order within a class),
    unit directions in the hemisphere facing away from the incoming ray,
    origin 0.001 off the surface"""
    r = sa.gen_random_rays(40, 5000, 0.0, 1.0)
    g = oracle.build_uniform(scene, top_divisor=1)
    prim, t = g.trace(r)
    b, src = oracle.bounce(scene, r, prim, t, 7)
    cls = _dir_class(b[3:6])
    assert _bounce_order_ok(cls, src)
    assert len(np.unique(cls)) > 16
    order = np.argsort(src)
    b, src = b[:, order], src[order]
    hit = np.flatnonzero(prim != MISS)
    assert np.array_equal(src, hit)
    d = b[3:6].astype(np.float64)
    assert np.allclose(np.linalg.norm(d, axis=0), 1.0, atol=1e-5)
    v = scene.astype(np.float64).reshape(3, 3, -1)[:, :, prim[hit]]
    n = np.cross((v[1] - v[0]).T, (v[2] - v[0]).T).T
    n /= np.linalg.norm(n, axis=0)
    din = r[3:6, hit].astype(np.float64)
    n *= -np.sign(np.sum(n * din, axis=0))
    assert (np.sum(n * d, axis=0) >= -1e-5).all()
    hitp = r[0:3, hit].astype(np.float64) + din * t[hit]
    off = np.sum((b[0:3].astype(np.float64) - hitp) * n, axis=0)
    assert np.allclose(off, 0.001, atol=2e-5)
    # cosine weighting: E[cos theta] = 2/3
    assert abs(np.mean(np.sum(n * d, axis=0)) - 2.0 / 3.0) < 0.03


def test_bounce_order_in_rounds(oracle, scene):
    """classes with more than 256 bounce rays: each class's rays come in
    rounds of 256 (sorted by round, class, input order), so a large class is
    not contiguous and the classes alternate"""
    r = sa.gen_random_rays(41, 60_000, 0.0, 1.0)
    g = oracle.build_uniform(scene, top_divisor=1)
    prim, t = g.trace(r)
    b, src = oracle.bounce(scene, r, prim, t, 9)
    cls = _dir_class(b[3:6])
    assert _bounce_order_ok(cls, src)
    cnt = np.bincount(cls, minlength=64)
    big = int(np.argmax(cnt))
    assert cnt[big] > 512
    at = np.flatnonzero(cls == big)
    assert at[256] - at[255] > 1 and at[255] - at[0] == 255
    # the first round holds min(count, 256) rays of every class, in class order
    n0 = int(np.minimum(cnt, 256).sum())
    assert (np.diff(cls[:n0]) >= 0).all()
    assert np.array_equal(np.bincount(cls[:n0], minlength=64), np.minimum(cnt, 256))
    # a plausible mistake -- plain grouping by class -- fails the check
    grouped = np.lexsort((src, cls))
    assert not _bounce_order_ok(cls[grouped], src[grouped])


def _shadow_rays(oracle, g, scene, rays, light):
    """segments from the closest hits of `rays` towards a point light:
    origin = hit point, direction = light - origin (unnormalised), tmax = 1"""
    prim, t = g.trace(rays)
    hit = prim != MISS
    o = rays[:3, hit] + rays[3:6, hit] * (t[hit] - 1e-4)
    d = np.asarray(light, np.float32)[:, None] - o
    out = np.empty((7, o.shape[1]), np.float32)
    out[:3], out[3:6], out[6] = o, d, 1.0
    return out


@pytest.mark.parametrize("kind", [0, 1])
def test_shadow_any_hit(oracle, scene, rays, kind):
    """any-hit traversal (TwoLevelGridShadowTraverser,
    src/Control/Algorithm/TwoLevelGridTraverser.h:677-828;
    UniformGridShadowTraverser, src/Control/Algorithm/SimpleGridTraverser.h:409-464):
    it visits every cell the closest-hit walk visits, so a segment is
    occluded (t < 0.9999) exactly when the closest hit along it is closer
    than 0.9999; a reported occluder is a real intersection that is never
    closer than the closest one."""
    g = (oracle.build_twolevel(scene, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, top_divisor=1))
    sh = _shadow_rays(oracle, g, scene, rays, (0.3, 1.7, 0.9))
    assert sh.shape[1] > 1000
    pa, ta = g.trace(sh, shadow=True)
    pc, tc = g.trace(sh)
    occ_any = ta < 0.9999
    occ_closest = tc < 0.9999
    assert np.array_equal(occ_any, occ_closest)
    assert 0.05 < occ_any.mean() < 0.95
    hit = pa != MISS
    assert (ta[hit] >= tc[hit]).all()
    # the reported triangle really is hit at the reported t (FP64 referee on that triangle alone)
    idx = np.flatnonzero(hit)[:400]
    for i in idx:
        tri = scene[:, pa[i]:pa[i] + 1]
        p64, t64 = oracle.brute_f64(np.ascontiguousarray(tri), np.ascontiguousarray(sh[:, i:i + 1]))
        assert p64[0] == 0 and abs(t64[0] - ta[i]) <= 1e-4 * max(1e-3, abs(t64[0]))
    # rays that start occluded-short (t < 0.9999) are not traversed at all
    short = sh.copy()
    short[6] = 0.5
    ps, ts = g.trace(short, shadow=True)
    assert (ps == MISS).all() and (ts == 0.5).all()


@pytest.mark.parametrize("kind", [0, 1])
def test_trace_store_geometry(oracle, scene, rays, kind):
    """SimpleRayBuffer::store (src/Engine/RayBuffers.h:52-77) pinned by
    geometry in float64, not by its own formula: the stored point plus
    EPS * dir lies on the plane of the hit triangle, the point sits EPS * dir
    before the hit (a dropped or flipped EPS moves it by 1e-5 |dir|), the
    stored direction is the ray's to within 1 ulp, t equals the trace's, and
    bestHit maps to the hit triangle through indices (misses: 0xFFFFFFFF)."""
    g = (oracle.build_twolevel(scene, lambda_leaf=2.0) if kind else oracle.build_uniform(scene, top_divisor=1))
    prim, t = g.trace(rays)
    buf = g.trace_store(rays)
    hit = prim != MISS
    assert hit.sum() > 100 and (~hit).sum() > 100
    np.testing.assert_array_equal(buf[6].view(np.uint32), t.view(np.uint32))
    best = buf[7].view(np.uint32)
    assert (best[~hit] == MISS).all()
    np.testing.assert_array_equal(g.indices[best[hit]], prim[hit])
    d = rays[3:6].astype(np.float64)
    D = buf[3:6].astype(np.float64)
    assert np.all(np.abs(D - d) <= 1.2e-7 * np.abs(d) + 1e-30)
    o = rays[0:3, hit].astype(np.float64)
    th = t[hit].astype(np.float64)
    P = buf[0:3, hit].astype(np.float64)
    onhit = o + th * D[:, hit]
    np.testing.assert_allclose(P, onhit - 1e-5 * D[:, hit], rtol=0, atol=2e-6)
    v = scene.astype(np.float64).reshape(3, 3, -1)[:, :, prim[hit]]
    nrm = np.cross(v[1] - v[0], v[2] - v[0], axis=0)
    nrm /= np.linalg.norm(nrm, axis=0)
    plane = np.abs(np.sum((P + 1e-5 * D[:, hit] - v[0]) * nrm, axis=0))
    assert np.max(plane) < 2e-6


@pytest.mark.parametrize("order", ["exit", "axis"])
def test_bounce_ordered(oracle, scene, order):
    """the ordered bounce rays are the class-round bounce rays, permuted: the
    same (source, ray) pairs; the float64 exit distance from the box
    (longest first) or |d . axis| (smallest first) of consecutive rays runs
    monotone up to one 16-bit step (1/65535 of the box diagonal, or of 1),
    and rays in one step keep input order"""
    r = sa.gen_random_rays(6, 3000, 0.0, 1.0)
    g = oracle.build_twolevel(scene, lambda_leaf=2.0)
    prim, t = g.trace(r)
    b0, s0 = oracle.bounce(scene, r, prim, t, 7)
    box = (tuple(g.bmin), tuple(g.bmax))
    axis = (0.0, 0.6, 0.8)
    b1, s1 = oracle.bounce(scene, r, prim, t, 7, order=order, box=box, axis=axis)
    assert b1.shape == b0.shape and b1.shape[1] > 500
    p0, p1 = np.argsort(s0), np.argsort(s1)
    np.testing.assert_array_equal(s0[p0], s1[p1])
    np.testing.assert_array_equal(b0[:, p0].view(np.uint32), b1[:, p1].view(np.uint32))
    o, d = b1[0:3].astype(np.float64), b1[3:6].astype(np.float64)
    if order == "exit":
        lo, hi = np.array(box[0])[:, None], np.array(box[1])[:, None]
        with np.errstate(divide="ignore", invalid="ignore"):
            far = np.fmax((lo - o) / d, (hi - o) / d)
        step = float(np.linalg.norm(np.array(box[1]) - np.array(box[0]))) / 65535.0
        key = -np.maximum(np.min(far, axis=0), 0.0) / step    # ascending = longest first, in steps
    else:
        key = np.abs(np.sum(d * np.array(axis)[:, None], axis=0)) * 65535.0
    dk = np.diff(key)
    assert (dk >= -1.0).all(), dk.min()                       # monotone to within one step
    assert (dk > 2.0).sum() > 100                              # and really ordered
    bucket = np.floor(key + 1e-9)
    same = np.abs(np.diff(key)) < 1e-3                         # clearly one step: input order
    assert (s1[1:][same] > s1[:-1][same]).all()
