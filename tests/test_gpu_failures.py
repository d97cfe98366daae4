"""Failure handling and concurrency of the C ABI on the GPU: a rebuild that
outgrows its grid's buffers is reported (single trace, multi-batch launch,
served request) instead of returning silent misses, a dynamic served mesh
that outgrows its grid is built again before it is swapped in, scenes can
be added to a server while its requests run, and the scratch of the
standalone primitives is per stream.  The overflow is forced from the host
(a mesh whose triangles grow 20x between build and rebuild); every kernel
clamps its writes to the buffers it was given, so it still exits normally."""
import ctypes as C

import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import _lib as L
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu

CAPACITY = 5   # SG_ERR_CAPACITY


def _grid(ctx, soa9, kind, **kw):
    soup = sa.FaceSoup(ctx, soa9)
    B = sa.FastTwoLevelGridBuilder if kind == 1 else sa.FastSimpleGridBuilder
    b = B(ctx, **kw)
    b.init(soup)
    g = b.build()
    assert g.info().status == 0
    return soup, g


@pytest.mark.parametrize("kind", [0, 1])
def test_rebuild_overflow_is_reported(ctx, kind):
    small = sa.gen_soup(61, 5000, 0.01)
    big = sa.gen_soup(61, 5000, 0.2)
    soup, g = _grid(ctx, small, kind, lam=5.0, lambda_leaf=2.0)
    rays = sa.gen_random_rays(62, 4000, 0.0, 1.0)
    print(f"overflow kind={kind}: rebuild of a 20x larger mesh", flush=True)
    soup.upload_vertices(big)
    g.rebuild()
    assert g.info().status == CAPACITY
    p, t = g.trace_host(rays)
    assert (p == sa.MISS).all()
    # a multi-batch launch reports the item's grid status
    rb = sa.RayBatch(ctx, rays)
    prim, tb = sa.DeviceBuffer(ctx, 4 * rb.n), sa.DeviceBuffer(ctx, 4 * rb.n)
    _, ok_grid = _grid(ctx, small, kind, lam=5.0, lambda_leaf=2.0)
    prim2, tb2 = sa.DeviceBuffer(ctx, 4 * rb.n), sa.DeviceBuffer(ctx, 4 * rb.n)
    batch = sa.Batch(ctx, 2)
    batch.add(ok_grid, rb, prim2, tb2)
    batch.add(g, rb, prim, tb)
    batch.commit()
    batch.launch()
    assert batch.status(2) == [0, CAPACITY]
    # a served request on the overflowed grid fails with its status; one on
    # a good grid in the same wave does not
    for mode in ("batched", "plain"):
        srv = sa.Server(ctx, mode=mode, hbm_budget=16 << 20, max_requests=8, max_waves=2)
        bad, good = srv.add_grid(g), srv.add_grid(ok_grid)
        r_bad, r_good = srv.submit(bad, rays), srv.submit(good, rays)
        with pytest.raises(sa.ShinobiError) as e:
            srv.wait(r_bad)
        assert e.value.status == CAPACITY
        pg, tg, _ = srv.wait(r_good)
        p1, t1 = ok_grid.trace_host(rays)
        np.testing.assert_array_equal(pg, p1)
        srv.close()
    # a new build sizes the buffers for the new mesh
    b = (sa.FastTwoLevelGridBuilder if kind else sa.FastSimpleGridBuilder)(ctx, lam=5.0, lambda_leaf=2.0)
    b.init(soup)
    assert b.build().info().status == 0


def test_dynamic_regrow_after_overflow(ctx, oracle):
    """update_dynamic with a mesh that outgrows the grid's capacity slack:
    the back grid is built again and the served results equal a fresh build
    of the new mesh (and the oracle)"""
    small = sa.gen_soup(71, 6000, 0.01)
    big = sa.gen_soup(71, 6000, 0.2)
    srv = sa.Server(ctx, mode="batched", hbm_budget=32 << 20, max_requests=16, max_waves=2)
    did = srv.add_dynamic(small, kind=1, lam=5.0, lambda_leaf=2.0)
    rays = sa.gen_random_rays(72, 6000, 0.0, 1.0)
    print("dynamic regrow: update with a 20x larger mesh", flush=True)
    srv.update_dynamic(did, big)
    assert srv.stats().regrown == 1
    p, t, _ = srv.wait(srv.submit(did, rays))
    _, g = _grid(ctx, big, 1, lam=5.0, lambda_leaf=2.0)
    p1, t1 = g.trace_host(rays)
    np.testing.assert_array_equal(p, p1)
    np.testing.assert_array_equal(t.view(np.uint32), t1.view(np.uint32))
    og = oracle.build_twolevel(big, lam=5.0, top_divisor=6, lambda_leaf=2.0)
    po, to = og.trace(rays)
    np.testing.assert_array_equal(p, po)
    np.testing.assert_array_equal(t.view(np.uint32), to.view(np.uint32))
    # back to the small mesh: fits the regrown buffers, no further regrow
    srv.update_dynamic(did, small)
    assert srv.stats().regrown == 1
    srv.close()


def test_scenes_added_while_requests_run(ctx):
    """sg_server_add_grid while submitted requests are in flight: the
    server's scene table grows under the running waves"""
    _, g0 = _grid(ctx, sa.gen_soup(81, 20_000, 0.01), 0, lam=5.0)
    _, g1 = _grid(ctx, sa.gen_heightfield(82, 128, 0.2), 1, lam=5.0, lambda_leaf=2.0)
    srv = sa.Server(ctx, mode="batched", hbm_budget=64 << 20, max_requests=64, max_waves=3)
    ids = [srv.add_grid(g0), srv.add_grid(g1)]
    reqs = [(k % 2, sa.gen_random_rays(900 + k, 5_000 + 3_000 * k, 0.0, 1.0)) for k in range(24)]
    rids = []
    for k, (s, r) in enumerate(reqs):
        rids.append(srv.submit(ids[s], r))
        for _ in range(8):                      # 192 scenes added in all
            ids.append(srv.add_grid(g1 if k % 2 else g0))
    late = srv.submit(ids[-1], reqs[0][1])      # a request on a scene added last
    for (s, r), i in zip(reqs, rids):
        p, t, _ = srv.wait(i)
        p1, t1 = (g1 if s else g0).trace_host(r)
        np.testing.assert_array_equal(p, p1)
        np.testing.assert_array_equal(t.view(np.uint32), t1.view(np.uint32))
    p, _, _ = srv.wait(late)
    np.testing.assert_array_equal(p, (g1 if 23 % 2 else g0).trace_host(reqs[0][1])[0])
    srv.close()


def test_scratch_is_per_stream(ctx):
    """two bounce calls of different sizes enqueued on two streams of one
    context at the same time give the results each gives alone"""
    w = W.heightfield2m(grid=96, res_x=160, res_y=120)
    soup = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(soup)
    g = b.build()
    full = sa.RayBatch(ctx, w.rays)
    part = sa.RayBatch(ctx, w.rays[:, : w.rays.shape[1] // 3].copy())
    outs = {}
    for name, rb in (("full", full), ("part", part)):
        prim, t = g.trace(rb)
        outs[name] = (rb, prim, t)
    ctx.sync()

    def run(name, seed, stream):
        rb, prim, t = outs[name]
        out = sa.RayBatch(ctx, capacity=rb.n)
        cnt = sa.bounce(ctx, soup, rb, prim, t, seed, out, stream=stream)
        return out, cnt

    alone = {}
    for name, seed in (("full", 5), ("part", 6)):
        o, c = run(name, seed, None)
        ctx.sync()
        m = int(c.download(np.uint32, 1)[0])
        alone[name] = o.download(m)
    streams = []
    for _ in range(2):
        s = C.c_void_p()
        sa.check(ctx._l.sg_stream_create(ctx.h, C.byref(s)), "sg_stream_create")
        streams.append(s)
    try:
        for rep in range(3):
            res = [run("full", 5, streams[0]), run("part", 6, streams[1])]
            for s in streams:
                sa.check(ctx._l.sg_stream_sync(ctx.h, s), "sg_stream_sync")
            for (o, c), name in zip(res, ("full", "part")):
                m = int(c.download(np.uint32, 1)[0])
                np.testing.assert_array_equal(o.download(m).view(np.uint32), alone[name].view(np.uint32),
                                              err_msg=f"{name} rep {rep}")
    finally:
        for s in streams:
            ctx._l.sg_stream_destroy(ctx.h, s)
