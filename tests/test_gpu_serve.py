"""Multi-batch launches, graph replay, the autotuner and the serving loop on
the GPU.  Every result is checked against the single-batch trace
(sg_trace_closest / sg_trace_shadow), which the trace tests pin to the
oracle bit for bit; the autotuned grid is checked against the oracle
directly."""
import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _grid(ctx, scene9, kind, **kw):
    soup = sa.FaceSoup(ctx, scene9)
    B = sa.FastTwoLevelGridBuilder if kind == 1 else sa.FastSimpleGridBuilder
    b = B(ctx, **kw)
    b.init(soup)
    g = b.build()
    assert g.info().status == 0
    return soup, g


@pytest.fixture(scope="module")
def scenes(ctx):
    out = []
    out.append(_grid(ctx, sa.gen_soup(41, 20_000, 0.01), 0, lam=5.0))
    out.append(_grid(ctx, sa.gen_heightfield(42, 192, 0.2), 1, lam=5.0, lambda_leaf=2.0))
    out.append(_grid(ctx, sa.gen_soup(43, 3_000, 0.05), 1, lam=3.0, lambda_leaf=1.2))
    return out


def _single(ctx, g, rays7, shadow=False):
    return g.trace_host(rays7, shadow=shadow)


@pytest.mark.parametrize("shadow", [False, True])
def test_batch_equals_single_launches(ctx, scenes, shadow):
    sizes = [1, 63, 64, 65, 4_000, 0, 50_001, 777]
    batch = sa.Batch(ctx, len(sizes))
    items = []
    for k, n in enumerate(sizes):
        _, g = scenes[k % len(scenes)]
        r7 = sa.gen_random_rays(100 + k, max(n, 1), 0.0, 1.0)[:, :n].copy()
        if shadow:
            r7[6] = 1.0
        rb = sa.RayBatch(ctx, r7) if n else sa.RayBatch(ctx, capacity=1)
        rb.n = n
        prim, t = sa.DeviceBuffer(ctx, 4 * max(n, 1)), sa.DeviceBuffer(ctx, 4 * max(n, 1))
        batch.add(g, rb, prim, t)
        items.append((g, r7, prim, t, n))
    batch.commit()
    batch.launch(shadow=shadow)
    ctx.sync()
    for g, r7, prim, t, n in items:
        if n == 0:
            continue
        p1, t1 = _single(ctx, g, r7, shadow)
        np.testing.assert_array_equal(prim.download(np.uint32, n), p1)
        np.testing.assert_array_equal(t.download(np.float32, n).view(np.uint32), t1.view(np.uint32))


def test_graph_replay_rebuild_and_batch(ctx, scenes):
    """a captured rebuild + multi-batch trace replays to the eager result on
    every replay, with eager work and synchronisation between replays"""
    soup, g = scenes[1]
    r7 = sa.gen_random_rays(7, 30_000, 0.0, 1.0)
    rb = sa.RayBatch(ctx, r7)
    prim, t = sa.DeviceBuffer(ctx, 4 * rb.n), sa.DeviceBuffer(ctx, 4 * rb.n)
    batch = sa.Batch(ctx, 2)
    batch.add(g, rb, prim, t)
    _, g2 = scenes[0]
    prim2, t2 = sa.DeviceBuffer(ctx, 4 * rb.n), sa.DeviceBuffer(ctx, 4 * rb.n)
    batch.add(g2, rb, prim2, t2)
    batch.commit()
    ctx.sync()
    eager = [_single(ctx, g, r7), _single(ctx, g2, r7)]
    with sa.Graph.capture(ctx) as graph:
        g.rebuild()
        batch.launch()
    for rep in range(4):
        # poison the outputs with eager work, then replay
        prim.upload(np.full(rb.n, 7, np.uint32))
        prim2.upload(np.full(rb.n, 7, np.uint32))
        _single(ctx, g2, r7[:, :100].copy())
        graph.launch()
        ctx.sync()
        for (p, tt), (pe, te) in zip(((prim, t), (prim2, t2)), eager):
            np.testing.assert_array_equal(p.download(np.uint32, rb.n), pe, err_msg=f"replay {rep}")
            np.testing.assert_array_equal(tt.download(np.float32, rb.n).view(np.uint32), te.view(np.uint32))
    graph.close()


def test_autotune_choice_is_exact(ctx, oracle):
    scene9 = sa.gen_soup(44, 30_000, 0.004)
    soup = sa.FaceSoup(ctx, scene9)
    sample = sa.RayBatch(ctx, sa.gen_random_rays(45, 20_000, 0.0, 1.0))
    res = sa.autotune(ctx, soup, sample, 1_000_000, kinds=3)
    p = res.best
    assert res.candidates >= 10
    assert p.radix_bits in (4, 6, 8) and p.trace_block in (64, 128, 256)
    assert res.cost_ms >= res.build_ms > 0 and res.trace_ms > 0
    # the chosen grid is bit-exact against the oracle with those parameters
    B = sa.FastTwoLevelGridBuilder if p.kind == 1 else sa.FastSimpleGridBuilder
    kw = dict(lam=p.lambda_, lambda_leaf=p.lambda_leaf, radix_bits=p.radix_bits, trace_block=p.trace_block)
    b = B(ctx, **kw)
    b.init(soup)
    g = b.build()
    if p.kind == 1:
        og = oracle.build_twolevel(scene9, lam=p.lambda_, top_divisor=p.top_divisor, lambda_leaf=p.lambda_leaf)
        np.testing.assert_array_equal(g.download("leaves"), og.leaves)
    else:
        og = oracle.build_uniform(scene9, lam=p.lambda_, top_divisor=1)
        np.testing.assert_array_equal(g.download("top_ranges"), og.top_ranges)
    np.testing.assert_array_equal(g.download("indices"), og.indices)
    r7 = sa.gen_random_rays(46, 5_000, 0.0, 1.0)
    po, to = og.trace(r7)
    pg, tg = g.trace_host(r7)
    np.testing.assert_array_equal(pg, po)


def _requests(scenes, n_req, seed):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_req):
        n = int(np.exp(rng.uniform(np.log(1_000), np.log(60_000))))
        sid = int(rng.integers(0, len(scenes)))
        out.append((sid, sa.gen_random_rays(1000 + seed * 100 + k, n, 0.0, 1.0)))
    return out


@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_server_result_alone_equals_in_mix(ctx, scenes, mode):
    srv = sa.Server(ctx, mode=mode, hbm_budget=64 << 20, max_requests=64, max_waves=3)
    ids = [srv.add_grid(g) for _, g in scenes]
    reqs = _requests(scenes, 40, 1)
    # request 0 alone
    sid0, r0 = reqs[0]
    p_alone, t_alone, lat = srv.wait(srv.submit(ids[sid0], r0))
    assert lat > 0
    # all of them at once, request 0 among them
    rids = [srv.submit(ids[s], r) for s, r in reqs]
    res = [srv.wait(i) for i in rids]
    np.testing.assert_array_equal(res[0][0], p_alone)
    np.testing.assert_array_equal(res[0][1].view(np.uint32), t_alone.view(np.uint32))
    # every request equals the single-batch trace of its grid
    for (s, r), (p, t, _) in zip(reqs, res):
        p1, t1 = _single(ctx, scenes[s][1], r)
        np.testing.assert_array_equal(p, p1)
        np.testing.assert_array_equal(t.view(np.uint32), t1.view(np.uint32))
    st = srv.stats()
    assert st.requests == 41 and st.rays == sum(r.shape[1] for _, r in reqs) + r0.shape[1]
    assert 0 < st.p50_us <= st.p99_us <= st.max_us
    if mode == "batched":
        assert st.waves < st.requests     # requests shared launches
    else:
        assert st.waves == st.requests
    srv.close()


def test_server_budget_admission(ctx, scenes):
    """a request larger than the whole budget is refused; requests that only
    fit one at a time are served one wave after another"""
    srv = sa.Server(ctx, mode="batched", hbm_budget=2 << 20, max_requests=8, max_waves=2)
    sid = srv.add_grid(scenes[0][1])
    with pytest.raises(sa.ShinobiError):
        srv.submit(sid, sa.gen_random_rays(1, 100_000, 0.0, 1.0))
    reqs = [sa.gen_random_rays(50 + k, 40_000, 0.0, 1.0) for k in range(6)]   # 1.4 MB each
    rids = [srv.submit(sid, r) for r in reqs]
    for r, i in zip(reqs, rids):
        p, t, _ = srv.wait(i)
        p1, _ = _single(ctx, scenes[0][1], r)
        np.testing.assert_array_equal(p, p1)
    assert srv.stats().waves == 6
    srv.close()


def test_server_dynamic_double_buffer(ctx):
    m0 = W.sphere_mesh(20_000, frame=0)
    m1 = W.sphere_mesh(20_000, frame=3)
    srv = sa.Server(ctx, mode="batched", hbm_budget=32 << 20, max_requests=32, max_waves=2)
    params = dict(kind=1, lam=5.0, lambda_leaf=2.0)
    did = srv.add_dynamic(m0, **params)
    # reference grids: m0 built, then rebuilt with m1's vertices (fixed resolution)
    soup = sa.FaceSoup(ctx, m0)
    b = sa.FastTwoLevelGridBuilder(ctx, lam=5.0, lambda_leaf=2.0)
    b.init(soup)
    g = b.build()
    rays = sa.gen_random_rays(77, 20_000, 0.2, 0.8)
    before = g.trace_host(rays)
    p, t, _ = srv.wait(srv.submit(did, rays))
    np.testing.assert_array_equal(p, before[0])
    inflight = [srv.submit(did, rays) for _ in range(4)]
    srv.update_dynamic(did, m1)
    after_ids = [srv.submit(did, rays) for _ in range(2)]
    soup.upload_vertices(m1)
    g.rebuild()
    after = g.trace_host(rays)
    assert not np.array_equal(before[0], after[0])
    for i in inflight:           # admitted before the update: old or new mesh, never a mix
        p, t, _ = srv.wait(i)
        assert np.array_equal(p, before[0]) or np.array_equal(p, after[0])
    for i in after_ids:
        p, t, _ = srv.wait(i)
        np.testing.assert_array_equal(p, after[0])
        np.testing.assert_array_equal(t.view(np.uint32), after[1].view(np.uint32))
    srv.close()
