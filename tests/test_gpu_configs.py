"""The batched and served paths at their full workload sizes, every result
checked against the oracle (not against another HIP launch):

* batched multi-scene launch: 64 uniform-in-cube soups, N log-uniform in
  [1e3, 5e5] (edge scale 0.01 N^-1/3), each with log-uniform [1e4, 2e6]
  random rays; density, radix width and block size from the autotuner; all
  64 rebuilds and the multi-batch trace captured in one graph and replayed
  twice.  Every item's ids and t equal the oracle's trace of the grid built
  with the autotuned parameters on a sample of its rays (its first 4,096 and
  4,096 more at random positions).
* served queries: 16 resident two-level scenes (50k-2M triangles, soups and
  height fields) plus a 200k-triangle sphere displaced every 33 ms; Poisson
  arrivals at 2,000 requests/s of log-uniform [1e3, 1e6] random rays against
  a random scene, until at least 400 requests are in and the mesh has been
  replaced at least 4 times.  Every result equals the oracle's trace of the
  mesh the request ran on (sg_server_wait_version names it) on all its rays
  up to 40,000 (a sample of 40,000 for larger requests), and a request run
  alone afterwards equals its result in the mix."""
import threading
import time

import numpy as np
import pytest

import shinobi_amd as sa
from shinobi_amd import workloads as W

pytestmark = pytest.mark.gpu


def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)


def _sample(n, rng, k=4096):
    if n <= 2 * k:
        return np.arange(n)
    return np.unique(np.concatenate([np.arange(k), rng.integers(k, n, k)]))


def test_batched_64_scenes_graph_matches_oracle(ctx, oracle):
    rng = np.random.default_rng(4)
    n_scenes = 64
    ns = _log_uniform(rng, 1_000, 500_000, n_scenes)
    nr = _log_uniform(rng, 10_000, 2_000_000, n_scenes)
    items = []
    for k in range(n_scenes):
        s9 = sa.gen_soup(4000 + k, int(ns[k]), float(0.01 * ns[k] ** (-1.0 / 3.0)))
        soup = sa.FaceSoup(ctx, s9)
        r7 = sa.gen_random_rays(6000 + k, int(nr[k]), 0.0, 1.0)
        rays = sa.RayBatch(ctx, r7)
        sample = sa.RayBatch(ctx, sa.gen_random_rays(8000 + k, min(20_000, int(nr[k])), 0.0, 1.0))
        p = sa.autotune(ctx, soup, sample, int(nr[k])).best
        B = sa.FastTwoLevelGridBuilder if p.kind == 1 else sa.FastSimpleGridBuilder
        b = B(ctx, lam=p.lambda_, lambda_leaf=p.lambda_leaf, radix_bits=p.radix_bits, trace_block=p.trace_block)
        b.init(soup)
        g = b.build()
        prim, t = sa.DeviceBuffer(ctx, 4 * rays.n), sa.DeviceBuffer(ctx, 4 * rays.n)
        items.append((s9, r7, p, soup, g, rays, prim, t))
    batch = sa.Batch(ctx, n_scenes)
    for it in items:
        batch.add(it[4], it[5], it[6], it[7])
    batch.commit()
    ctx.sync()
    # the rebuilds on 16 concurrent branches (sg_rebuild_many), as the
    # batched measurement runs them
    with sa.Graph.capture(ctx) as graph:
        batch.rebuild(branches=16)
        batch.launch()
    kinds = set()
    expect = []
    for s9, r7, p, *_ in items:
        kinds.add((p.kind, p.radix_bits))
        if p.kind == 1:
            og = oracle.build_twolevel(s9, lam=p.lambda_, top_divisor=p.top_divisor, lambda_leaf=p.lambda_leaf,
                                       leaf_min_refs=p.leaf_min_refs, nthreads=16)
        else:
            og = oracle.build_uniform(s9, lam=p.lambda_, top_divisor=1, nthreads=16)
        idx = _sample(r7.shape[1], rng)
        po, to = og.trace(r7[:, idx], nthreads=16)
        expect.append((idx, po, to))
        og.close()
    for rep in range(2):
        for it in items:
            it[6].upload(np.full(it[5].n, 7, np.uint32))
        graph.launch()
        ctx.sync()
        assert batch.status(n_scenes) == [0] * n_scenes
        for k, ((s9, r7, p, soup, g, rays, prim, t), (idx, po, to)) in enumerate(zip(items, expect)):
            pg = prim.download(np.uint32, rays.n)[idx]
            tg = t.download(np.float32, rays.n)[idx]
            np.testing.assert_array_equal(pg, po, err_msg=f"replay {rep} item {k} ({ns[k]} tris)")
            np.testing.assert_array_equal(tg.view(np.uint32), to.view(np.uint32), err_msg=f"replay {rep} item {k}")
    assert sum(len(e[0]) for e in expect) > 300_000
    graph.close()


def test_served_mix_with_dynamic_mesh_matches_oracle(ctx, oracle):
    rng = np.random.default_rng(5)
    sizes = _log_uniform(rng, 50_000, 2_000_000, 16)
    statics = []
    for k, nt in enumerate(sizes):
        if k % 2 == 0:
            s9 = sa.gen_soup(500 + k, int(nt), 0.01 * (1e4 / nt) ** (1.0 / 3.0))
        else:
            gsz = max(8, int(np.sqrt(nt / 2.0)) + 1)
            s9 = sa.gen_heightfield(500 + k, gsz, 0.2)
        soup = sa.FaceSoup(ctx, s9)
        b = sa.FastTwoLevelGridBuilder(ctx, lam=5.0, lambda_leaf=2.0)
        b.init(soup)
        statics.append((s9, soup, b.build()))
    frames = [W.sphere_mesh(200_000, frame=f) for f in range(6)]
    K = 24
    tsz = _log_uniform(rng, 1_000, 1_000_000, K)
    # request rays and results in pinned host memory, as a serving client
    # keeps them (pageable buffers make every copy a staged, blocking one)
    pinned = []          # owners of the pinned buffers, alive until the end
    temps = []
    for k, n in enumerate(tsz):
        r = sa.PinnedArray(ctx, (7, int(n)), np.float32)
        r.a[:] = sa.gen_random_rays(900 + k, int(n), 0.0, 1.0)
        pinned.append(r)
        temps.append(r.a)
    outs = []

    def out_for(n):
        p, t = sa.PinnedArray(ctx, (n,), np.uint32), sa.PinnedArray(ctx, (n,), np.float32)
        outs.append((p, t))
        return p.a, t.a

    srv = sa.Server(ctx, mode="batched", hbm_budget=3 << 30, max_requests=512, max_waves=4, max_wave_rays=1 << 22)
    ids = [srv.add_grid(g) for _, _, g in statics]
    dyn = srv.add_dynamic(frames[0], kind=1, lam=5.0, lambda_leaf=2.0)
    version_frame = {0: 0}
    stop = threading.Event()
    errors = []

    def producer():
        f = 0
        try:
            while not stop.is_set():
                t0 = time.perf_counter()
                f += 1
                srv.update_dynamic(dyn, frames[f % len(frames)])
                version_frame[f] = f % len(frames)
                stop.wait(max(0.0, 0.033 - (time.perf_counter() - t0)))
        except Exception as e:   # noqa: BLE001 -- reported by the test thread
            errors.append(e)

    # arrivals continue until at least 400 requests are in and the dynamic
    # mesh has been replaced at least 4 times
    n_max = 2000
    gaps = rng.exponential(1.0 / 2000.0, n_max)
    which = rng.integers(0, K, n_max)
    scene_of = rng.integers(0, 17, n_max)
    bufs = [out_for(int(tsz[which[k]])) for k in range(400)]
    th = threading.Thread(target=producer, daemon=True)
    th.start()
    rids = []
    t = time.perf_counter()
    k = 0
    while k < n_max and (k < 400 or len(version_frame) < 5):
        t += gaps[k]
        while True:        # sleep / yield, never spin holding the GIL: the mesh thread runs on time
            left = t - time.perf_counter()
            if left <= 0:
                break
            time.sleep(left - 1e-4 if left > 3e-4 else 0)
        sid = ids[scene_of[k]] if scene_of[k] < 16 else dyn
        p_out, t_out = bufs[k] if k < len(bufs) else out_for(int(tsz[which[k]]))
        rids.append(srv.submit(sid, temps[which[k]], p_out, t_out))
        k += 1
    n_req = k
    results = [srv.wait(r, with_version=True) for r in rids]
    stop.set()
    th.join()
    assert not errors, errors
    assert srv.stats().regrown == 0
    versions = {v for (_, _, _, v), s in zip(results, scene_of[:n_req]) if s == 16}
    assert len(version_frame) >= 4 and len(versions) >= 2, (len(version_frame), versions)

    # the oracle: static grids as built; the dynamic mesh of version v is
    # the sphere of frame version_frame[v] rebuilt at frame 0's resolution
    cache = {}

    def oracle_grid(s):
        key = ("g", s)
        if key not in cache:
            if s[0] < 16:
                cache[key] = oracle.build_twolevel(statics[s[0]][0], lam=5.0, top_divisor=6, lambda_leaf=2.0,
                                                   nthreads=16)
            else:
                res0 = oracle_grid((16, 0)).res if s[1] else None
                kw = dict(res_override=tuple(res0)) if res0 is not None else {}
                cache[key] = oracle.build_twolevel(frames[version_frame[s[1]]], lam=5.0, top_divisor=6,
                                                   lambda_leaf=2.0, nthreads=16, **kw)
        return cache[key]

    # every request, on all its rays up to 40,000 (larger ones: the first
    # 20,000 and 20,000 at random positions)
    checked = 0
    for k, (p, tt, lat, v) in enumerate(results):
        s = (int(scene_of[k]), v if scene_of[k] == 16 else 0)
        r7 = temps[which[k]]
        idx = _sample(r7.shape[1], rng, 20_000)
        po, to = oracle_grid(s).trace(r7[:, idx], nthreads=16)
        np.testing.assert_array_equal(p[idx], po, err_msg=f"request {k} scene {s}")
        np.testing.assert_array_equal(tt[idx].view(np.uint32), to.view(np.uint32), err_msg=f"request {k} scene {s}")
        checked += idx.size
    assert checked > 5_000_000
    # alone equals in the mix (static scenes: the same grid afterwards)
    for k in [i for i in range(n_req) if scene_of[i] < 16][:6]:
        p, tt, _ = srv.wait(srv.submit(ids[scene_of[k]], temps[which[k]]))
        np.testing.assert_array_equal(p, results[k][0])
        np.testing.assert_array_equal(tt.view(np.uint32), results[k][1].view(np.uint32))
    st = srv.stats()
    assert st.waves < st.requests
    srv.close()
    for g in cache.values():
        g.close()
