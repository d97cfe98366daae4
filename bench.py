#!/usr/bin/env python3
"""bench.py -- closest-hit Mrays/s + two-level grid build ms on one MI355X.

Workload (one step = one pass of the hot path over one batch):
  scene  2,093,058-triangle height field (1024x1024 vertices over [0,1]^2,
         8-octave value noise, amplitude 0.2, seed 2), resident in HBM
  step   1. two-level grid rebuild from the SoA triangles (bounds,
            resolution-fixed top level lambda 5 / 6, leaf lambda 2, every
            phase on the device, no host sync)
         2. closest-hit trace of 1920x1080 primary rays from (0.5,-0.3,0.6),
            generated inside the trace kernel (sg_trace_camera: the
            reference's Camera + RegularPixelSampler per pixel) and visited
            in 2x32 pixel tiles; hits and rays are written at their pixel
            ids (the row-major visit is reported as trace_primary_rowmajor_ms)
         3. one cosine-weighted diffuse bounce per hit (device generator,
            compacted and ordered longest first: descending exit distance
            from the grid's bounds, sg_bounce_ordered)
         4. closest-hit trace of the bounce rays (count read on the device)
value  = closest-hit Mrays/s = (primary + bounce rays) / (time of the two
         trace kernels), measured with HIP events on the stream the kernels
         run on.  build_ms (the rebuild of each step) and per-phase times (a
         separate loop of rebuilds with events between the phases) are
         reported beside it.

Usage: python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0          # MI355X HBM3E, GB/s
FP32_PEAK_TFLOPS = 157.3       # MI355X vector FP32, TFLOP/s

# The camera trace visits the pixels in 2x32-pixel column tiles: a wave then
# holds near and far rays, and the lanes that finish early test the
# remaining rays' cells with them (wave-shared tests).  Measured on the
# height field: 2x32 0.51 ms, 4x16 0.56, 8x8 0.59, row-major 0.62; on the
# ruins camera 2x32 0.55, 8x8 0.60, row-major 0.70.
PRIMARY_TILE = (2, 32)

PHASES = ["bounds", "top_emit", "top_sort", "top_ranges", "leaf_cells", "leaf_emit", "leaf_sort", "leaf_ranges"]


def hbm_copy_rate(ctx, nbytes: int = 1 << 30, reps: int = 5) -> dict:
    """The part's streaming rate as measured: the HIP runtime's own
    device-to-device copy of 1 GiB on the context's stream (read and write
    both counted), median of `reps` after one warm-up.  Reported beside the
    spec peak the roofline fractions use."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
    hip.hipMemcpyAsync.restype = C.c_int
    import shinobi_amd as sa
    src, dst = sa.DeviceBuffer(ctx, nbytes), sa.DeviceBuffer(ctx, nbytes)
    e0, e1 = ctx.event(), ctx.event()
    ts = []
    for k in range(reps + 1):
        e0.record()
        if hip.hipMemcpyAsync(dst.ptr, src.ptr, nbytes, 3, ctx.stream) != 0:   # 3: device to device
            raise RuntimeError("hipMemcpyAsync failed")
        e1.record()
        ctx.sync()
        if k:
            ts.append(e0.elapsed_ms(e1))
    ms = float(np.median(ts))
    del src, dst
    return {"how": "hipMemcpyAsync device to device, 1 GiB, read + write counted", "ms": round(ms, 4),
            "gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}


def build_bytes(N: int, P1: int, C1: int, P2: int, C2: int) -> dict:
    """Algorithmic HBM bytes of each build phase (36 B per triangle, 8 B per
    pair, 8-bit-digit sort convention); the reference's count / scan / emit
    stages are summed into the fused single-pass kernels that replace them."""
    k1 = max(1, math.ceil(math.log2(C1 + 1)))
    k2 = max(1, math.ceil(math.log2(C2 + 1)))
    return {
        "bounds": 36 * N,
        "top_emit": (36 * N + 4 * N) + 8 * N + (36 * N + 4 * N + 8 * P1),
        "top_sort": 16 * P1 * math.ceil(k1 / 8),
        "top_ranges": 8 * P1 + 8 * C1,
        "leaf_cells": (8 * C1 + 8 * C1 + 4 * C1) + 8 * C1,
        "leaf_emit": (8 + 8 + 36 + 4) * P1 + 8 * P1 + (8 + 8 + 36 + 4) * P1 + 8 * P2,
        "leaf_sort": 16 * P2 * math.ceil(k2 / 8),
        "leaf_ranges": 8 * P2 + 8 * C2 + 4 * P2,
    }


def trace_flops(stats: np.ndarray) -> float:
    """FP32 work of closest-hit traversal, F = 40 + 56 T + 20 S_leaf +
    20 S_top + 48 E per ray (T triangle tests, S DDA steps, E non-empty top
    cells entered; FMA = 2)."""
    s = stats.astype(np.float64)
    return float(np.sum(40.0 + 56.0 * s[:, 0] + 20.0 * s[:, 1] + 20.0 * s[:, 2] + 48.0 * s[:, 3]))


def traffic_of(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE scaled by the copy-kernel
    factor + WRITE_SIZE) from the committed rocprofv3 counter summary of this
    workload, or None when the summary does not list it."""
    path = os.path.join(ROOT, "profiles", "heightfield2m_pmc_traffic.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else int(k["fetch_bytes"] + k["write_bytes"])


def host_cpu() -> dict:
    """the host's CPU model and the cores this process may use: its affinity
    mask, limited by a cgroup CPU quota when one is set"""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    cores = len(os.sched_getaffinity(0))
    quota = None
    try:
        with open("/sys/fs/cgroup/cpu.max") as f:
            q, per = f.read().split()[:2]
            if q != "max":
                quota = max(1, int(int(q) / int(per)))
    except (OSError, ValueError):
        pass
    return {"model": model, "affinity_cores": cores, "cgroup_quota_cores": quota,
            "cores": min(cores, quota) if quota else cores}


def counters_of(launch: str) -> dict:
    """VALU / lane / L2 counters of a traversal launch from the committed
    rocprofv3 SQ / TCC summary of this workload ({} when it does not list it)"""
    path = os.path.join(ROOT, "profiles", "heightfield2m_pmc_traversal.json")
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(launch)
    except (OSError, ValueError, KeyError):
        return {}
    if not k:
        return {}
    return {"valu_util": k.get("valu_util"), "lanes_active": k.get("lanes_active"), "l2_hit": k.get("l2_hit"),
            "waves_per_simd": k.get("waves_per_simd")}


def cpu_baseline(w, n_primary: int, seed: int, threads: int, runs: int = 5, w_small=None) -> dict:
    """The oracle (the restatement of the reference's GPU path, compiled
    -O2 for the host) on this host's cores: two-level build, closest-hit
    trace of the primary rays, one bounce per hit, trace of the bounce rays.
    Median of `runs` runs after one warm-up, with all usable cores and with
    one core.  With `w_small` (the 10,000-triangle uniform soup and its
    256x256 camera rays), also that workload's uniform build and trace."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import py_oracle as O
    O.load()

    def small(nt):
        t0 = time.perf_counter()
        og = O.build_uniform(w_small.scene, nthreads=nt, **w_small.oracle)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        og.trace(w_small.rays, nthreads=nt)
        t_trace = time.perf_counter() - t0
        og.close()
        return t_build, t_trace

    def once(nt):
        t0 = time.perf_counter()
        og = O.build_twolevel(w.scene, nthreads=nt, **w.oracle)
        t_build = time.perf_counter() - t0
        t0 = time.perf_counter()
        prim, t = og.trace(w.rays, nthreads=nt)
        t_p = time.perf_counter() - t0
        b, _ = O.bounce(w.scene, w.rays, prim, t, seed)
        t0 = time.perf_counter()
        og.trace(b, nthreads=nt)
        t_b = time.perf_counter() - t0
        og.close()
        return t_build, t_p + t_b, b.shape[1]

    res = {}
    for nt in (threads, 1):
        once(nt)
        rows = [once(nt) for _ in range(runs)]
        t_build = float(np.median([r[0] for r in rows]))
        t_trace = float(np.median([r[1] for r in rows]))
        rays = n_primary + rows[0][2]
        res[nt] = (round(rays / t_trace / 1e6, 3), round(t_build * 1e3, 1), rows[0][2])
    cpu = host_cpu()
    out = {"value": res[threads][0], "unit": "Mrays/s", "cores": threads, "kind": "port",
           "build_ms": res[threads][1], "one_core": {"value": res[1][0], "unit": "Mrays/s", "build_ms": res[1][1]},
           "cpu_model": cpu["model"], "host": cpu, "runs": f"median of {runs} after 1 warm-up",
           "sample": f"full workload on the host: oracle two-level build + closest-hit trace of {n_primary} "
                     f"primary and {res[threads][2]} bounce rays (row-major pixel order), {threads} threads "
                     f"and 1 thread"}
    if w_small is not None:
        n_small = w_small.rays.shape[1]
        sm = {}
        for nt in (threads, 1):
            small(nt)
            rows = [small(nt) for _ in range(runs)]
            sm[nt] = {"build_ms": round(float(np.median([r[0] for r in rows])) * 1e3, 3),
                      "mrays_per_s": round(n_small / float(np.median([r[1] for r in rows])) / 1e6, 3)}
        out["soup10k"] = {"cores": threads, **sm[threads], "one_core": sm[1],
                          "sample": f"oracle uniform build + closest-hit trace of {n_small} camera rays"}
    return out


def _log_uniform(rng, lo, hi, n):
    return np.exp(rng.uniform(np.log(lo), np.log(hi), n)).astype(np.int64)


def wait_until(t: float) -> None:
    """sleep / yield until perf_counter() reaches t without holding the GIL
    in a busy loop (the dynamic-mesh thread must run on schedule)"""
    while True:
        left = t - time.perf_counter()
        if left <= 0:
            return
        time.sleep(left - 1e-4 if left > 3e-4 else 0)


def serving_bench(sa, W, ctx, seconds: float = 2.0, rate: float = 2000.0, burst: int = 400) -> dict:
    """Served ray queries: 16 resident static scenes (50k-2M triangles,
    uniform soups and height fields) plus a 200k-triangle sphere rebuilt every
    33 ms on the server's own stream; Poisson arrivals at `rate` req/s, each a
    batch of log-uniform [1e3, 1e6] random rays against a random scene, host
    data in pinned memory.  The same arrival schedule runs against the plain
    per-request path and the continuous-batching server; then a burst of
    `burst` requests submitted at once measures each path's throughput at
    saturation."""
    import threading
    rng = np.random.default_rng(5)
    sizes = _log_uniform(rng, 50_000, 2_000_000, 16)
    grids = []
    for k, nt in enumerate(sizes):
        if k % 2 == 0:
            s9 = sa.gen_soup(500 + k, int(nt), 0.01 * (1e4 / nt) ** (1.0 / 3.0))
        else:
            g = max(8, int(np.sqrt(nt / 2.0)) + 1)
            s9 = sa.gen_heightfield(500 + k, g, 0.2)
        soup = sa.FaceSoup(ctx, s9)
        b = sa.FastTwoLevelGridBuilder(ctx, lam=5.0, lambda_leaf=2.0)
        b.init(soup)
        grids.append((soup, b.build()))
    n_tris = [int(g.scene.n) for _, g in grids]
    # request templates in pinned memory (a request is a (7, n) block)
    K = 48
    tsz = _log_uniform(rng, 1_000, 1_000_000, K)
    temps = []
    for k, n in enumerate(tsz):
        r = sa.PinnedArray(ctx, (7, int(n)), np.float32)
        r.a[:] = sa.gen_random_rays(900 + k, int(n), 0.0, 1.0)
        outs = [(sa.PinnedArray(ctx, (int(n),), np.uint32), sa.PinnedArray(ctx, (int(n),), np.float32))
                for _ in range(2)]
        temps.append((r, outs))
    frames = [W.sphere_mesh(200_000, frame=f) for f in range(4)]
    n_req = int(seconds * rate)
    gaps = rng.exponential(1.0 / rate, n_req)
    which = rng.integers(0, K, n_req)
    scene_of = rng.integers(0, 17, n_req)

    def run(mode: str, timed: bool):
        srv = sa.Server(ctx, mode=mode, hbm_budget=3 << 30, max_requests=512, max_waves=4,
                        max_wave_rays=1 << 22)
        ids = [srv.add_grid(g) for _, g in grids]
        ids.append(srv.add_dynamic(frames[0], lam=5.0, lambda_leaf=2.0))
        stop = threading.Event()
        updates = []

        def producer():
            f = 0
            while not stop.is_set():
                t0 = time.perf_counter()
                f += 1
                srv.update_dynamic(ids[16], frames[f % len(frames)])
                updates.append(time.perf_counter())
                stop.wait(max(0.0, 0.033 - (time.perf_counter() - t0)))

        th = threading.Thread(target=producer, daemon=True)
        th.start()
        srv.reset_stats()
        t = time.perf_counter()
        n_sub = n_req if timed else burst
        t_begin = time.perf_counter()
        for k in range(n_sub):
            if timed:
                t += gaps[k]
                wait_until(t)
            r, outs = temps[which[k]]
            o = outs[k & 1]
            srv.submit(ids[scene_of[k]], r.a, o[0].a, o[1].a)
        srv.drain()
        t_end = time.perf_counter()
        st = srv.stats()
        stop.set()
        th.join()
        srv.close()
        during = [u for u in updates if t_begin <= u <= t_end]
        gaps_ms = np.diff(during) * 1e3 if len(during) > 1 else np.array([0.0])
        return {"requests": st.requests, "p50_us": round(st.p50_us, 1), "p99_us": round(st.p99_us, 1),
                "mean_us": round(st.mean_us, 1), "mrays_per_s": round(st.mrays_per_s, 1), "waves": st.waves,
                "mesh_updates": len(during), "update_interval_ms_mean": round(float(np.mean(gaps_ms)), 1),
                "update_interval_ms_max": round(float(np.max(gaps_ms)), 1)}

    out = {"scenes": {"static_tris": n_tris, "dynamic_tris": int(frames[0].shape[1]), "rebuild_ms_period": 33},
           "offered_req_per_s": rate,
           "offered_mrays_per_s": round(rate * float(np.mean(tsz[which])) / 1e6, 1),
           "mean_rays_per_request": round(float(np.mean(tsz[which])), 1)}
    out["plain"] = run("plain", True)
    out["batched"] = run("batched", True)
    sat_p = run("plain", False)
    sat_b = run("batched", False)
    out["saturated"] = {"requests": burst, "plain": sat_p, "batched": sat_b,
                        "throughput_ratio": round(sat_b["mrays_per_s"] / max(sat_p["mrays_per_s"], 1e-9), 2),
                        "p99_ratio": round(sat_b["p99_us"] / max(sat_p["p99_us"], 1e-9), 3)}
    return out


def cluster_bench(sa, W, ctx, spp: int = 4, depth: int = 4, reps: int = 3) -> dict:
    """Dense cluster in a large scene (500k tiny + 2k large triangles):
    two-level rebuild, then 1280x720 x `spp` camera paths of `depth`
    diffuse segments chained on the device (bounces longest first)."""
    w = W.cluster()
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    info = g.info()
    rays = sa.RayBatch(ctx, w.rays)
    order = sa.bounce_order("exit", grid=g)
    e0, e1, e2 = ctx.event(), ctx.event(), ctx.event()
    best = None
    for rep in range(reps + 1):
        e0.record()
        g.rebuild()
        e1.record()
        counts = []
        for smp in range(spp):
            counts += sa.trace_paths(ctx, scene, g, rays, depth, w.bounce_seed + 16 * smp, order=order)
        e2.record()
        ctx.sync()
        n_rays = spp * rays.n + sum(int(c.download(np.uint32, 1)[0]) for c in counts)
        b_ms, t_ms = e0.elapsed_ms(e1), e1.elapsed_ms(e2)
        if rep and (best is None or t_ms < best[1]):
            best = (b_ms, t_ms, n_rays)
    return {"n_tris": int(w.scene.shape[1]), "top_res": list(info.res), "C2": info.C2, "P2": info.P2,
            "build_ms": round(best[0], 3), "paths": f"{w.rays.shape[1]} pixels x {spp} spp x {depth} segments",
            "rays": best[2], "trace_ms": round(best[1], 3), "mrays_per_s": round(best[2] / best[1] / 1e3, 1)}


def ruins_bench(sa, W, ctx, reps: int = 10) -> dict:
    """The reference's own scene (scenes/ruins: 38,665 triangles, its
    1024x1024 camera.cfg view): two-level rebuild with the reference
    builder's defaults, camera trace (rays generated in the kernel), one
    diffuse bounce, bounce trace."""
    golden = os.path.join(ROOT, "tests", "golden")
    w = W.scene_file(os.path.join(golden, "ruins.obj.gz"), os.path.join(golden, "ruins_camera.cfg"))
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    info = g.info()
    n = w.rays.shape[1]
    cam = sa.pinhole(*w.camera)
    rays = sa.RayBatch(ctx, capacity=n)
    prim, t = sa.DeviceBuffer(ctx, 4 * n), sa.DeviceBuffer(ctx, 4 * n)
    out = sa.RayBatch(ctx, capacity=n)
    bp, bt = sa.DeviceBuffer(ctx, 4 * n), sa.DeviceBuffer(ctx, 4 * n)
    count = sa.DeviceBuffer(ctx, 4)
    order = sa.bounce_order("exit", grid=g)
    ev = [ctx.event() for _ in range(5)]
    rows = []
    for rep in range(reps + 2):
        ev[0].record()
        g.rebuild()
        ev[1].record()
        g.trace_camera(cam, PRIMARY_TILE, prim=prim, t=t, rays_out=rays)
        ev[2].record()
        sa.bounce(ctx, scene, rays, prim, t, w.bounce_seed, out, count=count, order=order)
        ev[3].record()
        g.trace(out, prim=bp, t=bt, d_count=count.ptr, n=n)
        ev[4].record()
        ctx.sync()
        if rep >= 2:
            rows.append([ev[i].elapsed_ms(ev[i + 1]) for i in range(4)])
    m = int(count.download(np.uint32, 1)[0])
    r = np.median(np.array(rows), axis=0)
    return {"n_tris": int(w.scene.shape[1]), "primary_rays": n, "bounce_rays": m, "res_top": list(info.res),
            "C1": info.C1, "P1": info.P1, "C2": info.C2, "P2": info.P2, "build_ms": round(float(r[0]), 4),
            "trace_primary_ms": round(float(r[1]), 4), "trace_bounce_ms": round(float(r[3]), 4),
            "mrays_per_s": round((n + m) / (r[1] + r[3]) / 1e3, 1)}


def soup10k_bench(sa, W, ctx, reps: int = 20) -> dict:
    """The reference's CPU-path configuration on the GPU: 10,000 triangles
    (centroids uniform in the unit cube, vertices within 0.01), uniform grid
    (lambda 5), 256x256 pinhole camera rays from (0.5, 0.5, -1) along +z,
    60 degrees; rebuild and camera trace (rays generated in the kernel)."""
    w = W.soup10k()
    scene = sa.FaceSoup(ctx, w.scene)
    b = sa.FastSimpleGridBuilder(ctx, **w.build)
    b.init(scene)
    g = b.build()
    info = g.info()
    n = w.rays.shape[1]
    cam = sa.pinhole(*w.camera)
    prim, t = sa.DeviceBuffer(ctx, 4 * n), sa.DeviceBuffer(ctx, 4 * n)
    ev = [ctx.event() for _ in range(3)]
    rows = []
    for rep in range(reps + 2):
        ev[0].record()
        g.rebuild()
        ev[1].record()
        g.trace_camera(cam, PRIMARY_TILE, prim=prim, t=t)
        ev[2].record()
        ctx.sync()
        if rep >= 2:
            rows.append([ev[0].elapsed_ms(ev[1]), ev[1].elapsed_ms(ev[2])])
    r = np.median(np.array(rows), axis=0)
    return {"n_tris": int(w.scene.shape[1]), "rays": n, "res": list(info.res), "C": info.C1, "P": info.P1,
            "build_ms": round(float(r[0]), 4), "trace_ms": round(float(r[1]), 4),
            "mrays_per_s": round(n / r[1] / 1e3, 1)}


BATCH_BRANCHES = 16   # concurrent rebuild branches of the batched graph


def batched_bench(sa, ctx, n_scenes: int = 64, reps: int = 5) -> dict:
    """Batched multi-scene launch: 64 uniform-in-cube soups, N log-uniform in
    [1e3, 5e5] with edge scale 0.01*N^(-1/3), each with log-uniform
    [1e4, 2e6] random rays; per-scene density, radix width and block size
    from the autotuner.  One captured graph rebuilds every grid and traces
    every batch in one multi-batch launch, the rebuilds on concurrent
    branches; the plain path runs the same rebuilds and one trace launch per
    scene.  graph_serial_rebuilds_ms is the graph with the rebuilds in
    sequence."""
    rng = np.random.default_rng(4)
    ns = _log_uniform(rng, 1_000, 500_000, n_scenes)
    nr = _log_uniform(rng, 10_000, 2_000_000, n_scenes)
    items = []
    t_tune = time.perf_counter()
    kinds = {0: 0, 1: 0}
    radix = {4: 0, 6: 0, 8: 0}
    blocks = {64: 0, 128: 0, 256: 0}
    for k in range(n_scenes):
        s9 = sa.gen_soup(4000 + k, int(ns[k]), float(0.01 * ns[k] ** (-1.0 / 3.0)))
        soup = sa.FaceSoup(ctx, s9)
        rays = sa.RayBatch(ctx, sa.gen_random_rays(6000 + k, int(nr[k]), 0.0, 1.0))
        sample = sa.RayBatch(ctx, sa.gen_random_rays(8000 + k, min(20_000, int(nr[k])), 0.0, 1.0))
        tr = sa.autotune(ctx, soup, sample, int(nr[k]))
        p = tr.best
        kinds[p.kind] += 1
        radix[p.radix_bits] += 1
        blocks[p.trace_block] += 1
        B = sa.FastTwoLevelGridBuilder if p.kind == 1 else sa.FastSimpleGridBuilder
        b = B(ctx, lam=p.lambda_, lambda_leaf=p.lambda_leaf, radix_bits=p.radix_bits, trace_block=p.trace_block)
        b.init(soup)
        g = b.build()
        prim, tb = sa.DeviceBuffer(ctx, 4 * rays.n), sa.DeviceBuffer(ctx, 4 * rays.n)
        items.append((soup, g, rays, prim, tb))
    t_tune = time.perf_counter() - t_tune
    batch = sa.Batch(ctx, n_scenes)
    for soup, g, rays, prim, tb in items:
        batch.add(g, rays, prim, tb)
    batch.commit()
    ctx.sync()
    # the rebuilds run on concurrent branches (sg_rebuild_many), the traces
    # in one multi-batch launch after the join
    with sa.Graph.capture(ctx) as graph:
        batch.rebuild(branches=BATCH_BRANCHES)
        batch.launch()
    with sa.Graph.capture(ctx) as serial:
        for soup, g, rays, prim, tb in items:
            g.rebuild()
        batch.launch()
    e0, e1 = ctx.event(), ctx.event()

    def timeit(fn):
        fn()
        ctx.sync()
        ts = []
        for _ in range(reps):
            e0.record()
            fn()
            e1.record()
            ctx.sync()
            ts.append(e0.elapsed_ms(e1))
        return float(np.median(ts))

    def plain():
        for soup, g, rays, prim, tb in items:
            g.rebuild()
            g.trace(rays, prim=prim, t=tb)

    def plain_trace():
        for soup, g, rays, prim, tb in items:
            g.trace(rays, prim=prim, t=tb)

    ms_graph = timeit(graph.launch)
    ms_serial = timeit(serial.launch)
    ms_plain = timeit(plain)
    ms_batch_trace = timeit(lambda: batch.launch())
    ms_plain_trace = timeit(plain_trace)
    # the batched results equal the per-scene launches
    ref = [(p.download(np.uint32, r.n)) for _, _, r, p, _ in items]
    graph.launch()
    ctx.sync()
    same = all(np.array_equal(p.download(np.uint32, r.n), q) for (_, _, r, p, _), q in zip(items, ref))
    total_rays = int(nr.sum())
    graph.close()
    serial.close()
    return {"scenes": n_scenes, "tris_total": int(ns.sum()), "rays_total": total_rays,
            "autotune_s": round(t_tune, 1), "chosen": {"two_level": kinds[1], "uniform": kinds[0],
                                                       "radix_bits": radix, "block": blocks},
            "rebuild_branches": BATCH_BRANCHES, "graph_ms": round(ms_graph, 3),
            "graph_serial_rebuilds_ms": round(ms_serial, 3), "plain_ms": round(ms_plain, 3),
            "batched_trace_ms": round(ms_batch_trace, 3), "plain_trace_ms": round(ms_plain_trace, 3),
            "graph_mrays_per_s": round(total_rays / ms_graph / 1e3, 1),
            "plain_mrays_per_s": round(total_rays / ms_plain / 1e3, 1),
            "results_equal": bool(same)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-serving", action="store_true", help="skip the served-queries measurement")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched multi-scene measurement")
    ap.add_argument("--no-cluster", action="store_true", help="skip the dense-cluster path measurement")
    ap.add_argument("--no-ruins", action="store_true", help="skip the reference ruins scene measurement")
    ap.add_argument("--no-soup10k", action="store_true", help="skip the 10,000-triangle uniform measurement")
    ap.add_argument("--no-rowmajor", action="store_true", help="skip the row-major primary trace measurement")
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a 64 Mi-element scan (k_scan_up reads 256 MiB, k_scan_down reads and writes "
                         "256 MiB): the reference kernels for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the grid hot path runs on a single MI355X here "
              "(traversal shards by replicating the grid; see the sharding note). Use --gpus 1.",
              file=sys.stderr)
        return 2

    import shinobi_amd as sa
    from shinobi_amd import workloads as W

    w = W.heightfield2m()
    n_tris = w.scene.shape[1]
    n_rays = w.rays.shape[1]
    cam = sa.pinhole(*w.camera)
    ctx = sa.Context(0)
    scene = sa.FaceSoup(ctx, w.scene)
    builder = sa.FastTwoLevelGridBuilder(ctx, **w.build)
    builder.init(scene)
    grid = builder.build()
    info = grid.info()
    if info.status != 0:
        raise RuntimeError(f"build failed: status {info.status}")

    rays = sa.RayBatch(ctx, capacity=n_rays)      # the camera trace writes the primary rays here
    prim = sa.DeviceBuffer(ctx, 4 * n_rays)
    tbuf = sa.DeviceBuffer(ctx, 4 * n_rays)
    brays = sa.RayBatch(ctx, capacity=n_rays)
    bprim = sa.DeviceBuffer(ctx, 4 * n_rays)
    bt = sa.DeviceBuffer(ctx, 4 * n_rays)
    count = sa.DeviceBuffer(ctx, 4)

    # untimed pass with traversal counters: the FLOP model of this workload
    # (per-ray counts, so the order in which the rays are visited does not
    # change it); the device-generated rays equal the host generator's
    st1 = sa.DeviceBuffer(ctx, 16 * n_rays)
    st2 = sa.DeviceBuffer(ctx, 16 * n_rays)
    grid.trace_camera(cam, PRIMARY_TILE, prim=prim, t=tbuf, rays_out=rays)
    ctx.sync()
    if not np.array_equal(rays.download().view(np.uint32), w.rays.view(np.uint32)):
        raise RuntimeError("device camera rays differ from the host generator")
    grid.trace(rays, stats=st1)
    order = sa.bounce_order("exit", grid=grid)
    sa.bounce(ctx, scene, rays, prim, tbuf, w.bounce_seed, brays, count=count, order=order)
    grid.trace(brays, prim=bprim, t=bt, stats=st2, d_count=count.ptr, n=n_rays)
    m = int(count.download(np.uint32, 1)[0])
    f_primary = trace_flops(st1.download(np.uint32, 4 * n_rays).reshape(-1, 4))
    f_bounce = trace_flops(st2.download(np.uint32, 4 * m).reshape(-1, 4))
    del st1, st2

    K, Wm = args.steps, args.warmup
    ev = [[ctx.event() for _ in range(5)] for _ in range(K)]
    ev_w = [ctx.event() for _ in range(5)]

    # a step: events only at its four boundaries (the per-phase split of the
    # rebuild comes from a separate phased loop below, whose nine events
    # inside the build add ~30 us that a production rebuild does not pay)
    def step(e):
        e[0].record()
        grid.rebuild()
        e[1].record()
        grid.trace_camera(cam, PRIMARY_TILE, prim=prim, t=tbuf, rays_out=rays)
        e[2].record()
        sa.bounce(ctx, scene, rays, prim, tbuf, w.bounce_seed, brays, count=count, order=order)
        e[3].record()
        grid.trace(brays, prim=bprim, t=bt, d_count=count.ptr, n=n_rays)
        e[4].record()

    for _ in range(Wm):
        step(ev_w)
    ctx.sync()
    e_start, e_end = ctx.event(), ctx.event()
    t_wall = time.perf_counter()
    e_start.record()
    for k in range(K):
        step(ev[k])
    e_end.record()
    ctx.sync()
    t_wall = time.perf_counter() - t_wall
    total_ms = e_start.elapsed_ms(e_end)

    build_ms, tp_ms, bn_ms, tb_ms = [], [], [], []
    for k in range(K):
        e = ev[k]
        build_ms.append(e[0].elapsed_ms(e[1]))
        tp_ms.append(e[1].elapsed_ms(e[2]))
        bn_ms.append(e[2].elapsed_ms(e[3]))
        tb_ms.append(e[3].elapsed_ms(e[4]))
    # per-phase split of the rebuild (sg_rebuild_phased), outside the timed steps
    ph = {p: [] for p in PHASES}
    phased_ms = []
    ev_p = [ctx.event() for _ in range(9)]
    for k in range(max(K, 5) + 1):
        grid.rebuild_phased(ev_p)
        ctx.sync()
        if k == 0:
            continue
        for i, p in enumerate(PHASES):
            ph[p].append(ev_p[i].elapsed_ms(ev_p[i + 1]))
        phased_ms.append(ev_p[0].elapsed_ms(ev_p[8]))
    # the same primary rays visited in row-major pixel order, for reference
    rowmajor_ms = None
    if not args.no_rowmajor:
        e_a, e_b = ctx.event(), ctx.event()
        rm = []
        for _ in range(6):
            e_a.record()
            grid.trace_camera(cam, (1, 1), prim=prim, t=tbuf, rays_out=rays)
            e_b.record()
            ctx.sync()
            rm.append(e_a.elapsed_ms(e_b))
        rowmajor_ms = round(float(np.median(rm[1:])), 4)
    info2 = grid.info()
    if info2.status != 0 or (info2.P1, info2.P2, info2.C2) != (info.P1, info.P2, info.C2):
        raise RuntimeError("rebuild changed the grid")
    m2 = int(count.download(np.uint32, 1)[0])
    if m2 != m:
        raise RuntimeError("bounce count changed")

    mean = lambda a: float(np.mean(a))  # noqa: E731
    rays_per_step = n_rays + m
    trace_ms = mean(tp_ms) + mean(tb_ms)
    mrays = rays_per_step / trace_ms / 1e3
    bb = build_bytes(n_tris, info.P1, info.C1, info.P2, info.C2)
    total_bytes = sum(bb.values())
    phase_ms = {p: round(mean(v), 4) for p, v in ph.items()}
    phase_gbs = {p: round(bb[p] / (mean(ph[p]) * 1e-3) / 1e9, 1) for p in PHASES}
    build_gbs = total_bytes / (mean(build_ms) * 1e-3) / 1e9

    # roofline of the dominant kernel: the two-level closest-hit traversal,
    # launched twice per step -- k_trace_gen<1,true> on the camera rays,
    # k_trace<1,false,false,false> on the bounce rays; the object describes
    # the longer launch, both launches are listed beside it
    kernel_of = {"primary": "k_trace_gen<1,true> (two-level closest hit, camera rays generated in the kernel)",
                 "bounce": "k_trace<1,false,false,false> (two-level closest hit, bounce rays)"}
    launches = {}
    for name, f, ms in (("primary", f_primary, mean(tp_ms)), ("bounce", f_bounce, mean(tb_ms))):
        tf = f / (ms * 1e-3) / 1e12
        launches[name] = {"ms": round(ms, 4), "flops": f, "achieved": round(tf, 3),
                          "frac": round(tf / FP32_PEAK_TFLOPS, 4), "traffic": traffic_of(f"k_trace {name}"),
                          **counters_of(f"k_trace {name}")}
    dom = max(launches, key=lambda k: launches[k]["ms"])
    dominant = max([("trace_primary", mean(tp_ms)), ("trace_bounce", mean(tb_ms))] +
                   [(p, mean(v)) for p, v in ph.items()], key=lambda x: x[1])
    roofline = {"kernel": kernel_of[dom], "bound": "valu",
                "achieved": launches[dom]["achieved"], "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": launches[dom]["frac"], "traffic": launches[dom]["traffic"],
                "valu_util": launches[dom].get("valu_util"), "lanes_active": launches[dom].get("lanes_active"),
                "l2_hit": launches[dom].get("l2_hit"),
                "flops_per_launch": launches[dom]["flops"], "launches": launches,
                "dominant_phase": dominant[0]}
    copy = hbm_copy_rate(ctx)
    roofline_build = {"bound": "hbm", "achieved": round(build_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                      "frac": round(build_gbs / HBM_PEAK_GBS, 4), "algorithmic_bytes": total_bytes,
                      "measured_copy_gbs": copy["gbs"], "frac_of_measured_copy": round(build_gbs / copy["gbs"], 4),
                      "phase_ms": phase_ms, "phase_gbs": phase_gbs}

    cpu = None
    if not args.no_cpu_baseline:
        threads = args.cpu_threads or host_cpu()["cores"]
        cpu = cpu_baseline(w, n_rays, w.bounce_seed, threads, w_small=W.soup10k())

    out = {
        "metric": "closest-hit Mrays/s + grid build ms (1 GPU)",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": Wm,
        "ms_per_step": round(total_ms / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "heightfield2m: 2,093,058-tri height field, two-level grid "
                               "(lambda 5 / 6 top, lambda 2 leaf), 1920x1080 primary + 1 diffuse bounce per hit",
                   "primary_rays_from": "sg_trace_camera (generated in the trace kernel)",
                   "primary_order": f"{PRIMARY_TILE[0]}x{PRIMARY_TILE[1]} pixel tiles",
                   "n_tris": n_tris, "primary_rays": n_rays, "bounce_rays": m, "res_top": list(info.res),
                   "C1": info.C1, "P1": info.P1, "C2": info.C2, "P2": info.P2},
        "grid_bytes": info.device_bytes,
        "build_ms": round(mean(build_ms), 4),
        "build_ms_median": round(float(np.median(build_ms)), 4),
        "build_ms_phased": round(mean(phased_ms), 4),
        "trace_primary_ms": round(mean(tp_ms), 4),
        "trace_primary_rowmajor_ms": rowmajor_ms,
        "bounce_gen_ms": round(mean(bn_ms), 4),
        "trace_bounce_ms": round(mean(tb_ms), 4),
        "wall_s": round(t_wall, 3),
        "roofline": roofline,
        "roofline_build": roofline_build,
        "hbm_copy": copy,
        "cpu_baseline": cpu,
    }
    if args.calibrate:
        n_cal = 64 << 20
        d_in = sa.DeviceBuffer(ctx, 4 * n_cal)
        d_out = sa.DeviceBuffer(ctx, 4 * n_cal)
        for _ in range(2):
            sa.check(ctx._l.sg_scan_u32(ctx.h, d_in.ptr, d_out.ptr, n_cal, 0, ctx.stream), "sg_scan_u32")
        ctx.sync()
        del d_in, d_out
    if not args.no_soup10k:
        out["soup10k"] = soup10k_bench(sa, W, ctx)
    if not args.no_ruins:
        out["ruins"] = ruins_bench(sa, W, ctx)
    if not args.no_cluster:
        out["cluster_paths"] = cluster_bench(sa, W, ctx)
    if not args.no_batched:
        out["batched"] = batched_bench(sa, ctx)
    if not args.no_serving:
        out["serving"] = serving_bench(sa, W, ctx)
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
