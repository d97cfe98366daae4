// sg_serve.hip -- multi-batch launches, graph capture, the autotuner and the
// serving loop of the C ABI (include/shinobi_grid.h), host side.
//
// The reference traces one ray buffer per persistent launch
// (src/Engine/RayTracingKernels.h:37-136) and its engines call it once per
// frame.  Here a table of (grid, ray batch) items is traced by one launch per
// grid kind (k_trace_items), which is what both the batched multi-scene runs
// and the serving loop build on.
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <vector>

#include "sg_arena.hpp"
#include "sg_internal.hpp"

namespace {

using Clock = std::chrono::steady_clock;

constexpr size_t kAlign = 256;

size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

uint32_t chunks_of(uint32_t n) { return (n + 63u) / 64u; }

}  // namespace

// ================================================================ batching
// The rebuild plan of a batch's uniform grids: the launches of the
// multi-grid rebuild (launch_multi / launch_multi_sort) in enqueue order,
// each over its entries [e0, e0 + ne) of the entry table.
struct PlanLaunch {
    bool sort;
    int phase;        // MultiPhase, or MultiSortStep for a sort launch
    int rb, pass;     // sort launches
    uint32_t e0, ne, blocks;
};

// entries of one job in the plan at most: the six phases plus three
// launches per sort pass (keys of at most 32 bits: 8 passes of 4 bits)
constexpr uint32_t kPlanEntriesPerJob = MP_COUNT + 3 * 8;

struct sg_batch {
    sg_ctx* ctx = nullptr;
    uint32_t cap = 0;
    std::vector<TraceItem> items;
    std::vector<sg_grid*> grids;   // the items' distinct grids, in sg_batch_add order
    uint32_t next_chunk = 0;
    TraceItem* pinned = nullptr;   // staging of the table for the upload
    TraceItem* dev = nullptr;      // device table
    uint32_t* dev_status = nullptr;// per item: the grid's status bits of the last launch
    hipEvent_t uploaded = nullptr; // the last upload out of `pinned`
    ItemSplit dev_split;           // layout of the device table
    // rebuild plan (sg_batch_commit): uniform grids as jobs of the
    // multi-grid launches, two-level grids rebuilt on branches
    UniJob* pinned_jobs = nullptr;
    MultiEntry* pinned_entries = nullptr;
    UniJob* dev_jobs = nullptr;
    MultiEntry* dev_entries = nullptr;
    std::vector<PlanLaunch> plan;
    std::vector<sg_grid*> plan_uniform, plan_two;
    // each planned grid's scene at the commit: the uniform jobs hold that
    // scene's device arrays, so a grid rebuilt from another scene since
    // must be committed again before sg_batch_rebuild
    std::vector<sg_scene*> plan_uniform_scene, plan_two_scene;
    uint64_t plan_tris = 0;
};

// Lays out the rebuild plan of the batch's grids in the pinned staging
// buffers; returns the number of entries.  The launch order is the one of
// a single uniform rebuild (zeroing with bounds, params, top count / emit, the
// sort passes, ranges, reference records); the sort passes go by radix
// width, each pass over the grids that sort with that width and have that
// many passes.
static uint32_t plan_rebuild(sg_batch* b) {
    b->plan.clear();
    b->plan_uniform.clear();
    b->plan_two.clear();
    b->plan_uniform_scene.clear();
    b->plan_two_scene.clear();
    b->plan_tris = 0;
    for (sg_grid* g : b->grids) {
        if (uni_job_of(g, &b->pinned_jobs[b->plan_uniform.size()])) {
            b->plan_uniform.push_back(g);
            b->plan_uniform_scene.push_back(g->scene);
            b->plan_tris += g->n_tris;
        } else {
            b->plan_two.push_back(g);
            b->plan_two_scene.push_back(g->scene);
        }
    }
    const UniJob* jobs = b->pinned_jobs;
    const uint32_t nj = (uint32_t)b->plan_uniform.size();
    uint32_t ne = 0;
    auto add = [&](bool sort, int phase, int rb, int pass, auto&& blocks_of) {
        PlanLaunch L{sort, phase, rb, pass, ne, 0, 0};
        for (uint32_t i = 0; i < nj; ++i) {
            const uint32_t nb = blocks_of(jobs[i]);
            if (nb == 0) continue;
            b->pinned_entries[ne++] = MultiEntry{i, L.blocks};
            L.blocks += nb;
        }
        L.ne = ne - L.e0;
        if (L.ne) b->plan.push_back(L);
    };
    for (int ph : {MP_INIT, MP_PARAMS, MP_TOP_COUNT, MP_TOP_EMIT})
        add(false, ph, 0, 0, [&](const UniJob& j) { return multi_blocks(j, (MultiPhase)ph); });
    for (int rb : {8, 6, 4}) {
        int passes = 0;   // at most 8: sort_passes of 32-bit keys at 4 bits
        for (uint32_t i = 0; i < nj; ++i)
            if (jobs[i].rb == rb) passes = std::max(passes, jobs[i].passes);
        for (int p = 0; p < passes; ++p)
            for (int step : {MS_UP, MS_SCAN, MS_DOWN})
                add(true, step, rb, p, [&](const UniJob& j) {
                    return j.rb == rb && j.passes > p ? multi_sort_blocks(j, (MultiSortStep)step) : 0u;
                });
    }
    for (int ph : {MP_RANGES, MP_REFS})
        add(false, ph, 0, 0, [&](const UniJob& j) { return multi_blocks(j, (MultiPhase)ph); });
    return ne;
}

static sg_status run_plan(void* arg, hipStream_t st) {
    const sg_batch* b = static_cast<const sg_batch*>(arg);
    for (const PlanLaunch& L : b->plan) {
        if (L.sort)
            launch_multi_sort((MultiSortStep)L.phase, L.rb, L.pass, b->dev_jobs, b->dev_entries + L.e0, L.ne,
                              L.blocks, st);
        else
            launch_multi((MultiPhase)L.phase, b->dev_jobs, b->dev_entries + L.e0, L.ne, L.blocks, st);
    }
    return SG_OK;
}

extern "C" {

sg_status sg_batch_create(sg_ctx* ctx, uint32_t max_items, sg_batch** out) {
    if (!ctx || !out || max_items == 0) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    sg_batch* b = new sg_batch();
    b->ctx = ctx;
    b->cap = max_items;
    b->items.reserve(max_items);
    const size_t n_entries = (size_t)kPlanEntriesPerJob * max_items;
    if (hipHostMalloc((void**)&b->pinned, sizeof(TraceItem) * max_items, hipHostMallocDefault) != hipSuccess ||
        hipMalloc((void**)&b->dev, sizeof(TraceItem) * max_items) != hipSuccess ||
        hipMalloc((void**)&b->dev_status, sizeof(uint32_t) * max_items) != hipSuccess ||
        hipHostMalloc((void**)&b->pinned_jobs, sizeof(UniJob) * max_items, hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc((void**)&b->pinned_entries, sizeof(MultiEntry) * n_entries, hipHostMallocDefault) !=
            hipSuccess ||
        hipMalloc((void**)&b->dev_jobs, sizeof(UniJob) * max_items) != hipSuccess ||
        hipMalloc((void**)&b->dev_entries, sizeof(MultiEntry) * n_entries) != hipSuccess ||
        hipEventCreate(&b->uploaded) != hipSuccess) {
        if (b->pinned) hipHostFree(b->pinned);
        if (b->dev) hipFree(b->dev);
        if (b->dev_status) hipFree(b->dev_status);
        if (b->pinned_jobs) hipHostFree(b->pinned_jobs);
        if (b->pinned_entries) hipHostFree(b->pinned_entries);
        if (b->dev_jobs) hipFree(b->dev_jobs);
        if (b->dev_entries) hipFree(b->dev_entries);
        delete b;
        return SG_ERR_OUT_OF_MEMORY;
    }
    *out = b;
    return SG_OK;
}

sg_status sg_batch_destroy(sg_batch* b) {
    if (!b) return SG_ERR_INVALID_ARG;
    hipSetDevice(b->ctx->device);
    hipEventSynchronize(b->uploaded);
    hipEventDestroy(b->uploaded);
    hipHostFree(b->pinned);
    hipFree(b->dev);
    hipFree(b->dev_status);
    hipHostFree(b->pinned_jobs);
    hipHostFree(b->pinned_entries);
    hipFree(b->dev_jobs);
    hipFree(b->dev_entries);
    delete b;
    return SG_OK;
}

sg_status sg_batch_clear(sg_batch* b) {
    if (!b) return SG_ERR_INVALID_ARG;
    b->items.clear();
    b->grids.clear();
    b->next_chunk = 0;
    // the uploaded table and the rebuild plan go with the items (their grids
    // may be destroyed next): a launch or rebuild before the next commit
    // does nothing
    b->dev_split = ItemSplit{};
    b->plan.clear();
    b->plan_uniform.clear();
    b->plan_two.clear();
    b->plan_uniform_scene.clear();
    b->plan_two_scene.clear();
    b->plan_tris = 0;
    return SG_OK;
}

sg_status sg_batch_add(sg_batch* b, sg_grid* g, const sg_rays* rays, uint32_t n, uint32_t* d_prim, float* d_t) {
    if (!b || !g || !rays || !d_prim || !d_t) return SG_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (!rays->o[i] || !rays->d[i]) return SG_ERR_INVALID_ARG;
    if (!rays->tmax || b->items.size() >= b->cap) return SG_ERR_INVALID_ARG;
    if ((uint64_t)b->next_chunk + chunks_of(n) >= (1ull << 32)) return SG_ERR_INVALID_ARG;
    TraceItem it;
    it.g = grid_view(g);
    it.rays = ray_soa(rays);
    it.prim = d_prim;
    it.t = d_t;
    it.status = b->dev_status + b->items.size();
    it.n = n;
    it.chunk0 = b->next_chunk;
    b->next_chunk += chunks_of(n);
    b->items.push_back(it);
    if (std::find(b->grids.begin(), b->grids.end(), g) == b->grids.end()) b->grids.push_back(g);
    return SG_OK;
}

sg_status sg_batch_commit(sg_batch* b, void* stream) {
    if (!b) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    HIPCHK(hipEventSynchronize(b->uploaded));   // the staging buffer is free again
    const size_t bytes = sizeof(TraceItem) * b->items.size();
    b->dev_split = split_items(b->items.data(), (uint32_t)b->items.size(), b->pinned);
    if (bytes) HIPCHK(hipMemcpyAsync(b->dev, b->pinned, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    const uint32_t n_entries = plan_rebuild(b);
    if (!b->plan_uniform.empty()) {
        HIPCHK(hipMemcpyAsync(b->dev_jobs, b->pinned_jobs, sizeof(UniJob) * b->plan_uniform.size(),
                              hipMemcpyHostToDevice, (hipStream_t)stream));
        HIPCHK(hipMemcpyAsync(b->dev_entries, b->pinned_entries, sizeof(MultiEntry) * n_entries,
                              hipMemcpyHostToDevice, (hipStream_t)stream));
    }
    // items without rays are never traced: their status reads 0
    if (!b->items.empty())
        HIPCHK(hipMemsetAsync(b->dev_status, 0, sizeof(uint32_t) * b->items.size(), (hipStream_t)stream));
    HIPCHK(hipEventRecord(b->uploaded, (hipStream_t)stream));
    return SG_OK;
}

sg_status sg_batch_get_status(sg_batch* b, void* stream, sg_status* out, uint32_t n) {
    if (!b || (!out && n) || n > b->items.size()) return SG_ERR_INVALID_ARG;
    if (n == 0) return SG_OK;
    HIPCHK(hipSetDevice(b->ctx->device));
    std::vector<uint32_t> bits(n);
    HIPCHK(hipMemcpyAsync(bits.data(), b->dev_status, sizeof(uint32_t) * n, hipMemcpyDeviceToHost,
                          (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    for (uint32_t i = 0; i < n; ++i) out[i] = untraversable(bits[i]) ? status_from_bits(bits[i]) : SG_OK;
    return SG_OK;
}

sg_status sg_batch_rebuild(sg_batch* b, uint32_t branches, void* stream) {
    if (!b || branches > SG_REBUILD_BRANCHES_MAX) return SG_ERR_INVALID_ARG;
    for (size_t i = 0; i < b->plan_uniform.size(); ++i)
        if (b->plan_uniform[i]->scene != b->plan_uniform_scene[i]) return SG_ERR_INVALID_ARG;
    for (size_t i = 0; i < b->plan_two.size(); ++i)
        if (b->plan_two[i]->scene != b->plan_two_scene[i]) return SG_ERR_INVALID_ARG;
    clear_last_error();
    const ForkTask task{run_plan, b, b->plan_tris};
    const sg_status r =
        rebuild_many_with(b->ctx, (uint32_t)b->plan_two.size(), b->plan_two_scene.data(), b->plan_two.data(),
                          branches, (hipStream_t)stream, b->plan.empty() ? nullptr : &task);
    // as sg_rebuild_many: the caller's stream orders the grids' later work
    for (sg_grid* g : b->plan_uniform) g->stream = (hipStream_t)stream;
    if (r != SG_OK) return r;
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_batch_launch(sg_batch* b, int shadow, void* stream) {
    if (!b) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(b->ctx->device));
    clear_last_error();
    launch_trace_items(b->dev, b->dev_split, shadow != 0, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

}  // extern "C"

// ================================================================== graphs
struct sg_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

extern "C" {

sg_status sg_graph_begin(sg_ctx* ctx, void* stream) {
    if (!ctx || !stream) return SG_ERR_INVALID_ARG;   // the null stream cannot be captured
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamBeginCapture((hipStream_t)stream, hipStreamCaptureModeThreadLocal));
    return SG_OK;
}

sg_status sg_graph_end(sg_ctx* ctx, void* stream, sg_graph** out) {
    if (!ctx || !stream || !out) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    sg_graph* g = new sg_graph();
    hipError_t e = hipStreamEndCapture((hipStream_t)stream, &g->graph);
    if (e == hipSuccess) e = hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        if (g->graph) hipGraphDestroy(g->graph);
        delete g;
        return hip_status(e);
    }
    *out = g;
    return SG_OK;
}

sg_status sg_graph_launch(sg_graph* g, void* stream) {
    if (!g) return SG_ERR_INVALID_ARG;
    HIPCHK(hipGraphLaunch(g->exec, (hipStream_t)stream));
    return SG_OK;
}

sg_status sg_graph_destroy(sg_graph* g) {
    if (!g) return SG_ERR_INVALID_ARG;
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    delete g;
    return SG_OK;
}

// =============================================================== autotuner
sg_status sg_autotune(sg_ctx* ctx, sg_scene* scene, const sg_rays* sample, uint32_t n_sample, uint64_t n_rays,
                      int kind_mask, void* stream, sg_tune_result* out) {
    if (!ctx || !scene || !sample || !out || n_sample == 0 || (kind_mask & 3) == 0) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = (hipStream_t)stream;
    uint32_t* d_prim = nullptr;
    float* d_t = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (hipMalloc((void**)&d_prim, 4ull * n_sample) != hipSuccess ||
        hipMalloc((void**)&d_t, 4ull * n_sample) != hipSuccess) {
        hipFree(d_prim);
        return SG_ERR_OUT_OF_MEMORY;
    }
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    const double scale = (double)n_rays / (double)n_sample;
    uint32_t measured = 0;

    // build once (sizing), then the best of three timed rebuilds and traces
    auto measure = [&](const sg_build_params& p, float* bms, float* tms) -> bool {
        sg_grid* g = nullptr;
        if (sg_build(ctx, scene, &p, stream, &g) != SG_OK) return false;
        sg_grid_info info;
        if (sg_grid_get_info(g, &info) != SG_OK || info.status != SG_OK) {
            sg_grid_destroy(g);
            return false;
        }
        float best_b = 1e30f, best_t = 1e30f, ms = 0.f;
        for (int r = 0; r < 3; ++r) {
            hipEventRecord(e0, st);
            sg_rebuild(ctx, scene, g, stream);
            hipEventRecord(e1, st);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
            best_b = std::min(best_b, ms);
            hipEventRecord(e0, st);
            sg_trace_closest(ctx, g, sample, n_sample, nullptr, d_prim, d_t, nullptr, stream);
            hipEventRecord(e1, st);
            hipEventSynchronize(e1);
            hipEventElapsedTime(&ms, e0, e1);
            best_t = std::min(best_t, ms);
        }
        sg_grid_destroy(g);
        ++measured;
        *bms = best_b;
        *tms = best_t;
        return true;
    };

    bool have = false;
    sg_build_params best;
    float best_b = 0.f, best_t = 0.f;
    double best_cost = 1e300;
    auto consider = [&](const sg_build_params& p) {
        float b, t;
        if (!measure(p, &b, &t)) return;
        const double cost = b + t * scale;
        if (!have || cost < best_cost) {
            have = true;
            best = p;
            best_b = b;
            best_t = t;
            best_cost = cost;
        }
    };
    // 1. density (default radix width 8 and block size 64)
    if (kind_mask & 1) {
        for (float lam : {1.f, 2.f, 3.f, 5.f, 8.f}) {
            sg_build_params p;
            sg_build_params_default(&p, SG_GRID_UNIFORM);
            p.lambda = lam;
            consider(p);
        }
    }
    if (kind_mask & 2) {
        for (float lam : {3.f, 5.f, 8.f})
            for (float ll : {1.2f, 2.f, 3.f}) {
                sg_build_params p;
                sg_build_params_default(&p, SG_GRID_TWO_LEVEL);
                p.lambda = lam;
                p.lambda_leaf = ll;
                consider(p);
            }
    }
    // 2. sort radix width, 3. traversal block size, each from the best so far
    if (have) {
        const sg_build_params base = best;
        for (int rb : {4, 6}) {
            sg_build_params p = base;
            p.radix_bits = rb;
            consider(p);
        }
        const sg_build_params base2 = best;
        for (int blk : {64, 128, 256}) {
            if (blk == base2.trace_block) continue;
            sg_build_params p = base2;
            p.trace_block = blk;
            consider(p);
        }
    }
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    hipFree(d_prim);
    hipFree(d_t);
    if (!have) return SG_ERR_INVALID_SCENE;
    out->best = best;
    out->build_ms = best_b;
    out->trace_ms = best_t;
    out->cost_ms = (float)best_cost;
    out->candidates = measured;
    return SG_OK;
}

}  // extern "C"

// ================================================================= serving
namespace {

struct Request {
    uint64_t id = 0;
    uint32_t scene = 0;
    uint32_t n = 0;
    const float* rays[7] = {};
    uint32_t* prim_out = nullptr;
    float* t_out = nullptr;
    size_t bytes = 0;            // device bytes while in flight
    int gi = 0;                  // grid of a dynamic scene the request runs on
    uint64_t version = 0;        // the mesh version of that grid (0 for static scenes)
    GridView view{};             // that grid, read under the server lock at admission
    int trace_block = 64;        // its traversal block size (plain mode)
    Clock::time_point t_submit, t_done;
    sg_status status = SG_OK;
    bool done = false;
};

struct Slot {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    TraceItem* pinned_items = nullptr;
    TraceItem* dev_items = nullptr;
    uint32_t* dev_status = nullptr;      // per request: its grid's status bits (batched mode)
    uint32_t* pinned_status = nullptr;   // the same, copied back with the results
    bool busy = false;
    std::vector<Request*> reqs;
    size_t off = 0, bytes = 0, rays_bytes = 0;   // arena region of the wave
};

struct SceneEntry {
    bool dynamic = false;
    sg_grid* grid[2] = {nullptr, nullptr};
    sg_scene* mesh[2] = {nullptr, nullptr};
    int front = 0;
    bool swap_pending = false;      // back grid rebuilt on the rebuild stream
    bool updating = false;          // an update_dynamic call owns the back copy
    sg_build_params params{};       // build parameters of a dynamic scene
    hipEvent_t rebuilt = nullptr;
    uint32_t in_use[2] = {0, 0};    // admitted requests that read each grid
    uint64_t version[2] = {0, 0};   // mesh version each grid holds: k after the k-th update
    uint64_t updates = 0;           // successful update_dynamic calls so far
};

size_t request_bytes(uint32_t n) { return align_up((size_t)n * 4, kAlign) * 9; }

}  // namespace

struct sg_server {
    sg_ctx* ctx = nullptr;
    sg_server_config cfg{};
    char* dev = nullptr;       // device arena (hbm_budget bytes)
    Arena arena;
    std::vector<Slot> slots;
    // a deque: adding a scene never moves the entries that waves and
    // update_dynamic calls refer to
    std::deque<SceneEntry> scenes;
    hipStream_t rebuild_st = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> pending;
    std::map<uint64_t, Request*> live;   // submitted, not yet waited for
    uint64_t next_id = 1;
    uint32_t admitted = 0;               // requests in busy slots
    bool stop = false;
    std::thread worker;

    // statistics since the last reset
    std::vector<double> lat_us;
    uint64_t st_rays = 0, st_waves = 0;
    uint64_t regrown = 0;   // dynamic rebuilds that outgrew their grid and were built again
    bool have_first = false;
    Clock::time_point first_submit, last_done;
};

namespace {

// upload the wave's rays straight from the callers' buffers (one copy when
// a request's seven arrays are one contiguous SoA block), trace, and copy
// the results straight back; the copies are DMA transfers when the callers'
// memory is pinned (sg_host_alloc)
sg_status launch_wave_body(sg_server* s, Slot& sl) {
    clear_last_error();
    const bool plain = s->cfg.mode == SG_SERVE_PLAIN;
    char* dbase = s->dev + sl.off;
    // device layout: [rays of every request][prim, t of every request]
    size_t ro = 0;
    std::vector<size_t> rofs(sl.reqs.size());
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        rofs[k] = ro;
        ro += align_up((size_t)sl.reqs[k]->n * 4, kAlign) * 7;
    }
    size_t oo = ro;
    std::vector<size_t> oofs(sl.reqs.size());
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        oofs[k] = oo;
        oo += align_up((size_t)sl.reqs[k]->n * 4, kAlign) * 2;
    }
    uint32_t chunk = 0;
    std::vector<TraceItem> wave_items(plain ? 0 : sl.reqs.size());
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        const Request* r = sl.reqs[k];
        const size_t nb = (size_t)r->n * 4, a = align_up(nb, kAlign);
        bool contiguous = a == nb;
        for (int c = 1; c < 7 && contiguous; ++c) contiguous = r->rays[c] == r->rays[0] + (size_t)c * r->n;
        if (contiguous) {
            HIPCHK(hipMemcpyAsync(dbase + rofs[k], r->rays[0], 7 * nb, hipMemcpyHostToDevice, sl.st));
        } else {
            for (int c = 0; c < 7; ++c)
                HIPCHK(hipMemcpyAsync(dbase + rofs[k] + c * a, r->rays[c], nb, hipMemcpyHostToDevice, sl.st));
        }
        TraceItem it;
        it.g = r->view;
        for (int c = 0; c < 7; ++c) it.rays.a[c] = reinterpret_cast<const float*>(dbase + rofs[k] + c * a);
        it.prim = reinterpret_cast<uint32_t*>(dbase + oofs[k]);
        it.t = reinterpret_cast<float*>(dbase + oofs[k] + a);
        it.status = sl.dev_status + k;
        it.n = r->n;
        it.chunk0 = chunk;
        chunk += chunks_of(r->n);
        if (plain) {
            launch_trace(it.g, it.rays, it.n, nullptr, it.prim, it.t, nullptr, false, r->trace_block, sl.st);
            HIPCHK(hipMemcpyAsync(sl.pinned_status + k, &it.g.hdr->status, sizeof(uint32_t), hipMemcpyDeviceToHost,
                                  sl.st));
        } else {
            wave_items[k] = it;
        }
    }
    if (!plain) {
        const ItemSplit split = split_items(wave_items.data(), (uint32_t)sl.reqs.size(), sl.pinned_items);
        HIPCHK(hipMemcpyAsync(sl.dev_items, sl.pinned_items, sizeof(TraceItem) * sl.reqs.size(),
                              hipMemcpyHostToDevice, sl.st));
        launch_trace_items(sl.dev_items, split, false, sl.st);
        HIPCHK(hipMemcpyAsync(sl.pinned_status, sl.dev_status, sizeof(uint32_t) * sl.reqs.size(),
                              hipMemcpyDeviceToHost, sl.st));
    }
    HIPCHK(hipGetLastError());
    for (size_t k = 0; k < sl.reqs.size(); ++k) {
        const Request* r = sl.reqs[k];
        const size_t nb = (size_t)r->n * 4, a = align_up(nb, kAlign);
        HIPCHK(hipMemcpyAsync(r->prim_out, dbase + oofs[k], nb, hipMemcpyDeviceToHost, sl.st));
        HIPCHK(hipMemcpyAsync(r->t_out, dbase + oofs[k] + a, nb, hipMemcpyDeviceToHost, sl.st));
    }
    HIPCHK(hipEventRecord(sl.ev, sl.st));
    return SG_OK;
}

// A wave that fails part way may have enqueued copies and kernels that
// still read and write its arena region and the callers' buffers: wait for
// them and record the slot's event on every exit path, so the slot is
// released only once the stream has drained.
sg_status launch_wave(sg_server* s, Slot& sl) {
    const sg_status e = launch_wave_body(s, sl);
    if (e != SG_OK) {
        (void)hipStreamSynchronize(sl.st);
        (void)hipEventRecord(sl.ev, sl.st);
        (void)hipEventSynchronize(sl.ev);
        clear_last_error();
    }
    return e;
}

void serve_loop(sg_server* s) {
    hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        bool progress = false;
        // 1. completed waves
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            if (hipEventQuery(sl.ev) != hipSuccess) continue;
            const auto now = Clock::now();
            for (size_t k = 0; k < sl.reqs.size(); ++k) {
                Request* r = sl.reqs[k];
                // a grid that failed or outgrew its buffers was traversed as
                // all misses: the request fails with the grid's status
                if (r->status == SG_OK && untraversable(sl.pinned_status[k]))
                    r->status = status_from_bits(sl.pinned_status[k]);
                s->scenes[r->scene].in_use[r->gi]--;
                r->t_done = now;
                r->done = true;
                s->lat_us.push_back(std::chrono::duration<double, std::micro>(now - r->t_submit).count());
                s->st_rays += r->n;
            }
            s->last_done = now;
            ++s->st_waves;
            s->arena.release(sl.off, sl.bytes);
            s->admitted -= (uint32_t)sl.reqs.size();
            sl.reqs.clear();
            sl.busy = false;
            progress = true;
            s->cv_done.notify_all();
        }
        // 2. dynamic scenes whose back grid finished rebuilding
        for (SceneEntry& se : s->scenes) {
            if (se.dynamic && se.swap_pending && hipEventQuery(se.rebuilt) == hipSuccess) {
                se.front = 1 - se.front;
                se.swap_pending = false;
                s->cv_done.notify_all();
            }
        }
        // 3. a new wave from the pending requests
        Slot* free_slot = nullptr;
        for (Slot& sl : s->slots)
            if (!sl.busy) {
                free_slot = &sl;
                break;
            }
        if (free_slot && !s->pending.empty()) {
            const bool plain = s->cfg.mode == SG_SERVE_PLAIN;
            size_t bytes = 0;
            uint64_t rays = 0;
            size_t take = 0;
            for (Request* r : s->pending) {
                if (take > 0 && plain) break;
                if (s->admitted + take + 1 > s->cfg.max_requests) break;
                if (take > 0 && s->cfg.max_wave_rays && rays + r->n > s->cfg.max_wave_rays) break;
                bytes += r->bytes;
                rays += r->n;
                ++take;
            }
            size_t off = 0;
            // shrink the wave until its region fits the free arena
            while (take > 0 && !s->arena.alloc(bytes, &off)) {
                --take;
                bytes -= s->pending[take]->bytes;
            }
            if (take > 0) {
                Slot& sl = *free_slot;
                sl.reqs.assign(s->pending.begin(), s->pending.begin() + take);
                s->pending.erase(s->pending.begin(), s->pending.begin() + take);
                for (Request* r : sl.reqs) {
                    SceneEntry& se = s->scenes[r->scene];
                    r->gi = se.dynamic ? se.front : 0;
                    r->version = se.dynamic ? se.version[r->gi] : 0;
                    se.in_use[r->gi]++;
                    r->view = grid_view(se.grid[r->gi]);
                    r->trace_block = se.grid[r->gi]->p.trace_block;
                }
                sl.off = off;
                sl.bytes = bytes;
                sl.busy = true;
                s->admitted += (uint32_t)take;
                lk.unlock();
                const sg_status e = launch_wave(s, sl);
                lk.lock();
                if (e != SG_OK)
                    for (Request* r : sl.reqs) r->status = e;
                progress = true;
            }
        }
        bool busy = false;
        for (Slot& sl : s->slots) busy |= sl.busy;
        for (const SceneEntry& se : s->scenes) busy |= se.swap_pending;
        if (s->stop && s->pending.empty() && !busy) break;
        if (!progress) {
            if (busy || (s->pending.size() && !free_slot)) {
                // the device is working: poll its events at a short interval
                s->cv_work.wait_for(lk, std::chrono::microseconds(20));
            } else {
                s->cv_work.wait_for(lk, std::chrono::milliseconds(5));
            }
        }
    }
}

}  // namespace

extern "C" {

sg_status sg_server_create(sg_ctx* ctx, const sg_server_config* cfg, sg_server** out) {
    if (!ctx || !cfg || !out || cfg->hbm_budget < (1u << 20) || cfg->max_requests == 0) return SG_ERR_INVALID_ARG;
    if (cfg->mode != SG_SERVE_BATCHED && cfg->mode != SG_SERVE_PLAIN) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    sg_server* s = new sg_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    if (s->cfg.mode == SG_SERVE_PLAIN) s->cfg.max_waves = 1;
    if (s->cfg.max_waves == 0) s->cfg.max_waves = 2;
    const size_t budget = align_up(cfg->hbm_budget, kAlign);
    s->cfg.hbm_budget = budget;
    s->arena = Arena(budget);
    bool ok = hipMalloc((void**)&s->dev, budget) == hipSuccess &&
              hipStreamCreateWithFlags(&s->rebuild_st, hipStreamNonBlocking) == hipSuccess;
    s->slots.resize(s->cfg.max_waves);
    for (Slot& sl : s->slots) {
        ok = ok && hipStreamCreateWithFlags(&sl.st, hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) == hipSuccess &&
             hipHostMalloc((void**)&sl.pinned_items, sizeof(TraceItem) * s->cfg.max_requests,
                           hipHostMallocDefault) == hipSuccess &&
             hipMalloc((void**)&sl.dev_items, sizeof(TraceItem) * s->cfg.max_requests) == hipSuccess &&
             hipHostMalloc((void**)&sl.pinned_status, sizeof(uint32_t) * s->cfg.max_requests,
                           hipHostMallocDefault) == hipSuccess &&
             hipMalloc((void**)&sl.dev_status, sizeof(uint32_t) * s->cfg.max_requests) == hipSuccess;
    }
    if (!ok) {
        s->stop = true;
        sg_server_destroy(s);
        return SG_ERR_OUT_OF_MEMORY;
    }
    s->worker = std::thread(serve_loop, s);
    *out = s;
    return SG_OK;
}

sg_status sg_server_destroy(sg_server* s) {
    if (!s) return SG_ERR_INVALID_ARG;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->ctx->device);
    if (s->rebuild_st) hipStreamSynchronize(s->rebuild_st);
    for (Slot& sl : s->slots) {
        if (sl.st) hipStreamSynchronize(sl.st);
        if (sl.st) hipStreamDestroy(sl.st);
        if (sl.ev) hipEventDestroy(sl.ev);
        if (sl.pinned_items) hipHostFree(sl.pinned_items);
        if (sl.dev_items) hipFree(sl.dev_items);
        if (sl.pinned_status) hipHostFree(sl.pinned_status);
        if (sl.dev_status) hipFree(sl.dev_status);
    }
    for (SceneEntry& se : s->scenes) {
        if (!se.dynamic) continue;
        for (int b = 0; b < 2; ++b) {
            if (se.grid[b]) sg_grid_destroy(se.grid[b]);
            if (se.mesh[b]) sg_scene_destroy(se.mesh[b]);
        }
        if (se.rebuilt) hipEventDestroy(se.rebuilt);
    }
    for (auto& kv : s->live) delete kv.second;
    for (Request* r : s->pending) delete r;
    if (s->rebuild_st) hipStreamDestroy(s->rebuild_st);
    if (s->dev) hipFree(s->dev);
    delete s;
    return SG_OK;
}

sg_status sg_server_add_grid(sg_server* s, sg_grid* g, uint32_t* id) {
    if (!s || !g || !id) return SG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    SceneEntry se;
    se.grid[0] = g;
    s->scenes.push_back(se);
    *id = (uint32_t)s->scenes.size() - 1;
    return SG_OK;
}

sg_status sg_server_add_dynamic(sg_server* s, const float* const soa9[9], uint32_t n, const sg_build_params* p,
                                uint32_t* id) {
    if (!s || !soa9 || !p || !id || n == 0) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(s->ctx->device));
    SceneEntry se;
    se.dynamic = true;
    se.params = *p;
    sg_status r = SG_OK;
    for (int b = 0; b < 2 && r == SG_OK; ++b) {
        r = sg_scene_create(s->ctx, soa9, n, 0, s->rebuild_st, &se.mesh[b]);
        if (r == SG_OK) r = sg_build(s->ctx, se.mesh[b], p, s->rebuild_st, &se.grid[b]);
    }
    if (r == SG_OK && hipEventCreateWithFlags(&se.rebuilt, hipEventDisableTiming) != hipSuccess) r = SG_ERR_HIP;
    if (r == SG_OK && hipStreamSynchronize(s->rebuild_st) != hipSuccess) r = SG_ERR_HIP;
    if (r != SG_OK) {
        for (int b = 0; b < 2; ++b) {
            if (se.grid[b]) sg_grid_destroy(se.grid[b]);
            if (se.mesh[b]) sg_scene_destroy(se.mesh[b]);
        }
        return r;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    s->scenes.push_back(se);
    *id = (uint32_t)s->scenes.size() - 1;
    return SG_OK;
}

sg_status sg_server_update_dynamic(sg_server* s, uint32_t id, const float* const soa9[9]) {
    if (!s || !soa9) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(s->ctx->device));
    int back;
    SceneEntry* sep = nullptr;
    {
        std::unique_lock<std::mutex> lk(s->mu);
        if (id >= s->scenes.size() || !s->scenes[id].dynamic) return SG_ERR_INVALID_ARG;
        // one rebuild at a time (the previous one has been swapped in), and
        // the requests still reading the back grid finish before it is
        // overwritten; new requests only take the front grid
        s->cv_done.wait(lk, [&] {
            const SceneEntry& se = s->scenes[id];
            return !se.updating && !se.swap_pending && se.in_use[1 - se.front] == 0;
        });
        s->scenes[id].updating = true;
        back = 1 - s->scenes[id].front;
        sep = &s->scenes[id];
    }
    // the entry does not move (deque) and, while `updating` is set, only
    // this call touches its back mesh and grid (waves read the front one)
    SceneEntry& se = *sep;
    sg_status r = sg_scene_update(se.mesh[back], soa9, 0, s->rebuild_st);
    if (r == SG_OK) r = sg_rebuild(s->ctx, se.mesh[back], se.grid[back], s->rebuild_st);
    // A mesh that deformed past the grid's capacity slack overflows the
    // rebuild (the grid would be traversed as all misses): build the back
    // grid again, sized for the new mesh, before it is swapped in.
    sg_grid_info info;
    if (r == SG_OK) r = sg_grid_get_info(se.grid[back], &info);
    if (r == SG_OK && info.status == SG_ERR_CAPACITY) {
        sg_grid* g2 = nullptr;
        r = sg_build(s->ctx, se.mesh[back], &se.params, s->rebuild_st, &g2);
        if (r == SG_OK) {
            sg_grid_destroy(se.grid[back]);
            std::lock_guard<std::mutex> g(s->mu);
            se.grid[back] = g2;
            ++s->regrown;
        }
    } else if (r == SG_OK && info.status != SG_OK && info.status != SG_ERR_LEAF_RES &&
               info.status != SG_ERR_NEG_EXTENT) {
        r = info.status;   // the new mesh cannot be gridded: keep the front one
    }
    if (r == SG_OK && hipEventRecord(se.rebuilt, s->rebuild_st) != hipSuccess) r = SG_ERR_HIP;
    std::unique_lock<std::mutex> lk(s->mu);
    if (r != SG_OK) {
        se.updating = false;
        s->cv_done.notify_all();
        return r;
    }
    se.updating = false;
    se.version[back] = ++se.updates;
    se.swap_pending = true;
    s->cv_work.notify_all();
    // returns once requests admitted from now on see the new mesh
    s->cv_done.wait(lk, [&] { return !s->scenes[id].swap_pending; });
    return SG_OK;
}

sg_status sg_server_submit(sg_server* s, uint32_t scene, const float* const rays7[7], uint32_t n,
                           uint32_t* prim_out, float* t_out, uint64_t* id) {
    if (!s || !rays7 || !prim_out || !t_out || !id || n == 0) return SG_ERR_INVALID_ARG;
    for (int c = 0; c < 7; ++c)
        if (!rays7[c]) return SG_ERR_INVALID_ARG;
    const size_t bytes = request_bytes(n);
    if (bytes > s->cfg.hbm_budget) return SG_ERR_BUDGET;
    Request* r = new Request();
    r->scene = scene;
    r->n = n;
    for (int c = 0; c < 7; ++c) r->rays[c] = rays7[c];
    r->prim_out = prim_out;
    r->t_out = t_out;
    r->bytes = bytes;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (scene >= s->scenes.size()) {
            delete r;
            return SG_ERR_INVALID_ARG;
        }
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        if (!s->have_first) {
            s->have_first = true;
            s->first_submit = r->t_submit;
        }
        s->pending.push_back(r);
        s->live[r->id] = r;
        *id = r->id;
    }
    s->cv_work.notify_all();
    return SG_OK;
}

sg_status sg_server_wait(sg_server* s, uint64_t id, double* latency_us) {
    return sg_server_wait_version(s, id, latency_us, nullptr);
}

sg_status sg_server_wait_version(sg_server* s, uint64_t id, double* latency_us, uint64_t* version) {
    if (!s) return SG_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->live.find(id);
    if (it == s->live.end()) return SG_ERR_INVALID_ARG;
    Request* r = it->second;
    s->cv_done.wait(lk, [&] { return r->done; });
    if (latency_us) *latency_us = std::chrono::duration<double, std::micro>(r->t_done - r->t_submit).count();
    if (version) *version = r->version;
    const sg_status st = r->status;
    s->live.erase(it);
    delete r;
    return st;
}

sg_status sg_server_drain(sg_server* s) {
    if (!s) return SG_ERR_INVALID_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] {
        for (auto& kv : s->live)
            if (!kv.second->done) return false;
        return true;
    });
    sg_status st = SG_OK;
    for (auto& kv : s->live) {
        if (kv.second->status != SG_OK) st = kv.second->status;
        delete kv.second;
    }
    s->live.clear();
    return st;
}

sg_status sg_server_get_stats(sg_server* s, sg_server_stats* out) {
    if (!s || !out) return SG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof(*out));
    std::vector<double> v = s->lat_us;
    out->requests = v.size();
    out->rays = s->st_rays;
    out->waves = s->st_waves;
    if (!v.empty()) {
        std::sort(v.begin(), v.end());
        auto pct = [&](double q) {
            const size_t k = (size_t)std::min<double>(v.size() - 1, std::max(0.0, q * (double)v.size() - 1e-9));
            return v[k];
        };
        out->p50_us = pct(0.50);
        out->p99_us = pct(0.99);
        double sum = 0;
        for (double x : v) sum += x;
        out->mean_us = sum / v.size();
        out->max_us = v.back();
        out->wall_s = std::chrono::duration<double>(s->last_done - s->first_submit).count();
        out->mrays_per_s = out->wall_s > 0 ? (double)s->st_rays / out->wall_s / 1e6 : 0.0;
    }
    out->regrown = s->regrown;
    return SG_OK;
}

sg_status sg_server_reset_stats(sg_server* s) {
    if (!s) return SG_ERR_INVALID_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_us.clear();
    s->st_rays = 0;
    s->st_waves = 0;
    s->have_first = false;
    return SG_OK;
}

}  // extern "C"
