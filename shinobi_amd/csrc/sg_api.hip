// sg_api.hip -- the C ABI (include/shinobi_grid.h): contexts, scenes, grid
// builds, traversal and device plumbing on top of the gfx950 kernels.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <vector>

#include "sg_internal.hpp"

static int bit_length(uint64_t x) {
    int b = 0;
    while (x) {
        ++b;
        x >>= 1;
    }
    return b < 1 ? 1 : b;
}

// Scratch of sg_bounce_noalloc: a fixed device array, sized for
// SG_BOUNCE_NOALLOC_MAX_RAYS input rays (the per-tile class counts and class
// totals of the bounce ordering; every word is written before it is read, so
// nothing needs clearing between calls).  Each GPU holds its own copy at its
// own address until the process ends.  In host code `scratch` names a
// host-side stand-in, not device memory: its device address on each GPU is
// looked up once by sg_ctx_create (noalloc_init), and only that address is
// ever passed to kernels.
static constexpr size_t kNoAllocScratchBytes = 16777728;   // >= bounce_scratch_bytes(MAX), multiple of 256
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_noalloc_mu;
static void* g_noalloc_ptr[kMaxDevices];   // device address of `scratch`, per device

// with `device` current: find the array's address on it (once per device)
static sg_status noalloc_init(int device) {
    if (device < 0 || device >= kMaxDevices) return SG_OK;   // sg_bounce_noalloc then reports SG_ERR_HIP
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if (g_noalloc_ptr[device]) return SG_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return e != hipSuccess ? hip_status(e) : SG_ERR_HIP;
    g_noalloc_ptr[device] = p;
    return SG_OK;
}

static void* noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    return g_noalloc_ptr[device];
}

extern "C" {

const char* sg_status_str(sg_status s) {
    switch (s) {
        case SG_OK: return "ok";
        case SG_ERR_INVALID_ARG: return "invalid argument";
        case SG_ERR_INVALID_SCENE: return "invalid scene (empty, zero-extent bounds or resolution < 1)";
        case SG_ERR_OUT_OF_MEMORY: return "out of device memory";
        case SG_ERR_HIP: return "HIP runtime error";
        case SG_ERR_CAPACITY: return "rebuild exceeded the grid's buffer capacity";
        case SG_ERR_LEAF_RES: return "leaf resolution outside [1,255]";
        case SG_ERR_TOO_MANY_CELLS: return "too many cells (>= 2^24)";
        case SG_ERR_NEG_EXTENT: return "negative leaf extent";
        case SG_ERR_TIMEOUT: return "device-side wait exceeded its bound";
        case SG_ERR_BUDGET: return "request exceeds the HBM budget";
        case SG_ERR_NO_DEVICE: return "no HIP device";
        case SG_ERR_IO: return "file cannot be read";
        case SG_ERR_SCRATCH_LIMIT: return "input exceeds the fixed scratch of an allocation-free entry point";
        default: return "unknown status";
    }
}

void sg_build_params_default(sg_build_params* p, int kind) {
    memset(p, 0, sizeof(*p));
    p->kind = kind;
    p->lambda = 5.0f;
    p->top_divisor = kind == SG_GRID_TWO_LEVEL ? 6 : 1;
    p->lambda_leaf = 1.2f;
    p->leaf_min_refs = 16;
    p->radix_bits = 8;
    p->capacity_slack = 1.25f;
    p->trace_block = 64;
}

sg_status sg_device_count(int* out) {
    if (!out) return SG_ERR_INVALID_ARG;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    *out = e == hipSuccess ? n : 0;
    return SG_OK;
}

sg_status sg_ctx_create(int device, sg_ctx** out) {
    if (!out) return SG_ERR_INVALID_ARG;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return SG_ERR_NO_DEVICE;
    if (device < 0 || device >= n) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(device));
    const sg_status ns = noalloc_init(device);
    if (ns != SG_OK) return ns;
    sg_ctx* c = new sg_ctx();
    c->device = device;
    bool ok = hipEventCreateWithFlags(&c->fork_start, hipEventDisableTiming) == hipSuccess;
    for (int k = 0; ok && k < SG_REBUILD_BRANCHES_MAX; ++k)
        ok = hipStreamCreateWithFlags(&c->fork_st[k], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&c->fork_join[k], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        sg_ctx_destroy(c);
        return SG_ERR_HIP;
    }
    *out = c;
    return SG_OK;
}

sg_status sg_ctx_destroy(sg_ctx* ctx) {
    if (!ctx) return SG_ERR_INVALID_ARG;
    hipSetDevice(ctx->device);
    if (!ctx->scratch.empty() || !ctx->retired.empty()) hipDeviceSynchronize();
    for (auto& kv : ctx->scratch) hipFree(kv.second.p);
    for (void* p : ctx->retired) hipFree(p);
    for (int k = 0; k < SG_REBUILD_BRANCHES_MAX; ++k) {
        if (ctx->fork_st[k]) hipStreamDestroy(ctx->fork_st[k]);
        if (ctx->fork_join[k]) hipEventDestroy(ctx->fork_join[k]);
    }
    if (ctx->fork_start) hipEventDestroy(ctx->fork_start);
    delete ctx;
    return SG_OK;
}

// The scratch buffer of `stream` with at least `bytes` bytes.  Growing it
// allocates, which a stream under graph capture must not do: such a call
// returns SG_ERR_INVALID_ARG (make the same call once before capturing).
static sg_status ctx_scratch(sg_ctx* ctx, hipStream_t stream, size_t bytes, void** out) {
    std::lock_guard<std::mutex> lk(ctx->scratch_mu);
    ScratchBuf& b = ctx->scratch[stream];
    if (b.bytes < bytes) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(stream, &cs));
        if (cs != hipStreamCaptureStatusNone) return SG_ERR_INVALID_ARG;
        const size_t nb = std::max(bytes, (size_t)1 << 20);
        void* p = nullptr;
        if (hipMalloc(&p, nb) != hipSuccess) return SG_ERR_OUT_OF_MEMORY;
        if (b.p) ctx->retired.push_back(b.p);
        b.p = p;
        b.bytes = nb;
    }
    *out = b.p;
    return SG_OK;
}

// ------------------------------------------------------------------ scenes
static TriSoA scene_soa(const sg_scene* s) {
    TriSoA t;
    for (int i = 0; i < 9; ++i) t.v[i] = s->soa + (size_t)i * s->stride;
    t.rec = s->rec;
    return t;
}

static sg_status scene_upload(sg_scene* s, const float* const soa9[9], int on_dev, hipStream_t st) {
    clear_last_error();
    const size_t b = (size_t)s->n * sizeof(float);
    for (int i = 0; i < 9; ++i) {
        if (!soa9[i]) return SG_ERR_INVALID_ARG;
        HIPCHK(hipMemcpyAsync(s->soa + (size_t)i * s->stride, soa9[i], b,
                              on_dev ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice, st));
    }
    launch_prep_tris(scene_soa(s), s->n, st);
    HIPCHK(hipGetLastError());
    if (!on_dev) HIPCHK(hipStreamSynchronize(st));
    return SG_OK;
}

sg_status sg_scene_create(sg_ctx* ctx, const float* const soa9[9], uint32_t n, int src_on_device, void* stream,
                          sg_scene** out) {
    if (!ctx || !soa9 || !out || n == 0) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    sg_scene* s = new sg_scene();
    s->ctx = ctx;
    s->device = ctx->device;
    s->n = n;
    s->stride = ((size_t)n + 63) & ~(size_t)63;
    if (hipMalloc(&s->soa, (size_t)9 * s->stride * sizeof(float)) != hipSuccess ||
        hipMalloc(&s->rec, (size_t)kTriRecWords * n * sizeof(float)) != hipSuccess) {
        if (s->soa) hipFree(s->soa);
        delete s;
        return SG_ERR_OUT_OF_MEMORY;
    }
    sg_status r = scene_upload(s, soa9, src_on_device, (hipStream_t)stream);
    if (r != SG_OK) {
        hipFree(s->soa);
        hipFree(s->rec);
        delete s;
        return r;
    }
    *out = s;
    return SG_OK;
}

sg_status sg_scene_update(sg_scene* s, const float* const soa9[9], int src_on_device, void* stream) {
    if (!s || !soa9) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(s->ctx->device));
    return scene_upload(s, soa9, src_on_device, (hipStream_t)stream);
}

sg_status sg_scene_destroy(sg_scene* s) {
    if (!s) return SG_ERR_INVALID_ARG;
    hipSetDevice(s->device);
    hipFree(s->soa);
    hipFree(s->rec);
    delete s;
    return SG_OK;
}

uint32_t sg_scene_size(const sg_scene* s) { return s ? s->n : 0u; }

sg_status sg_scene_device_soa(const sg_scene* s, const float* out9[9]) {
    if (!s || !out9) return SG_ERR_INVALID_ARG;
    for (int i = 0; i < 9; ++i) out9[i] = s->soa + (size_t)i * s->stride;
    return SG_OK;
}

// ------------------------------------------------------------------- grids
static size_t align_up(size_t x, size_t a) { return (x + a - 1) / a * a; }

// lay out (or measure, pool == nullptr) every array of the grid
static size_t grid_layout(sg_grid* g, char* pool) {
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        off = align_up(off, 256);
        char* p = pool ? pool + off : nullptr;
        off += align_up(bytes, 16);
        return p;
    };
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    const uint32_t N = g->n_tris;
    g->hdr = (GridHeader*)take(sizeof(GridHeader));
    for (int b = 0; b < 2; ++b) {
        g->k1[b] = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
        g->v1[b] = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
    }
    g->ranges1 = (uint32_t*)take((size_t)g->cap_C1 * 8 + 16);
    g->cells1 = two ? (uint32_t*)take((size_t)g->cap_C1 * 8 + 16) : nullptr;
    if (two) {
        for (int b = 0; b < 2; ++b) {
            g->k2[b] = (uint32_t*)take((size_t)g->cap_P2 * 4 + 16);
            g->v2[b] = (uint32_t*)take((size_t)g->cap_P2 * 4 + 16);
        }
        g->leaves = (uint32_t*)take((size_t)g->cap_C2 * 8 + 16);
    }
    g->refs = (float*)take((size_t)(two ? g->cap_P2 : g->cap_P1) * 36 + 16);
    if (two && g->large_path) {
        g->large.pacc = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
        g->large.prej = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
        g->large.sacc = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
        g->large.srej = (uint32_t*)take((size_t)g->cap_P1 * 4 + 16);
        g->large.scan_scratch = take(scan_scratch_bytes(g->cap_P1));
    }
    if (g->p.keep_debug) {
        g->dbg_tri = (uint32_t*)take((size_t)N * 4 + 16);
        g->dbg_leafc = two ? (uint32_t*)take((size_t)g->cap_C1 * 4 + 16) : nullptr;
        g->dbg_pair = two ? (uint32_t*)take((size_t)g->cap_P1 * 4 + 16) : nullptr;
    } else {
        g->dbg_tri = g->dbg_leafc = g->dbg_pair = nullptr;
    }
    g->bpart = (uint32_t*)take(bounds_blocks(N) * 6 * 4 + 16);
    g->sums_top = (uint32_t*)take(emit_tiles(N) * 4 + 16);
    g->sums_cells = two ? (uint32_t*)take(cell_tiles(g->cap_C1) * 4 + 16) : nullptr;
    g->sums_leaf = two ? (uint32_t*)take(leaf_sums_words(g->cap_P1) * 4) : nullptr;
    g->leaf_info = two ? (uint32_t*)take(leaf_info_words(g->cap_P1) * 4) : nullptr;
    g->key_bits1 = bit_length(g->cap_C1);
    g->passes1 = sort_passes(g->key_bits1, g->rb);
    g->sort1_words = sort_scratch_words(g->cap_P1, g->key_bits1, g->rb);
    g->sort1 = (uint32_t*)take(g->sort1_words * 4);
    if (two) {
        g->key_bits2 = bit_length(g->cap_C2);
        g->passes2 = sort_passes(g->key_bits2, g->rb);
        g->sort2_words = sort_scratch_words(g->cap_P2, g->key_bits2, g->rb);
        g->sort2 = (uint32_t*)take(g->sort2_words * 4);
    }
    return align_up(off, 256);
}

static sg_status grid_alloc(sg_grid* g) {
    const size_t bytes = grid_layout(g, nullptr);
    if (g->pool && g->pool_bytes >= bytes) {
        grid_layout(g, g->pool);
        return SG_OK;
    }
    if (g->pool) {
        HIPCHK(hipStreamSynchronize(g->stream));
        hipFree(g->pool);
        g->pool = nullptr;
        g->pool_bytes = 0;
    }
    char* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) return SG_ERR_OUT_OF_MEMORY;
    g->pool = p;
    g->pool_bytes = bytes;
    grid_layout(g, g->pool);
    return SG_OK;
}

// Enqueue the whole build on `st`.  marks (optional, n_marks entries) are
// hipEvent_t recorded at the phase boundaries: [0] start, [1] after
// init+bounds, [2] top emission, [3] top sort, [4] top ranges, then for the
// two-level grid [5] leaf cells, [6] leaf emission, [7] leaf sort, [8] leaf
// ranges.
// the build arguments of a grid's (re)build
static BuildArgs build_args(const sg_grid* g, bool fixed_res) {
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    BuildArgs a;
    a.lambda = g->p.lambda;
    a.top_divisor = two ? g->p.top_divisor : 1;
    a.lambda_leaf = g->p.lambda_leaf;
    a.leaf_min_refs = g->p.leaf_min_refs;
    for (int i = 0; i < 3; ++i) a.fixed_res[i] = fixed_res ? g->res[i] : 0;
    a.cap_C1 = g->cap_C1;
    a.cap_P1 = g->cap_P1;
    a.cap_C2 = g->cap_C2;
    a.cap_P2 = g->cap_P2;
    a.keep_debug = g->p.keep_debug ? 1u : 0u;
    return a;
}

// what the first launch of a build (k_init_bounds, k_multi_init) zeroes
static ZeroSegs zero_segs(const sg_grid* g) {
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    ZeroSegs z;
    z.count = 0;
    auto seg = [&](void* p, size_t bytes) {
        if (p && bytes) {
            z.ptr[z.count] = p;
            z.bytes[z.count] = align_up(bytes, 16);
            ++z.count;
        }
    };
    seg(g->ranges1, (size_t)g->cap_C1 * 8);
    if (two) {
        seg(g->leaves, (size_t)g->cap_C2 * 8);
        seg(g->sums_leaf, leaf_sums_words(g->cap_P1) * 4);
        seg(g->leaf_info, leaf_info_words(g->cap_P1) * 4);
    }
    // per-pair leaf counts are only written below P1_valid
    seg(g->dbg_pair, (size_t)g->cap_P1 * 4);
    return z;
}

}  // extern "C"

bool uni_job_of(sg_grid* g, UniJob* j) {
    if (!g || !j || g->kind != SG_GRID_UNIFORM || !g->scene || !g->pool) return false;
    const uint32_t N = g->n_tris;
    j->t = scene_soa(g->scene);
    j->n = N;
    j->h = g->hdr;
    j->a = build_args(g, true);
    j->z = zero_segs(g);
    j->part = g->bpart;
    j->nb_bounds = (uint32_t)bounds_blocks(N);
    j->sums = g->sums_top;
    j->dbg = g->dbg_tri;
    for (int b = 0; b < 2; ++b) {
        j->k[b] = g->k1[b];
        j->v[b] = g->v1[b];
    }
    j->sort = g->sort1;
    j->cap_P = g->cap_P1;
    j->T = (uint32_t)((g->cap_P1 + 4095u) / 4096u);
    j->passes = g->passes1;
    j->rb = g->rb;
    j->ranges = g->ranges1;
    j->refs = g->refs;
    return true;
}

extern "C" {

static void enqueue_build(sg_grid* g, bool fixed_res, hipStream_t st, void* const* marks = nullptr,
                          int n_marks = 0) {
    auto mark = [&](int i) {
        if (marks && i < n_marks && marks[i]) (void)hipEventRecord((hipEvent_t)marks[i], st);
    };
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    const TriSoA t = scene_soa(g->scene);
    const uint32_t N = g->n_tris;
    const BuildArgs a = build_args(g, fixed_res);
    const ZeroSegs z = zero_segs(g);
    mark(0);
    launch_init_bounds(g->hdr, N, z, t, g->bpart, st);
    launch_params(g->hdr, a, g->bpart, N, true, st);
    mark(1);
    launch_top_emit(t, N, g->hdr, a, g->k1[0], g->v1[0], g->dbg_tri, g->sums_top, st);
    mark(2);
    launch_sort_pairs(g->k1[0], g->v1[0], g->k1[1], g->v1[1], &g->hdr->P1, g->cap_P1, g->key_bits1, g->rb, g->sort1,
                      st);
    mark(3);
    const int f1 = g->passes1 & 1;
    launch_ranges(g->k1[f1], &g->hdr->P1, g->cap_P1, &g->hdr->C1, g->ranges1, &g->hdr->P1_valid, st);
    if (!two) launch_ref_records(g->v1[f1], &g->hdr->P1_valid, nullptr, g->cap_P1, g->scene->rec, N, g->refs, st);
    mark(4);
    if (!two) return;
    launch_leaf_cells(g->hdr, a, g->ranges1, g->cells1, g->dbg_leafc, g->sums_cells, st);
    mark(5);
    launch_leaf_emit(t, g->hdr, a, g->k1[f1], g->v1[f1], g->ranges1, g->cells1, g->k2[0], g->v2[0], g->k2[1],
                     g->v2[1], g->refs, g->dbg_pair, g->sums_leaf, g->leaf_info, g->large_path ? &g->large : nullptr,
                     g->leaves, st);
    mark(6);
    // Only a grid whose build met leaf tiles too large for the LDS sort has
    // the parallel emission of such tiles and needs the global sort after
    // it; any other grid sorts every tile inside the emission (an oversize
    // tile that a rebuild meets is sorted by its block), so its rebuilds
    // enqueue neither the sort passes nor the record pass behind them.
    if (g->large_path)
        launch_leaf_sort_fixup(g->k2[0], g->v2[0], g->k2[1], g->v2[1], g->hdr, g->cap_P2, g->key_bits2, g->rb,
                               g->sort2, st);
    mark(7);
    // without the global sort the emission wrote the leaf ranges itself
    if (g->large_path) {
        launch_ranges(g->k2[0], &g->hdr->P2, g->cap_P2, &g->hdr->C2, g->leaves, &g->hdr->P2_valid, st);
        launch_ref_records(g->v2[0], &g->hdr->P2_valid, &g->hdr->leaf_sort, g->cap_P2, g->scene->rec, N, g->refs,
                           st);
    }
    mark(8);
}

static sg_status read_header(sg_grid* g, GridHeader* h) {
    HIPCHK(hipStreamSynchronize(g->stream));
    HIPCHK(hipMemcpy(h, g->hdr, sizeof(GridHeader), hipMemcpyDeviceToHost));
    return SG_OK;
}

sg_status sg_build(sg_ctx* ctx, sg_scene* scene, const sg_build_params* p, void* stream, sg_grid** out) {
    if (!ctx || !scene || !p || !out) return SG_ERR_INVALID_ARG;
    if (p->kind != SG_GRID_UNIFORM && p->kind != SG_GRID_TWO_LEVEL) return SG_ERR_INVALID_ARG;
    if (p->radix_bits != 0 && p->radix_bits != 4 && p->radix_bits != 6 && p->radix_bits != 8)
        return SG_ERR_INVALID_ARG;
    if (p->trace_block != 0 && p->trace_block != 64 && p->trace_block != 128 && p->trace_block != 256)
        return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    sg_grid* g = new sg_grid();
    g->ctx = ctx;
    g->device = ctx->device;
    g->scene = scene;
    g->p = *p;
    g->kind = p->kind;
    g->n_tris = scene->n;
    g->rb = p->radix_bits ? p->radix_bits : 8;
    g->stream = st;
    const float slack = p->capacity_slack >= 1.0f ? p->capacity_slack : 1.25f;

    // 1. bounds and resolution (the one value the host needs to size the top level)
    {
        g->cap_C1 = 1;
        g->cap_P1 = 1;
        g->cap_C2 = 1;
        g->cap_P2 = 1;
        sg_status r = grid_alloc(g);
        if (r != SG_OK) {
            delete g;
            return r;
        }
        ZeroSegs z;
        z.count = 0;
        launch_init_bounds(g->hdr, g->n_tris, z, scene_soa(scene), g->bpart, st);
        launch_params(g->hdr, BuildArgs(), g->bpart, g->n_tris, false, st);
        GridHeader h;
        r = read_header(g, &h);
        if (r != SG_OK) {
            sg_grid_destroy(g);
            return r;
        }
        float mn[3], mx[3], cs[3], rc[3], cr[3];
        for (int i = 0; i < 3; ++i) {
            mn[i] = fkey_inv(h.bmin_key[i]);
            mx[i] = fkey_inv(h.bmax_key[i]);
        }
        const int two = p->kind == SG_GRID_TWO_LEVEL;
        if (!derive_params(mn, mx, g->n_tris, p->lambda, two ? p->top_divisor : 1, p->res_override, g->res, cs, rc,
                           cr)) {
            sg_grid_destroy(g);
            return SG_ERR_INVALID_SCENE;
        }
        const uint64_t C1 = (uint64_t)g->res[0] * g->res[1] * g->res[2];
        if (C1 >= (1ull << 24)) {
            sg_grid_destroy(g);
            return SG_ERR_TOO_MANY_CELLS;
        }
        g->cap_C1 = (uint32_t)C1;
        const uint64_t N = g->n_tris;
        g->cap_P1 = (uint32_t)std::min<uint64_t>(N * 2 + 4096, 0x3FFFFFFFull);
        g->cap_C2 = two ? (uint32_t)std::min<uint64_t>(C1 + N * 2 + 4096, 0x3FFFFFFFull) : 1;
        g->cap_P2 = two ? (uint32_t)std::min<uint64_t>(N * 4 + 4096, 0x3FFFFFFFull) : 1;
    }
    // 2. full build; grow the buffers to the measured sizes until it fits
    for (int attempt = 0; attempt < 6; ++attempt) {
        sg_status r = grid_alloc(g);
        if (r != SG_OK) {
            sg_grid_destroy(g);
            return r;
        }
        enqueue_build(g, true, st);
        HIPCHK(hipGetLastError());
        GridHeader h;
        r = read_header(g, &h);
        if (r != SG_OK) {
            sg_grid_destroy(g);
            return r;
        }
        auto grow = [&](uint32_t need) {
            return (uint32_t)std::min<uint64_t>((uint64_t)(need * (double)slack) + 4096, 0x3FFFFFFFull);
        };
        const uint32_t cap_bits = ST_CAP_P1 | ST_CAP_C1 | ST_CAP_C2 | ST_CAP_P2;
        if (!(h.status & cap_bits)) {
            // fits: keep it if every buffer has the requested headroom for
            // rebuilds of a moving scene, else grow once and build again
            const bool two = g->kind == SG_GRID_TWO_LEVEL;
            bool again = false;
            if ((double)g->cap_P1 < h.need_P1 * (double)slack) { g->cap_P1 = grow(h.need_P1); again = true; }
            if (two && (double)g->cap_C2 < h.need_C2 * (double)slack) { g->cap_C2 = grow(h.need_C2); again = true; }
            if (two && (double)g->cap_P2 < h.need_P2 * (double)slack) { g->cap_P2 = grow(h.need_P2); again = true; }
            // a tile too large for LDS: lay out the parallel emission and
            // build again with it (rebuilds then use it too)
            if (two && h.oversize > 0 && !g->large_path) { g->large_path = true; again = true; }
            if (!again) {
                *out = g;
                return SG_OK;
            }
            continue;
        }
        // the true total of an overflowing stage is in need_*; stages after
        // it ran on clamped data and are re-measured by the next attempt
        if (h.status & ST_CAP_P1) g->cap_P1 = grow(h.need_P1);
        else if (h.status & ST_CAP_C2) g->cap_C2 = grow(h.need_C2);
        else if (h.status & ST_CAP_P2) g->cap_P2 = grow(h.need_P2);
    }
    sg_grid_destroy(g);
    return SG_ERR_CAPACITY;
}

static sg_status rebuild_impl(sg_ctx* ctx, sg_scene* scene, sg_grid* g, void* stream, bool fixed,
                              void* const* marks = nullptr, int n_marks = 0) {
    if (!ctx || !scene || !g) return SG_ERR_INVALID_ARG;
    if (scene->n != g->n_tris) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    g->scene = scene;
    g->stream = (hipStream_t)stream;
    enqueue_build(g, fixed, g->stream, marks, n_marks);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_rebuild(sg_ctx* ctx, sg_scene* scene, sg_grid* g, void* stream) {
    return rebuild_impl(ctx, scene, g, stream, true);
}

sg_status sg_rebuild_free_res(sg_ctx* ctx, sg_scene* scene, sg_grid* g, void* stream) {
    return rebuild_impl(ctx, scene, g, stream, false);
}

// Fixed cost of one rebuild in triangles: a rebuild is about a dozen
// launches of a few microseconds each, which a scene of this size takes in
// work alone.  Weighs the branch assignment of sg_rebuild_many.
static constexpr uint64_t kRebuildLaunchTris = 50000;

}  // extern "C"

sg_status rebuild_many_with(sg_ctx* ctx, uint32_t n, sg_scene* const* scenes, sg_grid* const* grids,
                            uint32_t branches, hipStream_t st, const ForkTask* extra) {
    if (!ctx || branches > SG_REBUILD_BRANCHES_MAX) return SG_ERR_INVALID_ARG;
    if (n && (!scenes || !grids)) return SG_ERR_INVALID_ARG;
    for (uint32_t i = 0; i < n; ++i) {
        if (!scenes[i] || !grids[i] || scenes[i]->n != grids[i]->n_tris) return SG_ERR_INVALID_ARG;
        for (uint32_t j = 0; j < i; ++j)
            if (grids[j] == grids[i]) return SG_ERR_INVALID_ARG;
    }
    const uint32_t tasks = n + (extra ? 1u : 0u);
    if (tasks == 0) return SG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    const uint32_t K = std::min(branches ? branches : (uint32_t)SG_REBUILD_BRANCHES_MAX, tasks);
    if (K == 1) {
        if (extra) {
            const sg_status r = extra->run(extra->arg, st);
            if (r != SG_OK) return r;
        }
        for (uint32_t i = 0; i < n; ++i) {
            const sg_status r = rebuild_impl(ctx, scenes[i], grids[i], st, true);
            if (r != SG_OK) return r;
        }
        return SG_OK;
    }
    // largest first, each to the branch with the least work so far (the
    // extra task, when there is one, starts branch 0)
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return grids[a]->n_tris > grids[b]->n_tris; });
    std::vector<uint64_t> load(K, 0);
    if (extra) load[0] = extra->weight_tris + kRebuildLaunchTris;
    std::vector<uint32_t> branch(n);
    for (uint32_t i : order) {
        const uint32_t k = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        branch[i] = k;
        load[k] += grids[i]->n_tris + kRebuildLaunchTris;
    }
    std::lock_guard<std::mutex> lk(ctx->fork_mu);
    HIPCHK(hipEventRecord(ctx->fork_start, st));
    for (uint32_t k = 0; k < K; ++k) HIPCHK(hipStreamWaitEvent(ctx->fork_st[k], ctx->fork_start, 0));
    sg_status r = SG_OK;
    if (extra) r = extra->run(extra->arg, ctx->fork_st[0]);
    for (uint32_t i : order) {
        if (r == SG_OK) r = rebuild_impl(ctx, scenes[i], grids[i], ctx->fork_st[branch[i]], true);
        // the grid's stream (synchronised before its header is read) is the
        // caller's: the join below orders the rebuild before its later work,
        // and graph replays never run on the side streams
        grids[i]->stream = st;
    }
    // join every branch, also after an error, so that no branch is left
    // outside a capture
    for (uint32_t k = 0; k < K; ++k) {
        HIPCHK(hipEventRecord(ctx->fork_join[k], ctx->fork_st[k]));
        HIPCHK(hipStreamWaitEvent(st, ctx->fork_join[k], 0));
    }
    return r;
}

extern "C" {

sg_status sg_rebuild_many(sg_ctx* ctx, uint32_t n, sg_scene* const* scenes, sg_grid* const* grids, uint32_t branches,
                          void* stream) {
    return rebuild_many_with(ctx, n, scenes, grids, branches, (hipStream_t)stream, nullptr);
}

sg_status sg_rebuild_phased(sg_ctx* ctx, sg_scene* scene, sg_grid* g, void* stream, void* const* events,
                            int n_events) {
    if (!events || n_events < 0) return SG_ERR_INVALID_ARG;
    return rebuild_impl(ctx, scene, g, stream, true, events, n_events);
}

sg_status sg_grid_get_info(sg_grid* g, sg_grid_info* out) {
    if (!g || !out) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(g->ctx->device));
    GridHeader h;
    sg_status r = read_header(g, &h);
    if (r != SG_OK) return r;
    memset(out, 0, sizeof(*out));
    out->kind = g->kind;
    out->status = status_from_bits(h.status);
    for (int i = 0; i < 3; ++i) {
        out->res[i] = h.res[i];
        out->bmin[i] = h.bmin[i];
        out->bmax[i] = h.bmax[i];
        out->cell_size[i] = h.cs[i];
        out->cell_size_rcp[i] = h.cell_rcp[i];
        out->build_rcp[i] = h.rcp_build[i];
    }
    out->n_tris = g->n_tris;
    out->C1 = h.C1;
    out->P1 = h.P1;
    out->P1_valid = h.P1_valid;
    out->C2 = h.C2;
    out->P2 = h.P2;
    out->P2_valid = h.P2_valid;
    out->cap_C1 = g->cap_C1;
    out->cap_P1 = g->cap_P1;
    out->cap_C2 = g->cap_C2;
    out->cap_P2 = g->cap_P2;
    out->device_bytes = g->pool_bytes;
    return SG_OK;
}

sg_status sg_grid_get_device_view(sg_grid* g, sg_grid_device_view* out) {
    if (!g || !out) return SG_ERR_INVALID_ARG;
    sg_grid_info info;
    const sg_status r = sg_grid_get_info(g, &info);
    if (r != SG_OK) return r;
    const GridView v = grid_view(g);
    memset(out, 0, sizeof(*out));
    out->kind = g->kind;
    out->status = info.status;
    for (int i = 0; i < 3; ++i) {
        out->res[i] = info.res[i];
        out->bmin[i] = info.bmin[i];
        out->bmax[i] = info.bmax[i];
        out->cell_size[i] = info.cell_size[i];
        out->cell_size_rcp[i] = info.cell_size_rcp[i];
    }
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    out->cells = v.cells;
    out->leaves = two ? v.leaves : nullptr;
    out->n_cells = info.C1;
    out->n_leaves = two ? info.C2 : 0u;
    out->indices = v.indices;
    out->n_refs = two ? info.P2 : info.P1;
    const TriSoA t = scene_soa(g->scene);
    for (int i = 0; i < 9; ++i) out->tri[i] = t.v[i];
    out->n_tris = g->n_tris;
    return SG_OK;
}

sg_status sg_grid_download(sg_grid* g, int which, void* host, size_t bytes) {
    if (!g || !host) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(g->ctx->device));
    GridHeader h;
    sg_status r = read_header(g, &h);
    if (r != SG_OK) return r;
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    const int f1 = g->passes1 & 1, f2 = 0;   // leaf pairs always end in buffer 0
    const void* src = nullptr;
    size_t n = 0;
    switch (which) {
        case SG_ARRAY_TOP_KEYS: src = g->k1[f1]; n = (size_t)h.P1 * 4; break;
        case SG_ARRAY_TOP_VALS: src = g->v1[f1]; n = (size_t)h.P1 * 4; break;
        case SG_ARRAY_TOP_RANGES: src = g->ranges1; n = (size_t)h.C1 * 8; break;
        case SG_ARRAY_TOP_CELLS:
            if (!two) return SG_ERR_INVALID_ARG;
            src = g->cells1; n = (size_t)h.C1 * 8; break;
        case SG_ARRAY_LEAF_KEYS:
            // kept by keep_debug grids (and grids with the global leaf sort);
            // others write their leaf ranges without storing the keys
            if (!two || !(g->p.keep_debug || g->large_path)) return SG_ERR_INVALID_ARG;
            src = g->k2[f2]; n = (size_t)h.P2 * 4; break;
        case SG_ARRAY_LEAF_VALS:
            if (!two) return SG_ERR_INVALID_ARG;
            src = g->v2[f2]; n = (size_t)h.P2 * 4; break;
        case SG_ARRAY_LEAVES:
            if (!two) return SG_ERR_INVALID_ARG;
            src = g->leaves; n = (size_t)h.C2 * 8; break;
        case SG_ARRAY_INDICES:
            src = two ? g->v2[f2] : g->v1[f1];
            n = (size_t)(two ? h.P2 : h.P1) * 4;
            break;
        case SG_ARRAY_TRI_COUNTS: src = g->dbg_tri; n = (size_t)g->n_tris * 4; break;
        case SG_ARRAY_LEAF_COUNTS: src = g->dbg_leafc; n = (size_t)h.C1 * 4; break;
        case SG_ARRAY_PAIR_COUNTS: src = g->dbg_pair; n = (size_t)h.P1 * 4; break;
        default: return SG_ERR_INVALID_ARG;
    }
    if (!src) return SG_ERR_INVALID_ARG;
    if (bytes < n) return SG_ERR_INVALID_ARG;
    if (n) HIPCHK(hipMemcpy(host, src, n, hipMemcpyDeviceToHost));
    return SG_OK;
}

sg_status sg_grid_destroy(sg_grid* g) {
    if (!g) return SG_ERR_INVALID_ARG;
    hipSetDevice(g->device);
    if (g->pool) {
        hipStreamSynchronize(g->stream);
        hipFree(g->pool);
    }
    delete g;
    return SG_OK;
}

// ---------------------------------------------------------------- traversal
static sg_status trace_common(sg_ctx* ctx, sg_grid* g, const sg_rays* rays, uint32_t n, const uint32_t* d_count,
                              uint32_t* d_prim, float* d_t, uint32_t* d_stats, bool shadow, void* stream) {
    if (!ctx || !g || !rays || !d_prim || !d_t) return SG_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (!rays->o[i] || !rays->d[i]) return SG_ERR_INVALID_ARG;
    if (!rays->tmax) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    launch_trace(grid_view(g), ray_soa(rays), n, d_count, d_prim, d_t, d_stats, shadow, g->p.trace_block,
                 (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_trace_closest(sg_ctx* ctx, sg_grid* g, const sg_rays* rays, uint32_t n, const uint32_t* d_count,
                           uint32_t* d_prim, float* d_t, uint32_t* d_stats, void* stream) {
    return trace_common(ctx, g, rays, n, d_count, d_prim, d_t, d_stats, false, stream);
}

sg_status sg_trace_shadow(sg_ctx* ctx, sg_grid* g, const sg_rays* rays, uint32_t n, const uint32_t* d_count,
                          uint32_t* d_prim, float* d_t, uint32_t* d_stats, void* stream) {
    return trace_common(ctx, g, rays, n, d_count, d_prim, d_t, d_stats, true, stream);
}

sg_status sg_trace_closest_store(sg_ctx* ctx, sg_grid* g, const sg_rays* rays, uint32_t n, const uint32_t* d_count,
                                 float* d_raybuf, uint32_t* d_prim, float* d_t, void* stream) {
    if (!ctx || !g || !rays || !d_raybuf) return SG_ERR_INVALID_ARG;
    for (int i = 0; i < 3; ++i)
        if (!rays->o[i] || !rays->d[i]) return SG_ERR_INVALID_ARG;
    if (!rays->tmax) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    TraceOut o{};
    o.prim = d_prim;
    o.t = d_t;
    o.raybuf = d_raybuf;
    o.stride = n;
    launch_trace_gen(grid_view(g), ray_soa(rays), nullptr, n, d_count, o, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_trace_camera(sg_ctx* ctx, sg_grid* g, const sg_pinhole* cam, uint32_t tile_w, uint32_t tile_h,
                          uint32_t* d_prim, float* d_t, float* d_raybuf, const sg_rays_out* rays_out, void* stream) {
    if (!ctx || !g || !cam || cam->res_x == 0 || cam->res_y == 0) return SG_ERR_INVALID_ARG;
    if (tile_w == 0 || tile_h == 0) return SG_ERR_INVALID_ARG;
    const uint64_t n = (uint64_t)cam->res_x * cam->res_y;
    if (n >= (1ull << 31)) return SG_ERR_INVALID_ARG;
    if (rays_out) {
        for (int i = 0; i < 3; ++i)
            if (!rays_out->o[i] || !rays_out->d[i]) return SG_ERR_INVALID_ARG;
        if (!rays_out->tmax) return SG_ERR_INVALID_ARG;
    }
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    CamArgs c{};
    for (int k = 0; k < 3; ++k) {
        c.pos[k] = cam->position[k];
        c.tl[k] = cam->top_left[k];
        c.sx[k] = cam->step_x[k];
        c.sy[k] = cam->step_y[k];
    }
    c.res_x = cam->res_x;
    c.res_y = cam->res_y;
    c.tw = std::min(tile_w, cam->res_x);
    c.th = std::min(tile_h, cam->res_y);
    TraceOut o{};
    o.prim = d_prim;
    o.t = d_t;
    o.raybuf = d_raybuf;
    o.stride = (uint32_t)n;
    if (rays_out) {
        for (int i = 0; i < 3; ++i) {
            o.rays[i] = rays_out->o[i];
            o.rays[3 + i] = rays_out->d[i];
        }
        o.rays[6] = rays_out->tmax;
    }
    RaySoA none{};
    launch_trace_gen(grid_view(g), none, &c, (uint32_t)n, nullptr, o, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_bounce(sg_ctx* ctx, sg_scene* scene, const sg_rays* rays, const uint32_t* d_prim, const float* d_t,
                    uint32_t n, const uint32_t* d_n, uint64_t seed, const sg_rays_out* out, uint32_t* d_src,
                    uint32_t* d_count, void* stream) {
    if (!ctx || !scene || !rays || !d_prim || !d_t || !out || !d_count) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    void* buf = nullptr;
    sg_status r = ctx_scratch(ctx, (hipStream_t)stream, bounce_scratch_bytes(n), &buf);
    if (r != SG_OK) return r;
    RaySoAOut o;
    for (int i = 0; i < 3; ++i) {
        o.a[i] = out->o[i];
        o.a[3 + i] = out->d[i];
    }
    o.a[6] = out->tmax;
    launch_bounce(scene_soa(scene), ray_soa(rays), d_prim, d_t, n, d_n, seed, o, d_src, d_count, buf,
                  (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

sg_status sg_bounce_ordered(sg_ctx* ctx, sg_scene* scene, const sg_rays* rays, const uint32_t* d_prim,
                            const float* d_t, uint32_t n, const uint32_t* d_n, uint64_t seed,
                            const sg_bounce_order* order, const sg_rays_out* out, uint32_t* d_src, uint32_t* d_count,
                            void* stream) {
    if (!order) return SG_ERR_INVALID_ARG;
    if (order->order == SG_BOUNCE_CLASS_ROUNDS)
        return sg_bounce(ctx, scene, rays, d_prim, d_t, n, d_n, seed, out, d_src, d_count, stream);
    if (order->order != SG_BOUNCE_EXIT_FIRST && order->order != SG_BOUNCE_AXIS_FIRST) return SG_ERR_INVALID_ARG;
    if (!ctx || !scene || !rays || !d_prim || !d_t || !out || !d_count) return SG_ERR_INVALID_ARG;
    if (n >= (1u << 30)) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    void* buf = nullptr;
    sg_status r = ctx_scratch(ctx, (hipStream_t)stream, bounce_ordered_scratch_bytes(n), &buf);
    if (r != SG_OK) return r;
    BounceOrder q{};
    q.mode = order->order == SG_BOUNCE_EXIT_FIRST ? BOUNCE_EXIT_FIRST : BOUNCE_AXIS_FIRST;
    for (int k = 0; k < 3; ++k) {
        q.bmin[k] = order->box_min[k];
        q.bmax[k] = order->box_max[k];
        q.axis[k] = order->axis[k];
    }
    q.scale = bounce_exit_scale(q.bmin, q.bmax);
    RaySoAOut o;
    for (int i = 0; i < 3; ++i) {
        o.a[i] = out->o[i];
        o.a[3 + i] = out->d[i];
    }
    o.a[6] = out->tmax;
    launch_bounce_ordered(scene_soa(scene), ray_soa(rays), d_prim, d_t, n, d_n, seed, q, o, d_src, d_count, buf,
                          (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

// sg_bounce with the library's fixed per-device scratch array: no allocation
// of any kind, so it may be called under graph capture at any size up to
// the limit
sg_status sg_bounce_noalloc(sg_ctx* ctx, sg_scene* scene, const sg_rays* rays, const uint32_t* d_prim,
                            const float* d_t, uint32_t n, const uint32_t* d_n, uint64_t seed, const sg_rays_out* out,
                            uint32_t* d_src, uint32_t* d_count, void* stream) {
    if (!ctx || !scene || !rays || !d_prim || !d_t || !out || !d_count) return SG_ERR_INVALID_ARG;
    if (n > SG_BOUNCE_NOALLOC_MAX_RAYS || bounce_scratch_bytes(n) > kNoAllocScratchBytes)
        return SG_ERR_SCRATCH_LIMIT;
    HIPCHK(hipSetDevice(ctx->device));
    void* dev_scratch = noalloc_scratch(ctx->device);
    if (!dev_scratch) return SG_ERR_HIP;
    clear_last_error();
    RaySoAOut o;
    for (int i = 0; i < 3; ++i) {
        o.a[i] = out->o[i];
        o.a[3 + i] = out->d[i];
    }
    o.a[6] = out->tmax;
    launch_bounce(scene_soa(scene), ray_soa(rays), d_prim, d_t, n, d_n, seed, o, d_src, d_count, dev_scratch,
                  (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

// --------------------------------------------------------------- primitives
sg_status sg_scan_u32(sg_ctx* ctx, const uint32_t* d_in, uint32_t* d_out, uint32_t n, int inclusive, void* stream) {
    if (!ctx || (!d_in && n) || (!d_out && n)) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    void* buf = nullptr;
    sg_status r = ctx_scratch(ctx, (hipStream_t)stream, scan_scratch_bytes(n), &buf);
    if (r != SG_OK) return r;
    launch_scan(d_in, d_out, n, inclusive != 0, buf, (hipStream_t)stream);
    HIPCHK(hipGetLastError());
    return SG_OK;
}

__global__ void k_sort_setup(uint32_t* hdr_words, uint32_t n) {
    if (threadIdx.x < 64) hdr_words[threadIdx.x] = threadIdx.x == 0 ? n : 0u;
}

sg_status sg_sort_pairs(sg_ctx* ctx, uint32_t* d_keys, uint32_t* d_vals, uint32_t n, int key_bits, int radix_bits,
                        void* stream) {
    if (!ctx || !d_keys || !d_vals || key_bits < 1 || key_bits > 32) return SG_ERR_INVALID_ARG;
    if (radix_bits == 0) radix_bits = 8;
    if (radix_bits != 4 && radix_bits != 6 && radix_bits != 8) return SG_ERR_INVALID_ARG;
    if (n >= (1u << 30)) return SG_ERR_INVALID_ARG;
    if (n < 2) return SG_OK;
    HIPCHK(hipSetDevice(ctx->device));
    clear_last_error();
    hipStream_t st = (hipStream_t)stream;
    // scratch: [64 words: n][keys n][vals n][sort scratch]
    const size_t sw = sort_scratch_words(n, key_bits, radix_bits);
    const size_t bytes = 256 + align_up((size_t)n * 4, 256) * 2 + sw * 4;
    void* buf = nullptr;
    sg_status r = ctx_scratch(ctx, st, bytes, &buf);
    if (r != SG_OK) return r;
    uint32_t* words = (uint32_t*)buf;
    uint32_t* tk = (uint32_t*)((char*)buf + 256);
    uint32_t* tv = (uint32_t*)((char*)tk + align_up((size_t)n * 4, 256));
    uint32_t* ss = (uint32_t*)((char*)tv + align_up((size_t)n * 4, 256));
    hipLaunchKernelGGL(k_sort_setup, dim3(1), dim3(64), 0, st, words, n);
    launch_sort_pairs(d_keys, d_vals, tk, tv, words, n, key_bits, radix_bits, ss, st);
    if (sort_passes(key_bits, radix_bits) & 1) {
        HIPCHK(hipMemcpyAsync(d_keys, tk, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(hipMemcpyAsync(d_vals, tv, (size_t)n * 4, hipMemcpyDeviceToDevice, st));
    }
    HIPCHK(hipGetLastError());
    return SG_OK;
}

// ---------------------------------------------------------------- plumbing
sg_status sg_malloc(sg_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (hipMalloc(out, bytes ? bytes : 1) != hipSuccess) return SG_ERR_OUT_OF_MEMORY;
    return SG_OK;
}

sg_status sg_free(sg_ctx* ctx, void* p) {
    if (!ctx) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (p) HIPCHK(hipFree(p));
    return SG_OK;
}

sg_status sg_host_alloc(sg_ctx* ctx, size_t bytes, void** out) {
    if (!ctx || !out) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    if (hipHostMalloc(out, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) return SG_ERR_OUT_OF_MEMORY;
    return SG_OK;
}

sg_status sg_host_free(sg_ctx* ctx, void* p) {
    if (!ctx) return SG_ERR_INVALID_ARG;
    if (p) HIPCHK(hipHostFree(p));
    return SG_OK;
}

sg_status sg_memcpy_htod(sg_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return SG_OK;
}

sg_status sg_memcpy_dtoh(sg_ctx* ctx, void* dst, const void* src, size_t bytes, void* stream) {
    if (!ctx || (!dst && bytes) || (!src && bytes)) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return SG_OK;
}

sg_status sg_stream_create(sg_ctx* ctx, void** out) {
    if (!ctx || !out) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t s;
    HIPCHK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *out = s;
    return SG_OK;
}

sg_status sg_stream_destroy(sg_ctx* ctx, void* s) {
    if (!ctx || !s) return SG_ERR_INVALID_ARG;
    HIPCHK(hipStreamDestroy((hipStream_t)s));
    return SG_OK;
}

sg_status sg_stream_sync(sg_ctx* ctx, void* s) {
    if (!ctx) return SG_ERR_INVALID_ARG;
    HIPCHK(hipStreamSynchronize((hipStream_t)s));
    return SG_OK;
}

sg_status sg_event_create(sg_ctx* ctx, void** out) {
    if (!ctx || !out) return SG_ERR_INVALID_ARG;
    hipEvent_t e;
    HIPCHK(hipEventCreate(&e));
    *out = e;
    return SG_OK;
}

sg_status sg_event_destroy(sg_ctx* ctx, void* ev) {
    if (!ctx || !ev) return SG_ERR_INVALID_ARG;
    HIPCHK(hipEventDestroy((hipEvent_t)ev));
    return SG_OK;
}

sg_status sg_event_record(sg_ctx* ctx, void* ev, void* stream) {
    if (!ctx || !ev) return SG_ERR_INVALID_ARG;
    HIPCHK(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return SG_OK;
}

sg_status sg_stream_wait_event(sg_ctx* ctx, void* stream, void* ev) {
    if (!ctx || !ev) return SG_ERR_INVALID_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamWaitEvent((hipStream_t)stream, (hipEvent_t)ev, 0));
    return SG_OK;
}

sg_status sg_event_elapsed_ms(sg_ctx* ctx, void* a, void* b, float* ms) {
    if (!ctx || !a || !b || !ms) return SG_ERR_INVALID_ARG;
    HIPCHK(hipEventSynchronize((hipEvent_t)b));
    HIPCHK(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return SG_OK;
}

}  // extern "C"
