// sg_internal.hpp -- library-private definitions shared by the C ABI
// translation units (sg_api.hip, sg_serve.hip): the opaque handle structs
// and the device views the kernels take.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <vector>

#include "../../include/shinobi_grid.h"
#include "sg_kernels.hpp"

using namespace sg;

// Scratch of the standalone primitives (sg_scan_u32, sg_sort_pairs,
// sg_bounce): one buffer per stream, so calls on different streams never
// share device memory; calls on one stream are ordered by the stream.  A
// buffer that a larger call outgrows is retired, not freed, until the
// context is destroyed: work already enqueued, or captured into a graph,
// may still address it.
struct ScratchBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct sg_ctx {
    int device;
    std::mutex scratch_mu;
    std::map<hipStream_t, ScratchBuf> scratch;   // per stream (nullptr = the null stream)
    std::vector<void*> retired;
    // side streams and events of sg_rebuild_many (created with the context,
    // so that a call under graph capture creates nothing)
    std::mutex fork_mu;
    hipStream_t fork_st[SG_REBUILD_BRANCHES_MAX] = {};
    hipEvent_t fork_join[SG_REBUILD_BRANCHES_MAX] = {};
    hipEvent_t fork_start = nullptr;
};

struct sg_scene {
    sg_ctx* ctx;
    int device;               // the context's device, kept so destroy never reads the context
    uint32_t n;
    size_t stride = 0;        // floats between the SoA arrays (multiple of 64: 256-byte aligned arrays)
    float* soa = nullptr;     // 9*stride floats
    float* rec = nullptr;     // per-triangle records, kTriRecWords floats each
};

struct sg_grid {
    sg_ctx* ctx;
    int device;               // as sg_scene::device
    sg_scene* scene;
    sg_build_params p;
    int kind;
    uint32_t n_tris;
    int res[3];
    // device memory: one allocation carved into the arrays below
    char* pool = nullptr;
    size_t pool_bytes = 0;
    GridHeader* hdr = nullptr;
    uint32_t cap_C1 = 0, cap_P1 = 0, cap_C2 = 0, cap_P2 = 0;
    uint32_t *k1[2] = {nullptr, nullptr}, *v1[2] = {nullptr, nullptr};
    uint32_t *ranges1 = nullptr, *cells1 = nullptr;
    uint32_t *k2[2] = {nullptr, nullptr}, *v2[2] = {nullptr, nullptr};
    uint32_t* leaves = nullptr;
    float* refs = nullptr;    // per-reference records (write_ref_record)
    // parallel emission of oversize leaf tiles, laid out once a build of this
    // grid has met one (GridHeader::leaf_sort > 0)
    bool large_path = false;
    LeafLarge large{};
    uint32_t *dbg_tri = nullptr, *dbg_leafc = nullptr, *dbg_pair = nullptr;
    uint32_t *sums_top = nullptr, *sums_cells = nullptr, *sums_leaf = nullptr;   // per-tile sums
    uint32_t* leaf_info = nullptr;                                               // per leaf tile bounds
    uint32_t* bpart = nullptr;                                                   // bounds partials
    uint32_t *sort1 = nullptr, *sort2 = nullptr;
    size_t sort1_words = 0, sort2_words = 0;
    int key_bits1 = 1, key_bits2 = 1, passes1 = 1, passes2 = 1, rb = 8;
    hipStream_t stream = nullptr;
};

// ---------------------------------------------- shared by the C ABI units
// a grid's fixed-resolution rebuild as a job of the multi-grid launches
// (uniform grids only; false otherwise)
bool uni_job_of(sg_grid* g, UniJob* j);
// work enqueued on a stream by a caller of rebuild_many_with
struct ForkTask {
    sg_status (*run)(void* arg, hipStream_t st);
    void* arg;
    uint64_t weight_tris;   // its work, in triangles of rebuild
};
// sg_rebuild_many, with `extra` (optional) enqueued first on branch 0
sg_status rebuild_many_with(sg_ctx* ctx, uint32_t n, sg_scene* const* scenes, sg_grid* const* grids,
                            uint32_t branches, hipStream_t st, const ForkTask* extra);

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return hip_status(e_); \
    } while (0)

// hipGetLastError reports the last error of any earlier HIP call of this
// thread; an entry point that checks its own launches with it first drops
// what earlier, unrelated calls left behind
inline void clear_last_error() { (void)hipGetLastError(); }

// the status a caller sees for a grid's device status bits
inline sg_status status_from_bits(uint32_t b) {
    if (b & ST_INVALID) return SG_ERR_INVALID_SCENE;
    if (b & ST_TOO_MANY_CELLS) return SG_ERR_TOO_MANY_CELLS;
    if (b & (ST_CAP_P1 | ST_CAP_C1 | ST_CAP_C2 | ST_CAP_P2)) return SG_ERR_CAPACITY;
    if (b & ST_NEG_EXTENT) return SG_ERR_NEG_EXTENT;
    if (b & ST_LEAF_RES) return SG_ERR_LEAF_RES;
    return SG_OK;
}

// status bits with which the traversal treats every ray as a miss (a build
// that failed or outgrew its buffers); grids with a leaf resolution outside
// [1,255] or a negative leaf extent are traversed as the reference would
inline bool untraversable(uint32_t b) { return (b & ~(ST_LEAF_RES | ST_NEG_EXTENT)) != 0; }

inline sg_status hip_status(hipError_t e) {
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation) return SG_ERR_OUT_OF_MEMORY;
    if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return SG_ERR_NO_DEVICE;
    return SG_ERR_HIP;
}


// device view of a built grid (the buffers that hold its final arrays)
inline GridView grid_view(const sg_grid* g) {
    GridView v;
    const bool two = g->kind == SG_GRID_TWO_LEVEL;
    const int f1 = g->passes1 & 1, f2 = 0;   // leaf pairs always end in buffer 0
    v.hdr = g->hdr;
    v.cells = two ? g->cells1 : g->ranges1;
    v.leaves = two ? g->leaves : nullptr;
    v.indices = two ? g->v2[f2] : g->v1[f1];
    v.rec = g->scene->rec;
    v.refs = g->refs;
    v.kind = g->kind;
    return v;
}

inline RaySoA ray_soa(const sg_rays* r) {
    RaySoA s;
    for (int i = 0; i < 3; ++i) {
        s.a[i] = r->o[i];
        s.a[3 + i] = r->d[i];
    }
    s.a[6] = r->tmax;
    return s;
}

