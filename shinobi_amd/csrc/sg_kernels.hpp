// sg_kernels.hpp -- host-side launch interface of the gfx950 kernels.
#pragma once

#include "sg_common.hpp"

namespace sg {

// Unaligned 16-byte vector: packed records are read and written as two of
// these at 4-byte alignment plus one word.
typedef float f4u __attribute__((ext_vector_type(4), aligned(4)));

// nine SoA arrays: v0.xyz, v1.xyz, v2.xyz (device pointers), plus the
// per-triangle records that k_prep_tris writes whenever the scene's vertices
// are set (sg_scene_create / sg_scene_update): {v0, v1, v2}, nine floats
// packed at a 36-byte stride (kTriRecWords), so the gathers of the leaf level
// of the build and of the reference-record writes read one record instead of
// nine SoA arrays.  Edges are not stored: e1 = v1 - v0 and e2 = v2 - v0 are
// recomputed where needed, with the same binary32 subtraction, so every
// result is what stored edges gave.  (A 64-byte record with the edges moved
// 1.8x the bytes per gather.)
constexpr uint32_t kTriRecWords = 9;

struct TriSoA {
    const float* v[9];
    float* rec;
};

SG_D void store_tri_record(float* __restrict__ rec, uint32_t i, const v3& v0, const v3& v1, const v3& v2) {
    float* r = rec + kTriRecWords * (size_t)i;
    *(f4u*)r = f4u{v0.x, v0.y, v0.z, v1.x};
    *(f4u*)(r + 4) = f4u{v1.y, v1.z, v2.x, v2.y};
    r[8] = v2.z;
}

SG_D void load_tri_record(const float* __restrict__ rec, uint32_t i, v3& v0, v3& v1, v3& v2) {
    const float* r = rec + kTriRecWords * (size_t)i;
    const f4u a = *(const f4u*)r, b = *(const f4u*)(r + 4);
    v0 = mk(a.x, a.y, a.z);
    v1 = mk(a.w, b.x, b.y);
    v2 = mk(b.z, b.w, r[8]);
}

// seven SoA arrays: ox, oy, oz, dx, dy, dz, tmax (device pointers)
struct RaySoA {
    const float* a[7];
};
struct RaySoAOut {
    float* a[7];
};

// byte ranges zeroed by the first launch of a build (k_init_bounds,
// k_multi_init; multiples of 16 bytes, 16-byte aligned)
struct ZeroSegs {
    void* ptr[8];
    size_t bytes[8];
    int count;
};

// device view of a built grid, passed by value to the traversal kernels
struct GridView {
    const GridHeader* hdr;
    const uint32_t* cells;     // uniform: 2*C ranges; two-level: 2*C1 packed cells
    const uint32_t* leaves;    // two-level: 2*C2 ranges
    const uint32_t* indices;   // reference -> triangle id
    const float* rec;          // per-triangle record (9 floats, see TriSoA)
    const float* refs;         // per-reference records (9 floats each, see write_ref_record)
    int kind;                  // 0 uniform, 1 two-level
};

// ---------------------------------------------------------------- build
// Reference records: for every reference i < *d_valid of the final grid,
// refs[9i..9i+8] = v0, e1 = v1 - v0, e2 = v2 - v0 of triangle indices[i] --
// the 36 bytes a triangle test reads, in the order the traversal reads them
// (no index indirection; the triangle id comes from indices[] once per ray).
// Records are packed at a 36-byte stride and read as two 16-byte loads at
// 4-byte alignment plus one word: 25 % fewer bytes than 48-byte aligned
// records, which measured faster on gfx950 (more of them stay in the
// memory-side cache).
SG_D void write_ref_record(float* __restrict__ refs, uint32_t i, const float* __restrict__ rec, uint32_t tri,
                           uint32_t n_tris) {
    v3 v0 = mk(0.f, 0.f, 0.f), e1 = v0, e2 = v0;
    if (tri < n_tris) {   // always, unless the build overflowed (never traversed then)
        v3 v1, v2;
        load_tri_record(rec, tri, v0, v1, v2);
        e1 = v1 - v0;
        e2 = v2 - v0;
    }
    float* o = refs + 9 * (size_t)i;
    *(f4u*)o = f4u{v0.x, v0.y, v0.z, e1.x};
    *(f4u*)(o + 4) = f4u{e1.y, e1.z, e2.x, e2.y};
    o[8] = e2.z;
}
void launch_ref_records(const uint32_t* indices, const uint32_t* d_valid, const uint32_t* d_gate, uint32_t cap,
                        const float* rec, uint32_t n_tris, float* refs, hipStream_t st);
// zero the segments, reset the header and leave the per-block bounds
// partials (6 words per block, bounds_blocks(n) blocks), in one launch
size_t bounds_blocks(uint32_t n);
void launch_init_bounds(GridHeader* h, uint32_t n_tris, const ZeroSegs& segs, const TriSoA& t, uint32_t* part,
                        hipStream_t st);
// fold the partials into the header; with derive, resolution + cell sizes
void launch_params(GridHeader* h, const BuildArgs& a, const uint32_t* part, uint32_t n, bool derive,
                   hipStream_t st);
// per-tile sums needed by the emission kernels: emit_tiles(n) for n
// triangles or top pairs, cell_tiles(C1) for the leaf-cell kernels
size_t emit_tiles(uint32_t n);
size_t cell_tiles(uint32_t n);
void launch_top_emit(const TriSoA& t, uint32_t n, GridHeader* h, const BuildArgs& a, uint32_t* keys,
                     uint32_t* vals, uint32_t* dbg, uint32_t* sums, hipStream_t st);
void launch_ranges(const uint32_t* keys, const uint32_t* d_n, uint32_t cap, const uint32_t* d_C,
                   uint32_t* cells, uint32_t* d_pvalid, hipStream_t st);
void launch_leaf_cells(GridHeader* h, const BuildArgs& a, const uint32_t* ranges, uint32_t* cells,
                       uint32_t* dbg, uint32_t* sums, hipStream_t st);
// leaf pairs, sorted by key within each leaf tile.  sums
// (leaf_sums_words(cap_P1)) and info (leaf_info_words(cap_P1)) must be zero
// on entry (the build's first launch zeroes them).
size_t leaf_tiles(uint32_t n);
size_t leaf_sums_words(uint32_t n);
size_t leaf_info_words(uint32_t n);
// Buffers of the parallel emission of tiles too large for LDS (per top pair:
// accepted / rejected leaf-pair totals and their exclusive scans).  Without
// them such a tile is emitted by one block.
struct LeafLarge {
    uint32_t *pacc, *prej, *sacc, *srej;
    void* scan_scratch;   // scan_scratch_bytes(cap_P1)
};
// refs: reference records of the accepted pairs of sorted tiles.  Without
// `large`, a tile too large for LDS is emitted and sorted by one block
// (using tmp_keys / tmp_vals, buffers of the pairs' size) and no global leaf
// sort is needed; with `large`, such tiles are emitted in parallel and left
// to the global sort (launch_leaf_sort_fixup + launch_ref_records).
// Without `large` the tiles also write the leaf ranges (`leaves`, zeroed by
// the build's first launch) and the leaf keys only with a.keep_debug; with
// it, every key is
// written for the global sort and k_ranges.
void launch_leaf_emit(const TriSoA& t, GridHeader* h, const BuildArgs& a, const uint32_t* top_keys,
                      const uint32_t* top_vals, const uint32_t* ranges, const uint32_t* cells, uint32_t* keys,
                      uint32_t* vals, uint32_t* tmp_keys, uint32_t* tmp_vals, float* refs, uint32_t* dbg,
                      uint32_t* sums, uint32_t* info, const LeafLarge* large, uint32_t* leaves, hipStream_t st);
// global stable sort of the leaf pairs of tiles too large for LDS (no work
// when there are none); the result stays in (k0, v0)
void launch_leaf_sort_fixup(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, GridHeader* h, uint32_t cap,
                            int key_bits, int radix_bits, uint32_t* scratch, hipStream_t st);

// ---------------------------------------------------------------- sort / scan
// Scratch words the sort needs for `cap` pairs sorted over `key_bits` bits.
size_t sort_scratch_words(uint32_t cap, int key_bits, int radix_bits);
int sort_passes(int key_bits, int radix_bits);
// Stable LSD radix sort of (keys,vals) by the low key_bits bits.  n is read
// from device memory (*d_n, clamped to cap).  Input in (k0,v0); the result
// is in buffer index (sort_passes() & 1) of {(k0,v0),(k1,v1)}.  scratch holds
// sort_scratch_words() words (no initialisation needed).
void launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n,
                       uint32_t cap, int key_bits, int radix_bits, uint32_t* scratch, hipStream_t st);

// ------------------------------------------- many uniform grids per launch
// The rebuild of a set of uniform grids (fixed resolutions) with one launch
// per phase for the whole set: each launch covers the blocks every grid's
// own launch would have, numbered one grid after the other.  Entry i of a
// launch's table names a grid (job) and its first block b0; a block finds
// its entry by binary search and runs the single-grid kernel's body with
// its block index and count within the grid, so every grid computes exactly
// what its own launches compute.
struct UniJob {
    TriSoA t;
    uint32_t n;               // triangles
    GridHeader* h;
    BuildArgs a;              // as the grid's rebuild (fixed resolution)
    ZeroSegs z;               // what the grid's first build launch zeroes
    uint32_t* part;           // bounds partials
    uint32_t nb_bounds;       // bounds_blocks(n): blocks of the bounds pass
    uint32_t* sums;           // top emission tile sums
    uint32_t* dbg;            // per-triangle counts (keep_debug) or null
    uint32_t* k[2];           // sort buffers: pass p reads k[p & 1], writes k[(p + 1) & 1]
    uint32_t* v[2];
    uint32_t* sort;           // sort scratch (sort_scratch_words)
    uint32_t cap_P;           // pair capacity
    uint32_t T;               // sort tiles of cap_P
    int passes, rb;           // sort passes and radix width of the grid
    uint32_t* ranges;         // cell ranges
    float* refs;              // reference records
};
struct MultiEntry {
    uint32_t job, b0;
};
enum MultiPhase { MP_INIT, MP_PARAMS, MP_TOP_COUNT, MP_TOP_EMIT, MP_RANGES, MP_REFS, MP_COUNT };
// the job of block b of a launch over `total` blocks (the last entry with
// b0 <= b), with the block's index and the block count within the job
SG_D uint32_t multi_entry(const MultiEntry* __restrict__ e, uint32_t ne, uint32_t b, uint32_t total, uint32_t& bid,
                          uint32_t& nb) {
    uint32_t lo = 0, hi = ne;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (e[mid].b0 <= b) lo = mid;
        else hi = mid;
    }
    const uint32_t b0 = e[lo].b0, b1 = lo + 1 < ne ? e[lo + 1].b0 : total;
    bid = b - b0;
    nb = b1 - b0;
    return e[lo].job;
}
enum MultiSortStep { MS_UP, MS_SCAN, MS_DOWN };
// blocks a grid's own launch of the phase uses (0: no launch)
uint32_t multi_blocks(const UniJob& j, MultiPhase ph);
uint32_t multi_sort_blocks(const UniJob& j, MultiSortStep step);
void launch_multi(MultiPhase ph, const UniJob* jobs, const MultiEntry* e, uint32_t ne, uint32_t blocks,
                  hipStream_t st);
// pass `pass` of every listed job, all of radix width rb
void launch_multi_sort(MultiSortStep step, int rb, int pass, const UniJob* jobs, const MultiEntry* e, uint32_t ne,
                       uint32_t blocks, hipStream_t st);

// scan of n u32 values (reduce-then-scan); scratch: scan_scratch_bytes(n)
size_t scan_scratch_bytes(uint32_t n);
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, void* scratch,
                 hipStream_t st);

// ---------------------------------------------------------------- traversal
void launch_prep_tris(const TriSoA& t, uint32_t n, hipStream_t st);
// one (grid, ray batch) item of a multi-batch launch; chunk0 = first 64-ray
// chunk of the item in the concatenation of all items (items sorted by it)
struct TraceItem {
    GridView g;
    RaySoA rays;
    uint32_t* prim;
    float* t;
    uint32_t* status;   // optional: receives the grid's status bits (GridHeader::status)
    uint32_t n;
    uint32_t chunk0;
};
// A table of items split by grid kind: uniform items first, then two-level
// ones, each part with its own chunk numbering from 0.
struct ItemSplit {
    uint32_t n[2] = {0, 0};        // items of each kind
    uint32_t chunks[2] = {0, 0};   // 64-ray chunks of each kind
};
// lay out n host items in `out` as described by ItemSplit (chunk0 renumbered)
ItemSplit split_items(const TraceItem* in, uint32_t n, TraceItem* out);
// traversal of the items (device table laid out by split_items): one launch
// per kind present, one wave per 64-ray chunk, each kernel holding only its
// kind's walk (and so its registers); enqueues kernels only, so the
// launches can be captured in a graph and replayed
void launch_trace_items(const TraceItem* items, const ItemSplit& split, bool shadow, hipStream_t st);
// closest hit, or with shadow the any-hit traversal that stops at t < 0.9999
// (block: threads per block, 64, 128 or 256)
void launch_trace(const GridView& g, const RaySoA& r, uint32_t n, const uint32_t* d_n, uint32_t* prim,
                  float* t, uint32_t* stats, bool shadow, int block, hipStream_t st);

// Primary rays generated inside the trace kernel, as RegularPrimaryRayGenerator
// (src/Engine/RayGenerators.h:36-61) does per ray id: the pinhole Camera after
// init (src/Primitive/Camera.h:41-70) and RegularPixelSampler<1,1>
// (src/Primitive/PixelSampler.h:34-54); t starts at FLT_MAX.  The launch
// visits the pixels tile by tile (tiles of tw x th pixels in row-major order,
// row-major inside a tile, partial tiles at the right and bottom edges; 1 x 1
// is the reference's row-major order) and stores every output at the pixel
// id, so the order changes only which rays share a wave.
struct CamArgs {
    float pos[3], tl[3], sx[3], sy[3];   // position, top-left direction, pixel steps
    uint32_t res_x, res_y, tw, th;
};
SG_HD uint32_t tile_pixel(const CamArgs& c, uint32_t k) {
    const uint32_t band = c.th * c.res_x;            // pixels of one full row of tiles
    const uint32_t b = k / band, r = k - b * band;
    const uint32_t hb = c.res_y - b * c.th < c.th ? c.res_y - b * c.th : c.th;
    const uint32_t tsz = c.tw * hb;                  // pixels of a full-width tile of this row
    const uint32_t j = r / tsz, s = r - j * tsz;
    const uint32_t wj = c.res_x - j * c.tw < c.tw ? c.res_x - j * c.tw : c.tw;
    return (b * c.th + s / wj) * c.res_x + j * c.tw + s % wj;
}
// Outputs of one trace launch; every pointer is optional (null: not written).
struct TraceOut {
    uint32_t* prim;     // triangle id, 0xFFFFFFFF on a miss
    float* t;           // final t
    float* raybuf;      // SimpleRayBuffer::store (src/Engine/RayBuffers.h:52-77): 8 words per ray
                        // at stride `stride`: hit point org + (t - EPS) dir, dir, t, bestHit
    uint32_t stride;
    float* rays[7];     // camera launches: the generated rays (SoA, at the pixel id)
};
// closest hit of SoA rays (cam == nullptr) or of camera rays (n = res_x *
// res_y, d_n ignored) with the outputs of `out`
void launch_trace_gen(const GridView& g, const RaySoA& r, const CamArgs* cam, uint32_t n, const uint32_t* d_n,
                      const TraceOut& out, hipStream_t st);
// one cosine-weighted bounce per hit, compacted and ordered by direction
// class in rounds of 256 rays (input order within a class); *d_count
// receives the number of bounce rays.  scratch: bounce_scratch_bytes(n).
size_t bounce_scratch_bytes(uint32_t n);
// Bounce rays sorted by a key instead (see sg_bounce_ordered): compaction in
// input order, a stable sort of the keys, a gather.
enum { BOUNCE_CLASS_ROUNDS = 0, BOUNCE_EXIT_FIRST = 1, BOUNCE_AXIS_FIRST = 2 };
struct BounceOrder {
    int mode;
    float bmin[3], bmax[3];   // BOUNCE_EXIT_FIRST: the box the exit distance is measured in
    float scale;              // 65535 / the box diagonal (bounce_exit_scale)
    float axis[3];            // BOUNCE_AXIS_FIRST
};
// 65535 / sqrtf((dx * dx + dy * dy) + dz * dz), d = bmax - bmin, binary32
inline float bounce_exit_scale(const float bmin[3], const float bmax[3]) {
    const float dx = bmax[0] - bmin[0], dy = bmax[1] - bmin[1], dz = bmax[2] - bmin[2];
    return 65535.f / sqrtf((dx * dx + dy * dy) + dz * dz);
}
size_t bounce_ordered_scratch_bytes(uint32_t n);
void launch_bounce_ordered(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                           const uint32_t* d_n, uint64_t seed, const BounceOrder& q, const RaySoAOut& out,
                           uint32_t* src, uint32_t* d_count, void* scratch, hipStream_t st);
void launch_bounce(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                   const uint32_t* d_n, uint64_t seed, const RaySoAOut& out, uint32_t* src, uint32_t* d_count,
                   void* scratch, hipStream_t st);

}  // namespace sg
