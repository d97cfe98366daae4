// sg_kernels.hpp -- host-side launch interface of the gfx950 kernels.
#pragma once

#include "sg_common.hpp"

namespace sg {

// nine SoA arrays: v0.xyz, v1.xyz, v2.xyz (device pointers)
struct TriSoA {
    const float* v[9];
};

// seven SoA arrays: ox, oy, oz, dx, dy, dz, tmax (device pointers)
struct RaySoA {
    const float* a[7];
};
struct RaySoAOut {
    float* a[7];
};

// byte ranges zeroed by k_init (multiples of 16 bytes, 16-byte aligned)
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
    const float4* rec;         // per-triangle traversal record (3 x float4)
    int kind;                  // 0 uniform, 1 two-level
};

// ---------------------------------------------------------------- build
void launch_init(GridHeader* h, uint32_t n_tris, const ZeroSegs& segs, hipStream_t st);
void launch_bounds(const TriSoA& t, uint32_t n, GridHeader* h, hipStream_t st);
void launch_top_emit(const TriSoA& t, uint32_t n, GridHeader* h, const BuildArgs& a, uint32_t* keys,
                     uint32_t* vals, uint32_t* dbg, uint64_t* lb, hipStream_t st);
void launch_ranges(const uint32_t* keys, const uint32_t* d_n, uint32_t cap, const uint32_t* d_C,
                   uint32_t* cells, uint32_t* d_pvalid, hipStream_t st);
void launch_leaf_cells(GridHeader* h, const BuildArgs& a, const uint32_t* ranges, uint32_t* cells,
                       uint32_t* dbg, uint64_t* lb, hipStream_t st);
void launch_leaf_emit(const TriSoA& t, GridHeader* h, const BuildArgs& a, const uint32_t* top_keys,
                      const uint32_t* top_vals, const uint32_t* cells, uint32_t* keys, uint32_t* vals,
                      uint32_t* dbg, uint64_t* lb, hipStream_t st);

// ---------------------------------------------------------------- sort / scan
// Scratch words the sort needs for `cap` pairs sorted over `key_bits` bits.
size_t sort_scratch_words(uint32_t cap, int key_bits, int radix_bits);
int sort_passes(int key_bits, int radix_bits);
// Stable LSD radix sort of (keys,vals) by the low key_bits bits.  n is read
// from device memory (*d_n, clamped to cap).  Input in (k0,v0); the result
// is in buffer index (sort_passes() & 1) of {(k0,v0),(k1,v1)}.  scratch must
// be zero (k_init zeroes it) and holds sort_scratch_words() words;
// tile_ctr points at sort_passes() free, zeroed counters.
void launch_sort_pairs(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, const uint32_t* d_n,
                       uint32_t cap, int key_bits, int radix_bits, uint32_t* scratch, uint32_t* tile_ctr,
                       uint32_t* status, hipStream_t st);

// single-pass scan of n u32 values (decoupled look-back); scratch: scan_scratch_bytes(n),
// zeroed by the call itself.
size_t scan_scratch_bytes(uint32_t n);
void launch_scan(const uint32_t* in, uint32_t* out, uint32_t n, bool inclusive, void* scratch,
                 hipStream_t st);

// ---------------------------------------------------------------- traversal
void launch_prep_tris(const TriSoA& t, uint32_t n, float4* rec, hipStream_t st);
void launch_trace(const GridView& g, const RaySoA& r, uint32_t n, const uint32_t* d_n, uint32_t* prim,
                  float* t, uint32_t* stats, hipStream_t st);
// one cosine-weighted bounce per hit, compacted in primary order; *d_count
// receives the number of bounce rays.  scratch: bounce_scratch_bytes(n).
size_t bounce_scratch_bytes(uint32_t n);
void launch_bounce(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                   uint64_t seed, const RaySoAOut& out, uint32_t* src, uint32_t* d_count, void* scratch,
                   hipStream_t st);

}  // namespace sg
