// sg_build.hip -- grid construction kernels for gfx950.
//
// The build is device-resident: bounds, resolution, counts and totals live in
// the GridHeader in HBM, every kernel sizes itself from it, and no phase
// reads anything back to the host.  Counting, scanning and emission of each
// level are reduce-then-scan (a count kernel leaves one sum per tile; the
// emit kernel folds the sums of the tiles before it and rescans its own), so
// every pair lands at the slot that a serial scan of the counts in canonical
// order would give it: triangles (or sorted top pairs) ascending, then the
// reference's z/y/x loop order.
#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

constexpr int kItems = 4;                      // items per thread in the emission kernels
constexpr int kTile = kBlock * kItems;         // 1024 items per tile
constexpr uint32_t kStage = 4096;              // top pairs a tile stages in LDS before writing them out

// --------------------------------------------------------------- init
// Zero the cell ranges and leaves; reset the header.
// The kernels of the build are written as bodies that take their block
// index and block count (bid, nb): the kernels below call them with
// blockIdx.x and gridDim.x, the multi-grid kernels (k_multi_*) with a
// grid's share of one launch over many grids.
SG_D void init_body(GridHeader* h, uint32_t n_tris, const ZeroSegs& segs, uint32_t bid, uint32_t nb) {
    if (bid == 0 && threadIdx.x == 0) {
        for (int i = 0; i < 3; ++i) {
            h->bmin_key[i] = 0xFFFFFFFFu;
            h->bmax_key[i] = 0u;
        }
        h->n_tris = n_tris;
        h->C1 = h->P1 = h->P1_valid = h->C2 = h->P2 = h->P2_valid = 0;
        h->need_P1 = h->need_C2 = h->need_P2 = 0;
        h->status = 0;
        h->leaf_sort = 0;
        h->oversize = 0;
    }
    const size_t stride = (size_t)nb * blockDim.x;
    for (int s = 0; s < segs.count; ++s) {
        uint4* p = reinterpret_cast<uint4*>(segs.ptr[s]);
        const size_t n16 = segs.bytes[s] >> 4;
        for (size_t i = (size_t)bid * blockDim.x + threadIdx.x; i < n16; i += stride)
            p[i] = make_uint4(0, 0, 0, 0);
    }
}


// --------------------------------------------------------------- bounds
// Scene AABB (the host loop of FastTwoLevelGridBuilder.h:75-81) as a device
// reduction; exact because min/max never round.  Coalesced SoA loads, four
// consecutive triangles per lane (one 16-byte load per SoA array, as the top
// count and emission read them; the arrays are 16-byte aligned).  The
// per-triangle records (TriSoA::rec) are written when the vertices change
// (sg_scene_create / sg_scene_update, k_prep_tris), not here.  Each block
// leaves its six partial keys in `part`; k_params folds them (no
// same-address atomics).
SG_D void bounds_body(const TriSoA& t, uint32_t n, uint32_t* __restrict__ part, uint32_t bid, uint32_t nb) {
    __shared__ uint32_t s_red[6][kBlock / 64];
    uint32_t k[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t i0 = (bid * kBlock + threadIdx.x) * 4u; i0 < n; i0 += nb * kBlock * 4u) {
        if (i0 + 4u <= n) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const float4 p = reinterpret_cast<const float4*>(t.v[a])[i0 / 4];
                const float4 q = reinterpret_cast<const float4*>(t.v[3 + a])[i0 / 4];
                const float4 r = reinterpret_cast<const float4*>(t.v[6 + a])[i0 / 4];
                const uint32_t m0 = min(min(fkey(p.x), fkey(p.y)), min(fkey(p.z), fkey(p.w)));
                const uint32_t m1 = min(min(fkey(q.x), fkey(q.y)), min(fkey(q.z), fkey(q.w)));
                const uint32_t m2 = min(min(fkey(r.x), fkey(r.y)), min(fkey(r.z), fkey(r.w)));
                const uint32_t x0 = max(max(fkey(p.x), fkey(p.y)), max(fkey(p.z), fkey(p.w)));
                const uint32_t x1 = max(max(fkey(q.x), fkey(q.y)), max(fkey(q.z), fkey(q.w)));
                const uint32_t x2 = max(max(fkey(r.x), fkey(r.y)), max(fkey(r.z), fkey(r.w)));
                k[a] = min(k[a], min(m0, min(m1, m2)));
                k[3 + a] = max(k[3 + a], max(x0, max(x1, x2)));
            }
        } else {
            for (uint32_t i = i0; i < n; ++i)
#pragma unroll
                for (int a = 0; a < 3; ++a) {
                    const uint32_t k0 = fkey(t.v[a][i]), k1 = fkey(t.v[3 + a][i]), k2 = fkey(t.v[6 + a][i]);
                    k[a] = min(k[a], min(k0, min(k1, k2)));
                    k[3 + a] = max(k[3 + a], max(k0, max(k1, k2)));
                }
        }
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            k[a] = min(k[a], (uint32_t)__shfl_xor(k[a], d, 64));
            k[3 + a] = max(k[3 + a], (uint32_t)__shfl_xor(k[3 + a], d, 64));
        }
    }
    if (lane_id() == 0)
        for (int a = 0; a < 6; ++a) s_red[a][wave_id()] = k[a];
    __syncthreads();
    if (threadIdx.x < 6) {
        const int a = threadIdx.x;
        uint32_t r = s_red[a][0];
        for (int w = 1; w < kBlock / 64; ++w) r = a < 3 ? min(r, s_red[a][w]) : max(r, s_red[a][w]);
        part[6 * bid + a] = r;
    }
}

// The first launch of a build: the zeroing of the build's segments and the
// bounds reduction, which touch disjoint memory, share one launch (their
// writes and reads then overlap instead of following each other).  Blocks
// below nb_bounds also reduce the bounds; every block zeroes.
__global__ __launch_bounds__(kBlock) void k_init_bounds(GridHeader* h, uint32_t n_tris, ZeroSegs segs, TriSoA t,
                                                        uint32_t* __restrict__ part, uint32_t nb_bounds) {
    init_body(h, n_tris, segs, blockIdx.x, gridDim.x);
    if (blockIdx.x < nb_bounds) bounds_body(t, n_tris, part, blockIdx.x, nb_bounds);
}

// --------------------------------------------------------------- params
// thread 0 of k_params: the folded bounds keys, then (derive) resolution,
// cell sizes and status; s_crcp receives the traversal cell_rcp
SG_D void fold_params(GridHeader* h, const BuildArgs& a, const uint32_t (*s_red)[kBlock / 64], int derive,
                      float* s_crcp) {
    for (int i = 0; i < 6; ++i) {
        uint32_t r = s_red[i][0];
        for (int w = 1; w < kBlock / 64; ++w) r = i < 3 ? min(r, s_red[i][w]) : max(r, s_red[i][w]);
        if (i < 3) h->bmin_key[i] = r;
        else h->bmax_key[i - 3] = r;
    }
    if (!derive) return;
    float mn[3], mx[3], cs[3], rcp[3], crcp[3];
    int32_t res[3];
    for (int i = 0; i < 3; ++i) {
        mn[i] = fkey_inv(h->bmin_key[i]);
        mx[i] = fkey_inv(h->bmax_key[i]);
    }
    const bool ok = derive_params(mn, mx, h->n_tris, a.lambda, a.top_divisor, a.fixed_res, res, cs, rcp, crcp);
    const uint64_t C = ok ? (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2] : 0;
    uint32_t bad = 0;
    if (!ok) bad |= ST_INVALID;
    else if (C >= (1ull << 24)) bad |= ST_TOO_MANY_CELLS;
    else if (C > a.cap_C1) bad |= ST_CAP_C1;
    for (int i = 0; i < 3; ++i) {
        h->bmin[i] = mn[i];
        h->bmax[i] = mx[i];
        h->res[i] = ok ? res[i] : 0;
        h->cs[i] = ok ? cs[i] : 0.f;
        h->rcp_build[i] = ok ? rcp[i] : 0.f;
        h->cell_rcp[i] = ok ? crcp[i] : 0.f;
        s_crcp[i] = h->cell_rcp[i];
    }
    h->C1 = bad ? 0u : (uint32_t)C;
    if (bad) h->status |= bad;
}

// Resolution and cell sizes from the reduced bounds, once (one thread):
// GridBuilder::computeResolution (GridBuilder.h:41-62), the two-level top
// resolution (FastTwoLevelGridBuilder.h:87-102) and the cell sizes of
// FastSimpleGridBuilder.h:114-117 / SimpleGrid.h:99-107.  One block: first
// the bounds partials of the nb bounds blocks are folded into the header;
// with derive the block then fills the leaf cell size table.
SG_D void params_body(GridHeader* h, const BuildArgs& a, const uint32_t* __restrict__ part, uint32_t nb, int derive) {
    __shared__ uint32_t s_red[6][kBlock / 64];
    uint32_t k[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    // four partials per thread and step, loaded before any is folded (one
    // memory round trip per step, not per partial)
    for (uint32_t b0 = threadIdx.x; b0 < nb; b0 += 4u * kBlock) {
        uint32_t v[4][6];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const uint32_t b = b0 + (uint32_t)u * kBlock;
#pragma unroll
            for (int i = 0; i < 6; ++i) v[u][i] = b < nb ? part[6 * b + i] : (i < 3 ? 0xFFFFFFFFu : 0u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int i = 0; i < 6; ++i) k[i] = i < 3 ? min(k[i], v[u][i]) : max(k[i], v[u][i]);
    }
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
        for (int i = 0; i < 6; ++i) {
            const uint32_t o = (uint32_t)__shfl_xor(k[i], d, 64);
            k[i] = i < 3 ? min(k[i], o) : max(k[i], o);
        }
    if (lane_id() == 0)
        for (int i = 0; i < 6; ++i) s_red[i][wave_id()] = k[i];
    __syncthreads();
    __shared__ float s_crcp[3];
    if (threadIdx.x == 0) fold_params(h, a, s_red, derive, s_crcp);
    if (!derive) return;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < 3u * 256u; i += kBlock) {
        const uint32_t d = i >> 8, c = i & 255u;
        h->leaf_size[d][c] = 1.f / (s_crcp[d] * (float)c);
    }
}

__global__ __launch_bounds__(kBlock) void k_params(GridHeader* h, BuildArgs a, const uint32_t* __restrict__ part,
                                                   uint32_t nb, int derive) {
    params_body(h, a, part, nb, derive);
}

// grid parameters as every emission lane needs them (uniform loads)
struct Params {
    v3 gmin, cs, rcp, resf;
    uint32_t rx, rxy;
    bool ok;
};

SG_D Params read_params(const GridHeader* h) {
    Params p;
    p.gmin = mk(h->bmin[0], h->bmin[1], h->bmin[2]);
    p.cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    p.rcp = mk(h->rcp_build[0], h->rcp_build[1], h->rcp_build[2]);
    p.resf = mk((float)h->res[0], (float)h->res[1], (float)h->res[2]);
    p.rx = (uint32_t)h->res[0];
    p.rxy = f2u(p.resf.x * p.resf.y);
    p.ok = h->C1 != 0;
    return p;
}

// BBox::empty + extend x3, src/Primitive/BBox.hpp:78-82,109-119
SG_D void tri_bounds(v3 a, v3 b, v3 c, v3& mn, v3& mx) {
    mn = vmin(vmin(vmin(mk(3.402823466e+38f, 3.402823466e+38f, 3.402823466e+38f), a), b), c);
    mx = vmax(vmax(vmax(mk(-3.402823466e+38f, -3.402823466e+38f, -3.402823466e+38f), a), b), c);
}

// countRefs cell range, src/Control/Algorithm/FastBuildKernels.h:47-69
SG_D uint32_t top_range(const Params& p, v3 mn, v3 mx, uint32_t lo[3], uint32_t hi[3]) {
    v3 lof = vmax(mk(0.f, 0.f, 0.f), (mn - p.gmin) * p.rcp + mk(-kEps, -kEps, -kEps));
    v3 hif = vmin(p.resf - mk(1.f, 1.f, 1.f), (mx - p.gmin) * p.rcp + mk(kEps, kEps, kEps));
    lo[0] = f2u(lof.x); lo[1] = f2u(lof.y); lo[2] = f2u(lof.z);
    hi[0] = f2u(hif.x); hi[1] = f2u(hif.y); hi[2] = f2u(hif.z);
    if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return 0u;
    return (hi[0] - lo[0] + 1u) * (hi[1] - lo[1] + 1u) * (hi[2] - lo[2] + 1u);
}

// buildUnsortedGrid's per-triangle emission, src/Control/Algorithm/FastBuildKernels.h:130-212
SG_D void top_emit_one(const Params& P, const v3& v0, const v3& v1, const v3& v2, uint32_t i, uint32_t slot,
                       uint32_t* __restrict__ keys, uint32_t* __restrict__ vals) {
    v3 mn, mx;
    uint32_t lo[3], hi[3];
    tri_bounds(v0, v1, v2, mn, mx);
    top_range(P, mn, mx, lo, hi);
    if (hi[0] - lo[0] > 0u && hi[1] - lo[1] > 0u && hi[2] - lo[2] > 0u) {
        const uint32_t dummy = f2u(P.resf.x * P.resf.y * P.resf.z);
        const v3 n_ = normalize(cross(v1 - v0, v2 - v0));
        const v3 half = scl(P.cs, 0.5f);
        const v3 c0 = (mk((float)lo[0], (float)lo[1], (float)lo[2]) * P.cs + P.gmin) + half;
        v3 cc;
        cc.z = c0.z - P.cs.z;
        for (uint32_t z = lo[2]; z <= hi[2]; ++z) {
            cc.z += P.cs.z;
            cc.y = c0.y - P.cs.y;
            for (uint32_t y = lo[1]; y <= hi[1]; ++y) {
                cc.y += P.cs.y;
                cc.x = c0.x - P.cs.x;
                for (uint32_t x = lo[0]; x <= hi[0]; ++x, ++slot) {
                    cc.x += P.cs.x;
                    const v3 dv = scl(n_, dot(cc - v0, n_));
                    const bool acc = fabsf(dv.x) <= half.x + kEps && fabsf(dv.y) <= half.y + kEps &&
                                     fabsf(dv.z) <= half.z + kEps;
                    keys[slot] = acc ? x + y * P.rx + z * P.rxy : dummy;
                    vals[slot] = i;
                }
            }
        }
    } else {
        for (uint32_t z = lo[2]; z <= hi[2]; ++z)
            for (uint32_t y = lo[1]; y <= hi[1]; ++y)
                for (uint32_t x = lo[0]; x <= hi[0]; ++x, ++slot) {
                    keys[slot] = x + y * P.rx + z * P.rxy;
                    vals[slot] = i;
                }
    }
}

// ------------------------------------------------------ top level emission
// countRefs + InclusiveScan + buildUnsortedGrid of
// src/Control/Algorithm/FastBuildKernels.h:31-215 as reduce-then-scan: a
// count kernel leaves one sum per 1024-triangle tile, the emit kernel folds
// the sums of the tiles before it, rescans its own counts and writes every
// triangle's pairs at its slot.  Four consecutive triangles per lane (one
// 16-byte load per SoA array); the emit kernel keeps the lane's vertices in
// registers from the count to the emission (reloading them after the scans
// fetched the arrays 2.3 times: a wave of blocks touches more than an XCD's
// L2 holds).
SG_D void top_counts(const Params& P, const TriSoA& t, uint32_t n, uint32_t i0, uint32_t cnt[kItems],
                     float f[9][kItems]) {
    cnt[0] = cnt[1] = cnt[2] = cnt[3] = 0;
    if (!P.ok || i0 >= n) return;
    if (i0 + kItems <= n) {
#pragma unroll
        for (int c = 0; c < 9; ++c) {
            const float4 q = reinterpret_cast<const float4*>(t.v[c])[i0 / 4];
            f[c][0] = q.x; f[c][1] = q.y; f[c][2] = q.z; f[c][3] = q.w;
        }
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c)
#pragma unroll
            for (int j = 0; j < kItems; ++j) f[c][j] = i0 + j < n ? t.v[c][i0 + j] : 0.f;
    }
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (i0 + j < n) {
            v3 mn, mx;
            uint32_t lo[3], hi[3];
            tri_bounds(mk(f[0][j], f[1][j], f[2][j]), mk(f[3][j], f[4][j], f[5][j]), mk(f[6][j], f[7][j], f[8][j]),
                       mn, mx);
            cnt[j] = top_range(P, mn, mx, lo, hi);
        }
    }
}

SG_D void top_count_body(const TriSoA& t, uint32_t n, const GridHeader* h, uint32_t* __restrict__ sums,
                         uint32_t* __restrict__ dbg_counts, uint32_t bid) {
    __shared__ uint32_t s_red[kBlock / 64];
    const Params P = read_params(h);
    const uint32_t i0 = bid * kTile + threadIdx.x * kItems;
    uint32_t cnt[kItems];
    float f[9][kItems];
    top_counts(P, t, n, i0, cnt, f);
    if (dbg_counts)
        for (int j = 0; j < kItems; ++j)
            if (i0 + j < n) dbg_counts[i0 + j] = cnt[j];
    const uint32_t s = block_sum<kBlock / 64>(cnt[0] + cnt[1] + cnt[2] + cnt[3], s_red);
    if (threadIdx.x == 0) sums[bid] = s;
}

__global__ __launch_bounds__(kBlock) void k_top_count(TriSoA t, uint32_t n, GridHeader* h,
                                                      uint32_t* __restrict__ sums,
                                                      uint32_t* __restrict__ dbg_counts) {
    top_count_body(t, n, h, sums, dbg_counts, blockIdx.x);
}

SG_D void top_emit_body(const TriSoA& t, uint32_t n, GridHeader* h, const BuildArgs& a, uint32_t* __restrict__ keys,
                        uint32_t* __restrict__ vals, const uint32_t* __restrict__ sums, uint32_t bid, uint32_t nb) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_k[kStage], s_v[kStage];
    const Params P = read_params(h);
    const uint32_t prefix = block_prefix_of<kBlock / 64>(sums, bid, s_scan);
    const uint32_t i0 = bid * kTile + threadIdx.x * kItems;
    uint32_t cnt[kItems];
    float f[9][kItems];
    top_counts(P, t, n, i0, cnt, f);
    const uint32_t sum = cnt[0] + cnt[1] + cnt[2] + cnt[3];
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(sum, &agg, s_scan);
    if (bid == nb - 1 && threadIdx.x == 0) {
        const uint32_t P1 = prefix + agg;
        h->need_P1 = P1;
        h->P1 = P1 <= a.cap_P1 ? P1 : a.cap_P1;
        if (P1 > a.cap_P1) atomicOr(&h->status, ST_CAP_P1);
    }
    if ((uint64_t)prefix + agg > a.cap_P1) return;   // flagged by the last tile
    if (agg <= kStage) {
        // the tile's pairs are one contiguous run: build it in LDS, then
        // write it out with full cache lines
        uint32_t ls = local;
#pragma unroll
        for (int j = 0; j < kItems; ++j) {
            if (cnt[j])
                top_emit_one(P, mk(f[0][j], f[1][j], f[2][j]), mk(f[3][j], f[4][j], f[5][j]),
                             mk(f[6][j], f[7][j], f[8][j]), i0 + j, ls, s_k, s_v);
            ls += cnt[j];
        }
        __syncthreads();
        for (uint32_t j = threadIdx.x; j < agg; j += kBlock) {
            keys[prefix + j] = s_k[j];
            vals[prefix + j] = s_v[j];
        }
        return;
    }
    uint32_t slot = prefix + local;
#pragma unroll
    for (int j = 0; j < kItems; ++j) {
        if (cnt[j])
            top_emit_one(P, mk(f[0][j], f[1][j], f[2][j]), mk(f[3][j], f[4][j], f[5][j]),
                         mk(f[6][j], f[7][j], f[8][j]), i0 + j, slot, keys, vals);
        slot += cnt[j];
    }
}

__global__ __launch_bounds__(kBlock) void k_top_emit(TriSoA t, uint32_t n, GridHeader* h, BuildArgs a,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals,
                                                     const uint32_t* __restrict__ sums) {
    top_emit_body(t, n, h, a, keys, vals, sums, blockIdx.x, gridDim.x);
}

// ---------------------------------------------------------------- ranges
// prepareCellRanges / prepareLeafCellRanges
// (src/Control/Algorithm/FastBuildKernels.h:218-313,
// src/Control/Algorithm/TwoLevelBuildKernels.h:458-536): a neighbour compare
// on the sorted keys writes {first, last+1} of every key below C; the dummy
// key's range is not stored.  Also records P_valid = index of the first
// dummy key.  Cells were zeroed by the build's first launch.
SG_D void ranges_body(const uint32_t* __restrict__ keys, const uint32_t* d_n, uint32_t cap, const uint32_t* d_C,
                      uint32_t* __restrict__ cells, uint32_t* d_pvalid, uint32_t bid, uint32_t nb) {
    const uint32_t n = min(*d_n, cap);
    const uint32_t C = *d_C;
    if (n == 0) {
        if (bid == 0 && threadIdx.x == 0) *d_pvalid = 0;
        return;
    }
    for (uint32_t i = bid * kBlock + threadIdx.x; i < n; i += nb * kBlock) {
        const uint32_t key = keys[i];
        const uint32_t prev = i > 0 ? keys[i - 1] : 0xFFFFFFFFu;
        const uint32_t next = i + 1 < n ? keys[i + 1] : 0xFFFFFFFFu;
        if (key < C) {
            if (key != prev) cells[2 * (size_t)key] = i;
            if (key != next) cells[2 * (size_t)key + 1] = i + 1;
            if (i + 1 == n) *d_pvalid = n;
        } else if (i == 0 || prev < C) {
            *d_pvalid = i;
        }
    }
}

__global__ __launch_bounds__(kBlock) void k_ranges(const uint32_t* __restrict__ keys, const uint32_t* d_n,
                                                   uint32_t cap, const uint32_t* d_C,
                                                   uint32_t* __restrict__ cells, uint32_t* d_pvalid) {
    ranges_body(keys, d_n, cap, d_C, cells, d_pvalid, blockIdx.x, gridDim.x);
}

// ------------------------------------------------------------ leaf cells
// countLeafLevelCells<false> + ExclusiveScan + prepareTopLevelCellRanges
// (src/Control/Algorithm/TwoLevelBuildKernels.h:40-127,
// src/Control/Algorithm/FastTwoLevelGridBuilder.h:256-294): count kernel +
// write kernel over tiles of 256 top cells.
SG_D uint32_t leaf_cell(const GridHeader* h, const BuildArgs& a, const uint32_t* __restrict__ ranges, uint32_t k,
                        uint32_t* bits) {
    const v3 cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    const uint2 rg = reinterpret_cast<const uint2*>(ranges)[k];
    const uint32_t r = rg.y - rg.x;
    uint32_t rx = 1, ry = 1, rz = 1;
    if (r > a.leaf_min_refs) {
        const float vol = cs.x * cs.y * cs.z;
        const float m = cbrtpow(a.lambda_leaf * (float)r / vol);
        const v3 rf = scl(cs, m);
        rx = f2u(rf.x); ry = f2u(rf.y); rz = f2u(rf.z);
    }
    *bits = (r == 0 ? 0u : kNotEmpty) | (rx & 0xFFu) | ((ry & 0xFFu) << 8) | ((rz & 0xFFu) << 16);
    return rx * ry * rz;   // unmasked, as TwoLevelBuildKernels.h:81
}

// every leaf resolution in [1,255] (the packed bytes hold the whole value)
SG_D bool leaf_res_ok(uint32_t bits, uint32_t cnt) {
    const uint32_t rx = bits & 0xFFu, ry = (bits >> 8) & 0xFFu, rz = (bits >> 16) & 0xFFu;
    return rx >= 1 && ry >= 1 && rz >= 1 && rx * ry * rz == cnt;
}

__global__ __launch_bounds__(kBlock) void k_leaf_cells_count(GridHeader* h, BuildArgs a,
                                                             const uint32_t* __restrict__ ranges,
                                                             uint32_t* __restrict__ sums,
                                                             uint32_t* __restrict__ dbg_counts) {
    __shared__ uint32_t s_red[kBlock / 64];
    const uint32_t C1 = h->C1;
    const uint32_t base = blockIdx.x * kBlock;
    if (base >= C1) return;
    const uint32_t k = base + threadIdx.x;
    uint32_t cnt = 0, bits = 0;
    if (k < C1) {
        cnt = leaf_cell(h, a, ranges, k, &bits);
        if (!leaf_res_ok(bits, cnt)) atomicOr(&h->status, ST_LEAF_RES);
        if (dbg_counts) dbg_counts[k] = cnt;
    }
    const uint32_t s = block_sum<kBlock / 64>(cnt, s_red);
    if (threadIdx.x == 0) sums[blockIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_leaf_cells_write(GridHeader* h, BuildArgs a,
                                                             const uint32_t* __restrict__ ranges,
                                                             uint32_t* __restrict__ cells,
                                                             const uint32_t* __restrict__ sums) {
    __shared__ uint32_t s_scan[kBlock / 64];
    const uint32_t C1 = h->C1;
    const uint32_t base = blockIdx.x * kBlock;
    if (base >= C1) return;
    const uint32_t prefix = block_prefix_of<kBlock / 64>(sums, blockIdx.x, s_scan);
    const uint32_t k = base + threadIdx.x;
    uint32_t cnt = 0, bits = 0;
    if (k < C1) cnt = leaf_cell(h, a, ranges, k, &bits);
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(cnt, &agg, s_scan);
    if (k < C1) reinterpret_cast<uint2*>(cells)[k] = make_uint2(prefix + local, bits);
    if (base + kBlock >= C1 && threadIdx.x == 0) {
        const uint32_t C2 = prefix + agg;
        h->need_C2 = C2;
        if (C2 > a.cap_C2) {
            atomicOr(&h->status, ST_CAP_C2);
            h->C2 = a.cap_C2;
        } else {
            h->C2 = C2;
        }
    }
}

// ------------------------------------------------------------- leaf pairs
// countLeafLevelRefs + InclusiveScan + writeLeafLevelRefs and the leaf-level
// radix sort (src/Control/Algorithm/TwoLevelBuildKernels.h:129-455,
// src/Control/Algorithm/FastTwoLevelGridBuilder.h:487-488,711-712).
//
// The leaf pairs of top cell k carry keys in [cells[k].x, cells[k].x + leaf
// cells of k), a range no other top cell uses, and pairs the plane test
// rejects carry the key C2.  The stable sort of all leaf pairs by key is
// therefore the concatenation, over top cells in order, of the stable sort of
// each top cell's accepted pairs, followed by every rejected pair in emission
// order.  Leaf tiles hold whole top cells: tile t covers the sorted top pairs
// from the first top-cell start at or after t*kLeafTile to the first one at
// or after (t+1)*kLeafTile.  The count kernel runs the plane test and leaves
// the tile's accepted and rejected totals; the emit kernel builds the tile's
// pairs in LDS as (tile-local key << 12 | index of the top pair within the
// tile), sorts them by key there and writes accepted pairs to their final
// slots and rejected ones to the tail.  A tile too large for LDS writes
// its pairs in emission order and raises GridHeader::leaf_sort, which enables
// a global stable radix sort of the accepted pairs (a no-op otherwise).
constexpr uint32_t kLeafTile = 512;                   // nominal sorted top pairs per leaf tile
constexpr int kLeafBlock = 512;                       // threads of the leaf emission kernel
constexpr uint32_t kLeafStage = 4096;                 // leaf pairs a tile sorts in LDS
constexpr int kPayBits = 12;                          // payload: top pair index within the tile
constexpr int kLocalBits = 16;                        // tile-local key bits the LDS sort handles
constexpr int kDigitBits = 6;                         // widest digit of an LDS sort pass
constexpr int kLeafRounds = kLeafStage / kLeafBlock;  // LDS sort rounds per pass

struct LeafParams {
    v3 gmin, cs;
    uint32_t rx, rxy, C1, C2, n_tris;
};

struct LeafSetup {
    v3 v0, v1, v2, org, cres;
    int lo[3], hp[3];
    uint32_t px;
};

SG_D LeafParams read_leaf_params(const GridHeader* h) {
    LeafParams L;
    L.gmin = mk(h->bmin[0], h->bmin[1], h->bmin[2]);
    L.cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    L.rx = (uint32_t)h->res[0];
    L.rxy = (uint32_t)h->res[0] * (uint32_t)h->res[1];
    L.C1 = h->C1;
    L.C2 = h->C2;
    L.n_tris = h->n_tris;
    return L;
}

// leaf-cell box of one (top cell, triangle) pair; returns the number of leaf
// cells it spans, -1 for a negative extent
SG_D int leaf_range(const TriSoA& t, uint32_t key, uint32_t tri, const uint32_t* __restrict__ cells,
                    const LeafParams& L, LeafSetup& s) {
    const uint32_t tx = key % L.rx;
    const uint32_t ty = (key % L.rxy) / L.rx;
    const uint32_t tz = key / L.rxy;
    v3 mn, mx;
    load_tri_record(t.rec, tri, s.v0, s.v1, s.v2);
    tri_bounds(s.v0, s.v1, s.v2, mn, mx);
    const uint2 cell = reinterpret_cast<const uint2*>(cells)[key];
    s.px = cell.x;
    s.cres = mk((float)(cell.y & 0xFFu), (float)((cell.y >> 8) & 0xFFu), (float)((cell.y >> 16) & 0xFFu));
    s.org = mk((float)tx * L.cs.x + L.gmin.x, (float)ty * L.cs.y + L.gmin.y, (float)tz * L.cs.z + L.gmin.z);
    const v3 srcp = s.cres / L.cs;
    const v3 lof = (mn - s.org) * srcp + mk(-0.001f, -0.001f, -0.001f);
    const v3 hif = (mx - s.org) * srcp + mk(1.001f, 1.001f, 1.001f);
    int q;
    q = f2i(lof.x); s.lo[0] = q > 0 ? q : 0;
    q = f2i(lof.y); s.lo[1] = q > 0 ? q : 0;
    q = f2i(lof.z); s.lo[2] = q > 0 ? q : 0;
    int rc;
    rc = (int)s.cres.x; q = f2i(hif.x); s.hp[0] = rc < q ? rc : q;
    rc = (int)s.cres.y; q = f2i(hif.y); s.hp[1] = rc < q ? rc : q;
    rc = (int)s.cres.z; q = f2i(hif.z); s.hp[2] = rc < q ? rc : q;
    const int ex = s.hp[0] - s.lo[0], ey = s.hp[1] - s.lo[1], ez = s.hp[2] - s.lo[2];
    if (ex < 0 || ey < 0 || ez < 0) return -1;
    return ex * ey * ez;
}

// Visit the leaf cells of a pair in the reference's z/y/x order:
// f(key, accepted).  A triangle spanning more than one leaf cell on every axis
// keeps a cell only if the cell centre lies within half a cell (+EPS) of the
// triangle's plane on every axis.  Keys at or beyond C2 (a leaf level that
// overflowed its capacity) count as rejected.
template <class F>
SG_D void leaf_visit(const LeafParams& L, const LeafSetup& s, F&& f) {
    const uint32_t crx = f2u(s.cres.x), crxy = f2u(s.cres.x * s.cres.y);
    if (s.hp[0] - s.lo[0] > 1 && s.hp[1] - s.lo[1] > 1 && s.hp[2] - s.lo[2] > 1) {
        const v3 n_ = normalize(cross(s.v1 - s.v0, s.v2 - s.v0));
        const v3 scs = L.cs / s.cres;
        const v3 c0 = (mk((float)s.lo[0], (float)s.lo[1], (float)s.lo[2]) * scs + s.org) + scl(scs, 0.5f);
        v3 cc;
        cc.z = c0.z - scs.z;
        for (int z = s.lo[2]; z < s.hp[2]; ++z) {
            cc.z += scs.z;
            cc.y = c0.y - scs.y;
            for (int y = s.lo[1]; y < s.hp[1]; ++y) {
                cc.y += scs.y;
                cc.x = c0.x - scs.x;
                for (int x = s.lo[0]; x < s.hp[0]; ++x) {
                    cc.x += scs.x;
                    const v3 dv = scl(n_, dot(cc - s.v0, n_));
                    const bool acc = fabsf(dv.x) <= scs.x * 0.5f + kEps && fabsf(dv.y) <= scs.y * 0.5f + kEps &&
                                     fabsf(dv.z) <= scs.z * 0.5f + kEps;
                    const uint32_t key = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + s.px;
                    f(key, acc && key < L.C2);
                }
            }
        }
    } else {
        for (int z = s.lo[2]; z < s.hp[2]; ++z)
            for (int y = s.lo[1]; y < s.hp[1]; ++y)
                for (int x = s.lo[0]; x < s.hp[0]; ++x) {
                    const uint32_t key = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + s.px;
                    f(key, key < L.C2);
                }
    }
}

// set-up of sorted top pair i; 0 pairs for a negative extent (flagged) or a
// top level that overflowed its buffers (never follow an id outside the
// scene or the grid)
SG_D uint32_t leaf_pair(const TriSoA& t, const LeafParams& L, GridHeader* h, uint32_t key, uint32_t tri,
                        const uint32_t* __restrict__ cells, LeafSetup& s) {
    if (tri >= L.n_tris || key >= L.C1) return 0;
    const int c = leaf_range(t, key, tri, cells, L, s);
    if (c < 0) {
        atomicOr(&h->status, ST_NEG_EXTENT);
        return 0;
    }
    return (uint32_t)c;
}

// Leaf tile of a sorted top pair: tile t holds the top cells that start in
// [t*kLeafTile, (t+1)*kLeafTile), so pair i belongs to tile
// floor(start of its top cell / kLeafTile).  Per tile the count kernel
// records (zero-initialised words, combined with atomicMax):
//   info[4t+0] = ~first pair   info[4t+1] = end pair
//   info[4t+2] = ~first leaf key   info[4t+3] = last leaf key + 1
// and its accepted / rejected leaf-pair totals in sums[t] / sums[Tp + t].
SG_HD uint32_t leaf_sums_stride(uint32_t T) { return (T + 15u) & ~15u; }

// One sorted top pair per lane: the lane with the most leaf cells to visit
// sets a wave's time, and with four consecutive pairs per lane (one 16-byte
// load per array) those lanes ran four set-ups and visits in a row --
// height-field rebuild 0.387 -> 0.364 ms, ruins 0.392 -> 0.315 ms with one.
__global__ __launch_bounds__(kBlock) void k_leaf_count(TriSoA t, GridHeader* h,
                                                       const uint32_t* __restrict__ top_keys,
                                                       const uint32_t* __restrict__ top_vals,
                                                       const uint32_t* __restrict__ ranges,
                                                       const uint32_t* __restrict__ cells,
                                                       uint32_t* __restrict__ sums, uint32_t* __restrict__ info,
                                                       uint32_t T, uint32_t* __restrict__ dbg_counts,
                                                       uint32_t* __restrict__ pacc, uint32_t* __restrict__ prej) {
    constexpr int kSlots = 4;
    __shared__ uint32_t s_acc[kSlots], s_rej[kSlots];
    const uint32_t P1v = h->P1_valid;
    const uint32_t base = blockIdx.x * kBlock;
    if (base >= P1v) return;
    if (threadIdx.x < kSlots) s_acc[threadIdx.x] = s_rej[threadIdx.x] = 0;
    const LeafParams L = read_leaf_params(h);
    const uint32_t Tp = leaf_sums_stride(T);
    const uint32_t i = base + threadIdx.x;
    const uint32_t key = i < P1v ? top_keys[i] : 0u;
    const uint32_t tri = i < P1v ? top_vals[i] : 0u;
    // the block's first pair may sit in a top cell that starts far earlier:
    // slot 0 is that pair's tile, slots 1.. the tiles from base on
    const uint32_t first_key = top_keys[base];
    const uint32_t tile0 = first_key < L.C1 ? reinterpret_cast<const uint2*>(ranges)[first_key].x / kLeafTile : 0u;
    const uint32_t tile1 = base / kLeafTile;
    __syncthreads();
    if (i < P1v && key < L.C1) {
        const uint2 rg = reinterpret_cast<const uint2*>(ranges)[key];
        const uint32_t tile = rg.x / kLeafTile;
        LeafSetup st;
        const uint32_t c = leaf_pair(t, L, h, key, tri, cells, st);
        if (dbg_counts) dbg_counts[i] = c;
        uint32_t acc = 0;
        if (c) {
            // a box of one leaf cell on some axis has no plane test: every
            // cell is accepted unless its key reaches C2, and the largest
            // key is that of the far corner -- then the count is the volume
            const bool plane = st.hp[0] - st.lo[0] > 1 && st.hp[1] - st.lo[1] > 1 && st.hp[2] - st.lo[2] > 1;
            const uint32_t crx = f2u(st.cres.x), crxy = f2u(st.cres.x * st.cres.y);
            const uint32_t kmax = (uint32_t)(st.hp[0] - 1) + (uint32_t)(st.hp[1] - 1) * crx +
                                  (uint32_t)(st.hp[2] - 1) * crxy + st.px;
            if (!plane && kmax < L.C2)
                acc = c;
            else
                leaf_visit(L, st, [&](uint32_t, bool ok) { acc += ok ? 1u : 0u; });
        }
        if (pacc) {   // per-pair totals for the parallel emission of oversize tiles
            pacc[i] = acc;
            prej[i] = c - acc;
        }
        if (tile < T) {
            const uint32_t slot = tile == tile0 ? 0u : tile - tile1 + 1u;
            if (slot < kSlots) {
                if (acc) atomicAdd(&s_acc[slot], acc);
                if (c - acc) atomicAdd(&s_rej[slot], c - acc);
            } else {
                atomicAdd(&sums[tile], acc);
                atomicAdd(&sums[Tp + tile], c - acc);
            }
            if (rg.x == i) {   // first pair of its top cell
                const uint2 cl = reinterpret_cast<const uint2*>(cells)[key];
                const uint32_t nl = (cl.y & 0xFFu) * ((cl.y >> 8) & 0xFFu) * ((cl.y >> 16) & 0xFFu);
                atomicMax(&info[4 * (size_t)tile + 0], ~i);
                atomicMax(&info[4 * (size_t)tile + 1], rg.y);
                atomicMax(&info[4 * (size_t)tile + 2], ~cl.x);
                atomicMax(&info[4 * (size_t)tile + 3], cl.x + nl);
            }
        }
    }
    __syncthreads();
    if (threadIdx.x < kSlots) {
        const uint32_t slot = threadIdx.x;
        const uint32_t tile = slot == 0 ? tile0 : tile1 + slot - 1u;
        if (tile < T && (slot == 0 || tile != tile0)) {
            if (s_acc[slot]) atomicAdd(&sums[tile], s_acc[slot]);
            if (s_rej[slot]) atomicAdd(&sums[Tp + tile], s_rej[slot]);
        }
    }
}

// Exclusive scan, in place, of the accepted (sums[0..T)) and rejected
// (sums[Tp..Tp+T)) tile totals, both halves at once (waves 0-7 and 8-15);
// sums[2Tp], sums[2Tp+1] receive the grand totals and the header the
// leaf-pair counts (one block of 1024 threads).
__global__ __launch_bounds__(1024) void k_leaf_scan(GridHeader* h, BuildArgs a, uint32_t* __restrict__ sums,
                                                    uint32_t T) {
    constexpr int kPer = 16;
    __shared__ uint32_t s_w[16];
    __shared__ uint32_t s_tot[2];
    const uint32_t Tp = leaf_sums_stride(T);
    const int w = wave_id(), half = w >> 3;
    const uint32_t tt = threadIdx.x & 511u;
    uint32_t* row = sums + (size_t)half * Tp;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < T; c0 += 512u * kPer) {
        const uint32_t i0 = c0 + tt * kPer;
        uint32_t v[kPer], sum = 0;
        if (i0 + kPer <= Tp) {
#pragma unroll
            for (int q = 0; q < kPer / 4; ++q) {
                const uint4 x = reinterpret_cast<const uint4*>(row + i0)[q];
                v[4 * q] = x.x; v[4 * q + 1] = x.y; v[4 * q + 2] = x.z; v[4 * q + 3] = x.w;
            }
        } else {
#pragma unroll
            for (int k = 0; k < kPer; ++k) v[k] = 0;
        }
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (i0 + k >= T) v[k] = 0;
            sum += v[k];
        }
        const uint32_t inc = wave_incl_scan(sum);
        if (lane_id() == 63) s_w[w] = inc;
        __syncthreads();
        uint32_t pre = 0, tot = 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const uint32_t x = s_w[8 * half + j];
            pre += (8 * half + j < w) ? x : 0u;
            tot += x;
        }
        __syncthreads();
        uint32_t p = run + pre + inc - sum;
#pragma unroll
        for (int k = 0; k < kPer; ++k) {
            if (i0 + k < T) row[i0 + k] = p;
            p += v[k];
        }
        run += tot;
    }
    if (tt == 0) s_tot[half] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        const uint32_t ta = s_tot[0], tr = s_tot[1];
        const uint64_t total = (uint64_t)ta + tr;
        sums[2 * (size_t)Tp] = ta;
        sums[2 * (size_t)Tp + 1] = tr;
        h->need_P2 = (uint32_t)min(total, (uint64_t)0xFFFFFFFFu);
        if (total > a.cap_P2) {
            atomicOr(&h->status, ST_CAP_P2);
            h->P2 = a.cap_P2;
            h->P2_valid = min(ta, a.cap_P2);
        } else {
            h->P2 = (uint32_t)total;
            h->P2_valid = ta;
        }
    }
}

// One stable LSD pass over n <= kLeafStage words: digit (x >> shift) & (2^db - 1),
// db <= kDigitBits, in -> out, by a block of kLeafBlock threads.  Ranks within
// a wave come from ballots over the digit bits (computed once for the digit
// histogram and again for the scatter, which keeps registers low); waves are
// ordered through per-wave digit counts, rounds through running bases.
SG_D void lds_sort_pass(const uint32_t* in, uint32_t* out, uint32_t n, int shift, int db, uint32_t* s_base,
                        uint32_t* s_wc, uint32_t* s_scan) {
    constexpr int NW = kLeafBlock / 64;
    constexpr int RM = 1 << kDigitBits;
    static_assert(RM * NW <= kLeafBlock, "one zeroing store per thread");
    const uint32_t R = 1u << db, mask = R - 1u;
    const int w = wave_id();
    const uint64_t lt = lanemask_lt();
    if (threadIdx.x < RM * NW) s_wc[threadIdx.x] = 0;
    if (threadIdx.x < RM) s_base[threadIdx.x] = 0;
    __syncthreads();
    for (uint32_t r0 = 0; r0 < n; r0 += kLeafBlock) {
        const uint32_t i = r0 + threadIdx.x;
        const bool ok = i < n;
        const uint32_t d = ok ? (in[i] >> shift) & mask : 0u;
        const uint64_t peers = digit_peers(ok, d, db);
        if (ok && (peers & lt) == 0) atomicAdd(&s_base[d], (uint32_t)__popcll(peers));
    }
    __syncthreads();
    {
        const uint32_t v = threadIdx.x < R ? s_base[threadIdx.x] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<NW>(v, &tot, s_scan);
        if (threadIdx.x < R) s_base[threadIdx.x] = ex;
        __syncthreads();
    }
    for (uint32_t r0 = 0; r0 < n; r0 += kLeafBlock) {
        const uint32_t i = r0 + threadIdx.x;
        const bool ok = i < n;
        const uint32_t x = ok ? in[i] : 0u;
        const uint32_t d = (x >> shift) & mask;
        const uint64_t peers = digit_peers(ok, d, db);
        const uint32_t rk = (uint32_t)__popcll(peers & lt);
        if (ok && rk == 0) s_wc[w * RM + d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (ok) {
            uint32_t pos = s_base[d] + rk;
            for (int v = 0; v < w; ++v) pos += s_wc[v * RM + d];
            out[pos] = x;
        }
        __syncthreads();
        if (threadIdx.x < R) {
            uint32_t s = 0;
#pragma unroll
            for (int v = 0; v < NW; ++v) {
                s += s_wc[v * RM + threadIdx.x];
                s_wc[v * RM + threadIdx.x] = 0;
            }
            s_base[threadIdx.x] += s;
        }
        __syncthreads();
    }
}

// Where tile b's pairs go: prefixes from k_leaf_scan, tile bounds and leaf
// key range from k_leaf_count; `staged` when the tile fits the LDS sort.
struct LeafTilePlan {
    uint32_t a0, a1, pre_acc, pre_rej, tot_acc, kbase, span;
    bool staged;
};

SG_D bool leaf_tile_plan(const BuildArgs& a, const uint32_t* __restrict__ sums, const uint32_t* __restrict__ info,
                         uint32_t T, uint32_t b, LeafTilePlan& q) {
    const uint32_t Tp = leaf_sums_stride(T);
    const uint32_t tot_acc = sums[2 * (size_t)Tp], tot_rej = sums[2 * (size_t)Tp + 1];
    if ((uint64_t)tot_acc + tot_rej > a.cap_P2) return false;   // flagged by k_leaf_scan
    const uint4 inf = reinterpret_cast<const uint4*>(info)[b];
    if (inf.x == 0) return false;   // no top cell starts in this tile
    q.a0 = ~inf.x;
    q.a1 = inf.y > q.a0 ? inf.y : q.a0;
    if (q.a0 >= q.a1) return false;
    q.tot_acc = tot_acc;
    q.pre_acc = sums[b];
    q.pre_rej = sums[Tp + b];
    const uint32_t n_tile = ((b + 1 < T ? sums[b + 1] : tot_acc) - q.pre_acc) +
                            ((b + 1 < T ? sums[Tp + b + 1] : tot_rej) - q.pre_rej);
    q.kbase = ~inf.z;
    q.span = inf.w >= q.kbase ? inf.w - q.kbase : 0xFFFFFFFFu;
    q.staged = n_tile <= kLeafStage && q.span < (1u << kLocalBits) && q.a1 - q.a0 <= (1u << kPayBits);
    return true;
}

// tiles that fit the LDS sort
__global__ __launch_bounds__(kLeafBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_leaf_emit(TriSoA t, GridHeader* h, BuildArgs a,
                                                         const uint32_t* __restrict__ top_keys,
                                                         const uint32_t* __restrict__ top_vals,
                                                         const uint32_t* __restrict__ cells,
                                                         uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                         float* __restrict__ refs,
                                                         const uint32_t* __restrict__ sums,
                                                         const uint32_t* __restrict__ info, uint32_t T,
                                                         uint32_t* __restrict__ leaves, int write_keys) {
    constexpr int NW = kLeafBlock / 64;
    __shared__ uint32_t s_scan[NW];
    __shared__ uint32_t s_a[kLeafStage], s_b[kLeafStage];
    __shared__ uint32_t s_base[1 << kDigitBits], s_wc[(1 << kDigitBits) * NW];
    static_assert(kLeafStage >= (1u << kPayBits), "the spare sort buffer holds the tile's triangle ids");
    LeafTilePlan q;
    if (!leaf_tile_plan(a, sums, info, T, blockIdx.x, q) || !q.staged) return;
    const LeafParams L = read_leaf_params(h);
    const uint32_t a0 = q.a0, a1 = q.a1, kbase = q.kbase, span = q.span;
    const uint32_t cap = a.cap_P2;
    const uint32_t rej_base = q.tot_acc + q.pre_rej;

    // pairs in emission order as (local key << kPayBits | top pair); a
    // rejected pair takes the local key `span`, after every accepted one
    uint32_t run = 0, n_acc = 0;
    for (uint32_t c0 = a0; c0 < a1; c0 += kLeafBlock) {
        const uint32_t i = c0 + threadIdx.x;
        LeafSetup s;
        uint32_t c = 0;
        if (i < a1) c = leaf_pair(t, L, h, top_keys[i], top_vals[i], cells, s);
        uint32_t agg;
        uint32_t p = run + block_excl_scan<NW>(c, &agg, s_scan);
        const uint32_t pay = i - a0;
        if (c)
            leaf_visit(L, s, [&](uint32_t key, bool ok) {
                const uint32_t lk = (ok && key >= kbase && key - kbase < span) ? key - kbase : span;
                n_acc += lk < span ? 1u : 0u;
                if (p < kLeafStage) s_a[p] = (lk << kPayBits) | pay;
                ++p;
            });
        run += agg;
    }
    const uint32_t acc_t = block_sum<NW>(n_acc, s_scan);   // syncs: staging complete
    const uint32_t n = run < kLeafStage ? run : kLeafStage;
    const int bits = 32 - __builtin_clz(span | 1u);   // local keys 0..span
    const int passes = (bits + kDigitBits - 1) / kDigitBits;
    const int db = (bits + passes - 1) / passes;
    uint32_t* src = s_a;
    uint32_t* dst = s_b;
    for (int k = 0, lo = 0; k < passes; ++k, lo += db) {
        lds_sort_pass(src, dst, n, kPayBits + lo, min(db, bits - lo), s_base, s_wc, s_scan);
        uint32_t* tmp = src;
        src = dst;
        dst = tmp;
    }
    // the triangle of each top pair of the tile, in the sort buffer the last
    // pass read (free now): the write-out reads it there instead of gathering
    // it from the sorted top pairs.  Sharing the buffer keeps the block's LDS
    // at 35 KB, four blocks per CU.
    uint32_t* s_tri = dst;
    for (uint32_t j = threadIdx.x; j < a1 - a0; j += kLeafBlock) s_tri[j] = top_vals[a0 + j];
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += kLeafBlock) {
        const uint32_t x = src[j];
        const uint32_t v = s_tri[x & ((1u << kPayBits) - 1u)];
        const uint32_t lk = x >> kPayBits;
        uint32_t idx, key;
        if (j < acc_t) {
            idx = q.pre_acc + j;
            key = kbase + lk;
        } else {
            idx = rej_base + (j - acc_t);
            key = L.C2;
        }
        if (idx < cap) {
            if (write_keys) keys[idx] = key;
            vals[idx] = v;
            if (j < acc_t) {
                // accepted pairs are at their final slots: their reference
                // records too (the tile's triangles were read while staging)
                write_ref_record(refs, idx, t.rec, v, L.n_tris);
                // and the ranges of the tile's leaf cells, which no other
                // tile holds: {first, last + 1} of each key, as k_ranges
                // writes them from the sorted keys
                if (leaves) {
                    if (j == 0 || (src[j - 1] >> kPayBits) != lk) leaves[2 * (size_t)key] = idx;
                    if (j + 1 == acc_t || (src[j + 1] >> kPayBits) != lk) leaves[2 * (size_t)key + 1] = idx + 1u;
                }
            }
        }
    }
}

// One stable LSD pass by a block of kBlock threads over n (key, value)
// pairs in global memory: digit ((k - kbase) >> shift) & (2^db - 1), db <= 8,
// (k, v) -> (ko, vo).  Ranks as lds_sort_pass: a digit histogram, its scan,
// then rounds of kBlock pairs ranked by wave ballots and ordered across
// waves by per-wave digit counts.
static_assert(kBlock == 256, "one digit bin per thread");
SG_D void block_sort_pass_global(const uint32_t* k, const uint32_t* v, uint32_t* ko, uint32_t* vo, uint32_t n,
                                 uint32_t kbase, int shift, int db, uint32_t* s_base, uint32_t* s_wc,
                                 uint32_t* s_scan) {
    constexpr int NW = kBlock / 64;
    constexpr int RM = 256;
    const uint32_t R = 1u << db, mask = R - 1u;
    const int w = wave_id();
    const uint64_t lt = lanemask_lt();
    for (uint32_t j = threadIdx.x; j < RM * NW; j += kBlock) s_wc[j] = 0;
    s_base[threadIdx.x] = 0;   // kBlock == RM
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) atomicAdd(&s_base[((k[i] - kbase) >> shift) & mask], 1u);
    __syncthreads();
    {
        const uint32_t c = threadIdx.x < R ? s_base[threadIdx.x] : 0u;
        uint32_t tot;
        const uint32_t ex = block_excl_scan<NW>(c, &tot, s_scan);
        if (threadIdx.x < R) s_base[threadIdx.x] = ex;
        __syncthreads();
    }
    for (uint32_t r0 = 0; r0 < n; r0 += kBlock) {
        const uint32_t i = r0 + threadIdx.x;
        const bool ok = i < n;
        const uint32_t x = ok ? k[i] : 0u, y = ok ? v[i] : 0u;
        const uint32_t d = ((x - kbase) >> shift) & mask;
        const uint64_t peers = digit_peers(ok, d, db);
        const uint32_t rk = (uint32_t)__popcll(peers & lt);
        if (ok && rk == 0) s_wc[w * RM + d] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (ok) {
            uint32_t pos = s_base[d] + rk;
            for (int u = 0; u < w; ++u) pos += s_wc[u * RM + d];
            ko[pos] = x;
            vo[pos] = y;
        }
        __syncthreads();
        if (threadIdx.x < R) {
            uint32_t c = 0;
#pragma unroll
            for (int u = 0; u < NW; ++u) {
                c += s_wc[u * RM + threadIdx.x];
                s_wc[u * RM + threadIdx.x] = 0;
            }
            s_base[threadIdx.x] += c;
        }
        __syncthreads();
    }
}

// Tiles too large for LDS (more than kLeafStage pairs, or a local key range
// or top-pair count beyond the LDS sort's bit budget): one block per tile
// emits the tile's pairs in emission order.  With self_sort, the grid has no
// global leaf sort scheduled, so the block then sorts its accepted pairs
// itself (stable LSD passes over the tile-local key, ping-ponging with
// tk / tv) and writes their reference records -- correct but slow, for a
// rebuild that meets an oversize tile the grid's build never met.  Without
// self_sort the global pass sorts them (GridHeader::leaf_sort receives the
// number of accepted pairs).  Either way GridHeader::oversize counts the tile.
__global__ __launch_bounds__(kBlock) void k_leaf_emit_large(TriSoA t, GridHeader* h, BuildArgs a,
                                                            const uint32_t* __restrict__ top_keys,
                                                            const uint32_t* __restrict__ top_vals,
                                                            const uint32_t* __restrict__ cells,
                                                            uint32_t* __restrict__ keys,
                                                            uint32_t* __restrict__ vals,
                                                            const uint32_t* __restrict__ sums,
                                                            const uint32_t* __restrict__ info, uint32_t T,
                                                            int self_sort, uint32_t* __restrict__ tk,
                                                            uint32_t* __restrict__ tv, float* __restrict__ refs,
                                                            uint32_t* __restrict__ leaves) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_base[256], s_wc[256 * (kBlock / 64)];
    LeafTilePlan q;
    if (!leaf_tile_plan(a, sums, info, T, blockIdx.x, q) || q.staged) return;
    const LeafParams L = read_leaf_params(h);
    const uint32_t cap = a.cap_P2;
    const uint32_t rej_base = q.tot_acc + q.pre_rej;
    if (threadIdx.x == 0) {
        atomicAdd(&h->oversize, 1u);
        if (!self_sort) h->leaf_sort = q.tot_acc;
    }
    uint32_t run_acc = 0, run_rej = 0;
    for (uint32_t c0 = q.a0; c0 < q.a1; c0 += kBlock) {
        const uint32_t i = c0 + threadIdx.x;
        LeafSetup s;
        uint32_t tri = 0, c = 0, ca = 0;
        if (i < q.a1) {
            tri = top_vals[i];
            c = leaf_pair(t, L, h, top_keys[i], tri, cells, s);
            if (c) leaf_visit(L, s, [&](uint32_t, bool ok) { ca += ok ? 1u : 0u; });
        }
        uint32_t agg_a, agg_r;
        uint32_t pa = q.pre_acc + run_acc + block_excl_scan<kBlock / 64>(ca, &agg_a, s_scan);
        uint32_t pr = rej_base + run_rej + block_excl_scan<kBlock / 64>(c - ca, &agg_r, s_scan);
        if (c)
            leaf_visit(L, s, [&](uint32_t key, bool ok) {
                const uint32_t idx = ok ? pa++ : pr++;
                if (idx < cap) {
                    keys[idx] = ok ? key : L.C2;
                    vals[idx] = tri;
                }
            });
        run_acc += agg_a;
        run_rej += agg_r;
    }
    if (!self_sort) return;
    // the accepted pairs of this tile: [pre_acc, pre_acc + run_acc), keys in
    // [kbase, kbase + span) (leaf_tile_plan guarantees pre_acc + run_acc <= cap)
    __threadfence_block();
    __syncthreads();
    const uint32_t n = run_acc;
    uint32_t* k0 = keys + q.pre_acc;
    uint32_t* v0 = vals + q.pre_acc;
    uint32_t* k1 = tk + q.pre_acc;
    uint32_t* v1 = tv + q.pre_acc;
    const uint32_t span = q.span == 0xFFFFFFFFu ? 0xFFFFFFFFu : q.span;
    const int bits = 32 - __builtin_clz(span | 1u);
    const int passes = (bits + 7) / 8;
    const int db = (bits + passes - 1) / passes;
    for (int p = 0, lo = 0; p < passes; ++p, lo += db) {
        if (p & 1) block_sort_pass_global(k1, v1, k0, v0, n, q.kbase, lo, min(db, bits - lo), s_base, s_wc, s_scan);
        else block_sort_pass_global(k0, v0, k1, v1, n, q.kbase, lo, min(db, bits - lo), s_base, s_wc, s_scan);
        __threadfence_block();
        __syncthreads();
    }
    if (passes & 1) {
        for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
            k0[i] = k1[i];
            v0[i] = v1[i];
        }
    }
    for (uint32_t i = threadIdx.x; i < n; i += kBlock)
        write_ref_record(refs, q.pre_acc + i, t.rec, (passes & 1) ? v1[i] : v0[i], L.n_tris);
    if (!leaves) return;
    // the ranges of the tile's leaf cells from its sorted keys (as k_leaf_emit)
    __threadfence_block();
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += kBlock) {
        const uint32_t key = k0[i];
        if (i == 0 || k0[i - 1] != key) leaves[2 * (size_t)key] = q.pre_acc + i;
        if (i + 1 == n || k0[i + 1] != key) leaves[2 * (size_t)key + 1] = q.pre_acc + i + 1u;
    }
}

// Oversize tiles with one thread per top pair: pair i of tile t writes its
// accepted pairs from pre_acc(t) + (sacc[i] - sacc[first pair of t]) and its
// rejected ones likewise (sacc / srej: exclusive scans of the per-pair totals
// of k_leaf_count), so the output is k_leaf_emit_large's, in parallel -- a
// top cell of 500k pairs (a dense cluster) no longer runs on one block.
__global__ __launch_bounds__(kBlock) void k_leaf_emit_par(TriSoA t, GridHeader* h, BuildArgs a,
                                                          const uint32_t* __restrict__ top_keys,
                                                          const uint32_t* __restrict__ top_vals,
                                                          const uint32_t* __restrict__ ranges,
                                                          const uint32_t* __restrict__ cells,
                                                          uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                          const uint32_t* __restrict__ sums,
                                                          const uint32_t* __restrict__ info, uint32_t T,
                                                          const uint32_t* __restrict__ sacc,
                                                          const uint32_t* __restrict__ srej) {
    const LeafParams L = read_leaf_params(h);
    const uint32_t P1v = min(h->P1_valid, a.cap_P1);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < P1v; i += gridDim.x * kBlock) {
        const uint32_t key = top_keys[i];
        if (key >= L.C1) continue;
        const uint32_t tile = reinterpret_cast<const uint2*>(ranges)[key].x / kLeafTile;
        LeafTilePlan q;
        if (tile >= T || !leaf_tile_plan(a, sums, info, T, tile, q) || q.staged) continue;
        if (i == q.a0) {
            h->leaf_sort = q.tot_acc;
            atomicAdd(&h->oversize, 1u);
        }
        uint32_t pa = q.pre_acc + (sacc[i] - sacc[q.a0]);
        uint32_t pr = q.tot_acc + q.pre_rej + (srej[i] - srej[q.a0]);
        const uint32_t tri = top_vals[i];
        LeafSetup s;
        const uint32_t c = leaf_pair(t, L, h, key, tri, cells, s);
        if (c)
            leaf_visit(L, s, [&](uint32_t k2, bool ok) {
                const uint32_t idx = ok ? pa++ : pr++;
                if (idx < a.cap_P2) {
                    keys[idx] = ok ? k2 : L.C2;
                    vals[idx] = tri;
                }
            });
    }
}

// The global stable sort of the accepted leaf pairs works in two buffers;
// when its pass count is odd this copies its result back to the first one.
// n = GridHeader::leaf_sort, zero when every tile was sorted in LDS.
__global__ __launch_bounds__(kBlock) void k_copy_pairs(const uint32_t* __restrict__ ks,
                                                       const uint32_t* __restrict__ vs, uint32_t* __restrict__ kd,
                                                       uint32_t* __restrict__ vd, const uint32_t* d_n, uint32_t cap) {
    const uint32_t n = min(*d_n, cap);
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
        kd[i] = ks[i];
        vd[i] = vs[i];
    }
}

// ---------------------------------------------------------------- launchers
static uint32_t init_blocks(const ZeroSegs& segs) {
    size_t total = 0;
    for (int i = 0; i < segs.count; ++i) total += segs.bytes[i] >> 4;
    uint32_t blocks = (uint32_t)((total + kBlock - 1) / kBlock);
    return blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
}

static uint32_t ranges_blocks(uint32_t cap) {
    const uint32_t blocks = (cap + kBlock * 4 - 1) / (kBlock * 4);
    return blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
}

// one reference per thread (each record is a dependent gather)
static uint32_t refs_blocks(uint32_t cap) {
    const uint32_t blocks = (cap + kBlock - 1) / kBlock;
    return blocks < 1 ? 1 : (blocks > 32768 ? 32768 : blocks);
}

size_t bounds_blocks(uint32_t n) {
    const size_t b = ((size_t)n + kBlock * 4 - 1) / (kBlock * 4);   // 1024 triangles per block, four per lane
    return b < 1 ? 1 : (b > 4096 ? 4096 : b);
}

void launch_init_bounds(GridHeader* h, uint32_t n_tris, const ZeroSegs& segs, const TriSoA& t, uint32_t* part,
                        hipStream_t st) {
    const uint32_t nb = (uint32_t)bounds_blocks(n_tris);
    const uint32_t zb = init_blocks(segs);
    hipLaunchKernelGGL(k_init_bounds, dim3(nb > zb ? nb : zb), dim3(kBlock), 0, st, h, n_tris, segs, t, part, nb);
}

void launch_params(GridHeader* h, const BuildArgs& a, const uint32_t* part, uint32_t n, bool derive,
                   hipStream_t st) {
    hipLaunchKernelGGL(k_params, dim3(1), dim3(kBlock), 0, st, h, a, part, (uint32_t)bounds_blocks(n),
                       derive ? 1 : 0);
}

void launch_top_emit(const TriSoA& t, uint32_t n, GridHeader* h, const BuildArgs& a, uint32_t* keys,
                     uint32_t* vals, uint32_t* dbg, uint32_t* sums, hipStream_t st) {
    const uint32_t blocks = (n + kTile - 1) / kTile;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_top_count, dim3(blocks), dim3(kBlock), 0, st, t, n, h, sums, dbg);
    hipLaunchKernelGGL(k_top_emit, dim3(blocks), dim3(kBlock), 0, st, t, n, h, a, keys, vals, sums);
}

void launch_ranges(const uint32_t* keys, const uint32_t* d_n, uint32_t cap, const uint32_t* d_C,
                   uint32_t* cells, uint32_t* d_pvalid, hipStream_t st) {
    hipLaunchKernelGGL(k_ranges, dim3(ranges_blocks(cap)), dim3(kBlock), 0, st, keys, d_n, cap, d_C, cells,
                       d_pvalid);
}

void launch_leaf_cells(GridHeader* h, const BuildArgs& a, const uint32_t* ranges, uint32_t* cells,
                       uint32_t* dbg, uint32_t* sums, hipStream_t st) {
    const uint32_t blocks = (a.cap_C1 + kBlock - 1) / kBlock;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_leaf_cells_count, dim3(blocks), dim3(kBlock), 0, st, h, a, ranges, sums, dbg);
    hipLaunchKernelGGL(k_leaf_cells_write, dim3(blocks), dim3(kBlock), 0, st, h, a, ranges, cells, sums);
}

void launch_leaf_emit(const TriSoA& t, GridHeader* h, const BuildArgs& a, const uint32_t* top_keys,
                      const uint32_t* top_vals, const uint32_t* ranges, const uint32_t* cells, uint32_t* keys,
                      uint32_t* vals, uint32_t* tmp_keys, uint32_t* tmp_vals, float* refs, uint32_t* dbg,
                      uint32_t* sums, uint32_t* info, const LeafLarge* large, uint32_t* leaves, hipStream_t st) {
    const uint32_t T = (uint32_t)leaf_tiles(a.cap_P1);
    const uint32_t blocks = (a.cap_P1 + kBlock - 1) / kBlock;
    if (blocks)
        hipLaunchKernelGGL(k_leaf_count, dim3(blocks), dim3(kBlock), 0, st, t, h, top_keys, top_vals, ranges, cells,
                           sums, info, T, dbg, large ? large->pacc : nullptr, large ? large->prej : nullptr);
    hipLaunchKernelGGL(k_leaf_scan, dim3(1), dim3(1024), 0, st, h, a, sums, T);
    // with `large`, the global sort and k_ranges need every key; without it
    // the tiles write the leaf ranges and the keys are kept for keep_debug
    uint32_t* lv = large ? nullptr : leaves;
    const int write_keys = (large || a.keep_debug) ? 1 : 0;
    hipLaunchKernelGGL(k_leaf_emit, dim3(T), dim3(kLeafBlock), 0, st, t, h, a, top_keys, top_vals, cells, keys, vals,
                       refs, sums, info, T, lv, write_keys);
    if (large) {
        launch_scan(large->pacc, large->sacc, a.cap_P1, false, large->scan_scratch, st);
        launch_scan(large->prej, large->srej, a.cap_P1, false, large->scan_scratch, st);
        uint32_t pb = (a.cap_P1 + kBlock - 1) / kBlock;
        pb = pb < 1 ? 1 : (pb > 4096 ? 4096 : pb);
        hipLaunchKernelGGL(k_leaf_emit_par, dim3(pb), dim3(kBlock), 0, st, t, h, a, top_keys, top_vals, ranges, cells,
                           keys, vals, sums, info, T, large->sacc, large->srej);
    } else {
        hipLaunchKernelGGL(k_leaf_emit_large, dim3(T), dim3(kBlock), 0, st, t, h, a, top_keys, top_vals, cells, keys,
                           vals, sums, info, T, 1, tmp_keys, tmp_vals, refs, lv);
    }
}

void launch_leaf_sort_fixup(uint32_t* k0, uint32_t* v0, uint32_t* k1, uint32_t* v1, GridHeader* h, uint32_t cap,
                            int key_bits, int radix_bits, uint32_t* scratch, hipStream_t st) {
    launch_sort_pairs(k0, v0, k1, v1, &h->leaf_sort, cap, key_bits, radix_bits, scratch, st);
    if (sort_passes(key_bits, radix_bits) & 1) {
        uint32_t blocks = (cap + kBlock * 4 - 1) / (kBlock * 4);
        blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
        hipLaunchKernelGGL(k_copy_pairs, dim3(blocks), dim3(kBlock), 0, st, k1, v1, k0, v0, &h->leaf_sort, cap);
    }
}

// d_gate (optional): nothing to do while *d_gate == 0 (two-level grids: the
// LDS-sorted leaf tiles wrote their records in k_leaf_emit; only a global
// leaf sort, GridHeader::leaf_sort > 0, leaves them to this pass)
SG_D void ref_records_body(const uint32_t* __restrict__ indices, const uint32_t* d_valid, const uint32_t* d_gate,
                           uint32_t cap, const float* __restrict__ rec, uint32_t n_tris, float* __restrict__ refs,
                           uint32_t bid, uint32_t nb) {
    if (d_gate && *d_gate == 0) return;
    const uint32_t n = min(*d_valid, cap);
    for (uint32_t i = bid * kBlock + threadIdx.x; i < n; i += nb * kBlock)
        write_ref_record(refs, i, rec, indices[i], n_tris);
}

__global__ __launch_bounds__(kBlock) void k_ref_records(const uint32_t* __restrict__ indices, const uint32_t* d_valid,
                                                        const uint32_t* d_gate, uint32_t cap,
                                                        const float* __restrict__ rec, uint32_t n_tris,
                                                        float* __restrict__ refs) {
    ref_records_body(indices, d_valid, d_gate, cap, rec, n_tris, refs, blockIdx.x, gridDim.x);
}

void launch_ref_records(const uint32_t* indices, const uint32_t* d_valid, const uint32_t* d_gate, uint32_t cap,
                        const float* rec, uint32_t n_tris, float* refs, hipStream_t st) {
    hipLaunchKernelGGL(k_ref_records, dim3(refs_blocks(cap)), dim3(kBlock), 0, st, indices, d_valid, d_gate, cap,
                       rec, n_tris, refs);
}

// ------------------------------------------------ many uniform grids, one launch
// k_multi_*: one launch of a phase for every job of an entry table; each
// block runs the phase's body for its job (see UniJob)
// zeroing and bounds of every job in one launch, as k_init_bounds
__global__ __launch_bounds__(kBlock) void k_multi_init(const UniJob* __restrict__ jobs,
                                                       const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    init_body(j.h, j.n, j.z, bid, nb);
    if (bid < j.nb_bounds) {
        const TriSoA t = j.t;
        bounds_body(t, j.n, j.part, bid, j.nb_bounds);
    }
}

__global__ __launch_bounds__(kBlock) void k_multi_params(const UniJob* __restrict__ jobs,
                                                         const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const BuildArgs a = j.a;
    params_body(j.h, a, j.part, j.nb_bounds, 1);
}

__global__ __launch_bounds__(kBlock) void k_multi_top_count(const UniJob* __restrict__ jobs,
                                                            const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const TriSoA t = j.t;
    top_count_body(t, j.n, j.h, j.sums, j.dbg, bid);
}

__global__ __launch_bounds__(kBlock) void k_multi_top_emit(const UniJob* __restrict__ jobs,
                                                           const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const TriSoA t = j.t;
    const BuildArgs a = j.a;
    top_emit_body(t, j.n, j.h, a, j.k[0], j.v[0], j.sums, bid, nb);
}

__global__ __launch_bounds__(kBlock) void k_multi_ranges(const UniJob* __restrict__ jobs,
                                                         const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const int f = j.passes & 1;
    ranges_body(j.k[f], &j.h->P1, j.cap_P, &j.h->C1, j.ranges, &j.h->P1_valid, bid, nb);
}

__global__ __launch_bounds__(kBlock) void k_multi_refs(const UniJob* __restrict__ jobs,
                                                       const MultiEntry* __restrict__ e, uint32_t ne) {
    uint32_t bid, nb;
    const UniJob& j = jobs[multi_entry(e, ne, blockIdx.x, gridDim.x, bid, nb)];
    const int f = j.passes & 1;
    ref_records_body(j.v[f], &j.h->P1_valid, nullptr, j.cap_P, j.t.rec, j.n, j.refs, bid, nb);
}

uint32_t multi_blocks(const UniJob& j, MultiPhase ph) {
    switch (ph) {
        case MP_INIT: return init_blocks(j.z) > j.nb_bounds ? init_blocks(j.z) : j.nb_bounds;
        case MP_PARAMS: return 1;
        case MP_TOP_COUNT:
        case MP_TOP_EMIT: return (j.n + kTile - 1) / kTile;
        case MP_RANGES: return ranges_blocks(j.cap_P);
        case MP_REFS: return refs_blocks(j.cap_P);
        default: return 0;
    }
}

void launch_multi(MultiPhase ph, const UniJob* jobs, const MultiEntry* e, uint32_t ne, uint32_t blocks,
                  hipStream_t st) {
    if (ne == 0 || blocks == 0) return;
    const dim3 g(blocks), b(kBlock);
    switch (ph) {
        case MP_INIT: hipLaunchKernelGGL(k_multi_init, g, b, 0, st, jobs, e, ne); break;
        case MP_PARAMS: hipLaunchKernelGGL(k_multi_params, g, b, 0, st, jobs, e, ne); break;
        case MP_TOP_COUNT: hipLaunchKernelGGL(k_multi_top_count, g, b, 0, st, jobs, e, ne); break;
        case MP_TOP_EMIT: hipLaunchKernelGGL(k_multi_top_emit, g, b, 0, st, jobs, e, ne); break;
        case MP_RANGES: hipLaunchKernelGGL(k_multi_ranges, g, b, 0, st, jobs, e, ne); break;
        case MP_REFS: hipLaunchKernelGGL(k_multi_refs, g, b, 0, st, jobs, e, ne); break;
        default: break;
    }
}

// number of per-tile sums the emission kernels need for n items
size_t emit_tiles(uint32_t n) { return ((size_t)n + kTile - 1) / kTile; }
size_t cell_tiles(uint32_t n) { return ((size_t)n + kBlock - 1) / kBlock; }
// leaf tiles over n sorted top pairs
size_t leaf_tiles(uint32_t n) {
    const size_t t = ((size_t)n + kLeafTile - 1) / kLeafTile;
    return t ? t : 1;
}
size_t leaf_sums_words(uint32_t n) { return 2 * (size_t)leaf_sums_stride((uint32_t)leaf_tiles(n)) + 4; }
size_t leaf_info_words(uint32_t n) { return 4 * leaf_tiles(n); }


}  // namespace sg
