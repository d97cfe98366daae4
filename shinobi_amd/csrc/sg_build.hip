// sg_build.hip -- grid construction kernels for gfx950.
//
// The build is device-resident: bounds, resolution, counts and totals live in
// the GridHeader in HBM, every kernel sizes itself from it, and no phase
// reads anything back to the host.  Counting, scanning and emission of each
// level run as ONE single-pass kernel (per-tile block scan + decoupled
// look-back), so every pair lands at the slot that a serial scan of the
// counts in canonical order would give it: triangles (or sorted top pairs)
// ascending, then the reference's z/y/x loop order.
#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

// --------------------------------------------------------------- init
// Zero the look-back state, cell ranges and leaves; reset the header.
__global__ void k_init(GridHeader* h, uint32_t n_tris, ZeroSegs segs) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        for (int i = 0; i < 3; ++i) {
            h->bmin_key[i] = 0xFFFFFFFFu;
            h->bmax_key[i] = 0u;
        }
        h->n_tris = n_tris;
        h->C1 = h->P1 = h->P1_valid = h->C2 = h->P2 = h->P2_valid = 0;
        h->need_P1 = h->need_C2 = h->need_P2 = 0;
        h->status = 0;
        for (int i = 0; i < 32; ++i) h->tile_ctr[i] = 0;
    }
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (int s = 0; s < segs.count; ++s) {
        uint4* p = reinterpret_cast<uint4*>(segs.ptr[s]);
        const size_t n16 = segs.bytes[s] >> 4;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride)
            p[i] = make_uint4(0, 0, 0, 0);
    }
}

// --------------------------------------------------------------- bounds
// Scene AABB (the host loop of FastTwoLevelGridBuilder.h:75-81) as a device
// reduction; exact because min/max never round.
__global__ __launch_bounds__(kBlock) void k_bounds(TriSoA t, uint32_t n, GridHeader* h) {
    __shared__ uint32_t s_red[6][kBlock / 64];
    uint32_t k[6] = {0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0u, 0u, 0u};
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            uint32_t k0 = fkey(t.v[a][i]), k1 = fkey(t.v[3 + a][i]), k2 = fkey(t.v[6 + a][i]);
            k[a] = min(k[a], min(k0, min(k1, k2)));
            k[3 + a] = max(k[3 + a], max(k0, max(k1, k2)));
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
        if (a < 3) atomicMin(&h->bmin_key[a], r);
        else atomicMax(&h->bmax_key[a - 3], r);
    }
}

// Resolve bounds and grid parameters from the header into LDS (one thread),
// block 0 also records them in the header for later kernels and the host.
struct Params {
    v3 gmin, gmax, cs, rcp, crcp, resf;
    int32_t res[3];
    uint32_t rx, rxy, C;
    bool ok;
};

SG_D void load_params(GridHeader* h, const BuildArgs& a, Params* p) {
    if (threadIdx.x == 0) {
        float mn[3], mx[3], cs[3], rcp[3], crcp[3];
        int32_t res[3];
        for (int i = 0; i < 3; ++i) {
            mn[i] = fkey_inv(__hip_atomic_load(&h->bmin_key[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
            mx[i] = fkey_inv(__hip_atomic_load(&h->bmax_key[i], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT));
        }
        bool ok = derive_params(mn, mx, h->n_tris, a.lambda, a.top_divisor, a.fixed_res, res, cs, rcp, crcp);
        uint64_t C = ok ? (uint64_t)res[0] * (uint64_t)res[1] * (uint64_t)res[2] : 0;
        p->gmin = mk(mn[0], mn[1], mn[2]);
        p->gmax = mk(mx[0], mx[1], mx[2]);
        p->cs = mk(cs[0], cs[1], cs[2]);
        p->rcp = mk(rcp[0], rcp[1], rcp[2]);
        p->crcp = mk(crcp[0], crcp[1], crcp[2]);
        for (int i = 0; i < 3; ++i) p->res[i] = res[i];
        p->resf = mk((float)res[0], (float)res[1], (float)res[2]);
        p->rx = (uint32_t)res[0];
        p->rxy = f2u(p->resf.x * p->resf.y);
        p->C = (uint32_t)C;
        uint32_t bad = 0;
        if (!ok) bad |= ST_INVALID;
        else if (C >= (1ull << 24)) bad |= ST_TOO_MANY_CELLS;
        else if (C > a.cap_C1) bad |= ST_CAP_C1;
        p->ok = bad == 0;
        if (blockIdx.x == 0) {
            for (int i = 0; i < 3; ++i) {
                h->bmin[i] = mn[i];
                h->bmax[i] = mx[i];
                h->res[i] = ok ? res[i] : 0;
                h->cs[i] = cs[i];
                h->rcp_build[i] = rcp[i];
                h->cell_rcp[i] = crcp[i];
            }
            h->C1 = bad ? 0u : (uint32_t)C;
            if (bad) atomicOr(&h->status, bad);
        }
    }
    __syncthreads();
}

SG_D void tri_load(const TriSoA& t, uint32_t i, v3& a, v3& b, v3& c) {
    a = mk(t.v[0][i], t.v[1][i], t.v[2][i]);
    b = mk(t.v[3][i], t.v[4][i], t.v[5][i]);
    c = mk(t.v[6][i], t.v[7][i], t.v[8][i]);
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

// ------------------------------------------------------ top level emission
// countRefs + InclusiveScan + buildUnsortedGrid of
// src/Control/Algorithm/FastBuildKernels.h:31-215 fused into one pass.  One
// triangle per thread, one tile per block.
__global__ __launch_bounds__(kBlock) void k_top_emit(TriSoA t, uint32_t n, GridHeader* h, BuildArgs a,
                                                     uint32_t* __restrict__ keys,
                                                     uint32_t* __restrict__ vals,
                                                     uint32_t* __restrict__ dbg_counts, uint64_t* lb) {
    __shared__ Params P;
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_word;
    load_params(h, a, &P);
    const uint32_t tile = next_tile(&h->tile_ctr[0], &s_word);
    const uint32_t i = tile * kBlock + threadIdx.x;
    v3 v0 = mk(0, 0, 0), v1 = v0, v2 = v0, mn, mx;
    uint32_t lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    uint32_t cnt = 0;
    if (P.ok && i < n) {
        tri_load(t, i, v0, v1, v2);
        tri_bounds(v0, v1, v2, mn, mx);
        cnt = top_range(P, mn, mx, lo, hi);
        if (dbg_counts) dbg_counts[i] = cnt;
    }
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(cnt, &agg, s_scan);
    const uint32_t prefix = lb_tile_prefix(lb, tile, agg, &h->status, &s_word);
    const uint32_t total_tiles = (n + kBlock - 1) / kBlock;
    if (tile == total_tiles - 1 && threadIdx.x == 0) {
        uint32_t P1 = prefix + agg;
        h->need_P1 = P1;
        h->P1 = P1 <= a.cap_P1 ? P1 : a.cap_P1;
        if (P1 > a.cap_P1) atomicOr(&h->status, ST_CAP_P1);
    }
    if (cnt == 0) return;
    uint32_t slot = prefix + local;
    if ((uint64_t)slot + cnt > a.cap_P1) return;   // flagged above
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
                    uint32_t key;
                    if (fabsf(dv.x) <= half.x + kEps && fabsf(dv.y) <= half.y + kEps &&
                        fabsf(dv.z) <= half.z + kEps)
                        key = x + y * P.rx + z * P.rxy;
                    else
                        key = dummy;
                    keys[slot] = key;
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

// ---------------------------------------------------------------- ranges
// prepareCellRanges / prepareLeafCellRanges
// (src/Control/Algorithm/FastBuildKernels.h:218-313,
// src/Control/Algorithm/TwoLevelBuildKernels.h:458-536): a neighbour compare
// on the sorted keys writes {first, last+1} of every key below C; the dummy
// key's range is not stored.  Also records P_valid = index of the first
// dummy key.  Cells were zeroed by k_init.
__global__ __launch_bounds__(kBlock) void k_ranges(const uint32_t* __restrict__ keys, const uint32_t* d_n,
                                                   uint32_t cap, const uint32_t* d_C,
                                                   uint32_t* __restrict__ cells, uint32_t* d_pvalid) {
    const uint32_t n = min(*d_n, cap);
    const uint32_t C = *d_C;
    if (n == 0) {
        if (blockIdx.x == 0 && threadIdx.x == 0) *d_pvalid = 0;
        return;
    }
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock) {
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

// ------------------------------------------------------------ leaf cells
// countLeafLevelCells<false> + ExclusiveScan + prepareTopLevelCellRanges
// (src/Control/Algorithm/TwoLevelBuildKernels.h:40-127,
// src/Control/Algorithm/FastTwoLevelGridBuilder.h:256-294) in one pass over
// the top cells.
__global__ __launch_bounds__(kBlock) void k_leaf_cells(GridHeader* h, BuildArgs a,
                                                       const uint32_t* __restrict__ ranges,
                                                       uint32_t* __restrict__ cells,
                                                       uint32_t* __restrict__ dbg_counts, uint64_t* lb) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_word;
    const uint32_t C1 = h->C1;
    const uint32_t tile = next_tile(&h->tile_ctr[1], &s_word);
    const uint32_t base = tile * kBlock;
    if (base >= C1) return;
    const uint32_t k = base + threadIdx.x;
    const v3 cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    uint32_t cnt = 0, bits = 0;
    if (k < C1) {
        const uint32_t r = ranges[2 * (size_t)k + 1] - ranges[2 * (size_t)k];
        uint32_t rx = 1, ry = 1, rz = 1;
        if (r > a.leaf_min_refs) {
            const float vol = cs.x * cs.y * cs.z;
            const float m = cbrtpow(a.lambda_leaf * (float)r / vol);
            const v3 rf = scl(cs, m);
            rx = f2u(rf.x); ry = f2u(rf.y); rz = f2u(rf.z);
        }
        if (rx < 1 || ry < 1 || rz < 1 || rx > 255 || ry > 255 || rz > 255) atomicOr(&h->status, ST_LEAF_RES);
        cnt = rx * ry * rz;
        bits = (r == 0 ? 0u : kNotEmpty) | (rx & 0xFFu) | ((ry & 0xFFu) << 8) | ((rz & 0xFFu) << 16);
        if (dbg_counts) dbg_counts[k] = cnt;
    }
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(cnt, &agg, s_scan);
    const uint32_t prefix = lb_tile_prefix(lb, tile, agg, &h->status, &s_word);
    if (k < C1) {
        cells[2 * (size_t)k] = prefix + local;
        cells[2 * (size_t)k + 1] = bits;
    }
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
// countLeafLevelRefs + InclusiveScan + writeLeafLevelRefs
// (src/Control/Algorithm/TwoLevelBuildKernels.h:129-455) in one pass over the
// sorted top pairs; the slot of each leaf pair follows the sorted top-pair
// order, then the reference's z/y/x loop order.
struct LeafSetup {
    v3 v0, v1, v2, org, cres;
    int lo[3], hp[3];
    uint32_t px;
};

SG_D int leaf_range(const TriSoA& t, uint32_t key, uint32_t tri, const uint32_t* __restrict__ cells,
                    v3 gmin, v3 cs, uint32_t rx, uint32_t rxy, LeafSetup& s) {
    const uint32_t tx = key % rx;
    const uint32_t ty = (key % rxy) / rx;
    const uint32_t tz = key / rxy;
    v3 mn, mx;
    tri_load(t, tri, s.v0, s.v1, s.v2);
    tri_bounds(s.v0, s.v1, s.v2, mn, mx);
    const uint2 cell = *reinterpret_cast<const uint2*>(cells + 2 * (size_t)key);
    s.px = cell.x;
    s.cres = mk((float)(cell.y & 0xFFu), (float)((cell.y >> 8) & 0xFFu), (float)((cell.y >> 16) & 0xFFu));
    s.org = mk((float)tx * cs.x + gmin.x, (float)ty * cs.y + gmin.y, (float)tz * cs.z + gmin.z);
    const v3 srcp = s.cres / cs;
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

__global__ __launch_bounds__(kBlock) void k_leaf_emit(TriSoA t, GridHeader* h, BuildArgs a,
                                                      const uint32_t* __restrict__ top_keys,
                                                      const uint32_t* __restrict__ top_vals,
                                                      const uint32_t* __restrict__ cells,
                                                      uint32_t* __restrict__ keys, uint32_t* __restrict__ vals,
                                                      uint32_t* __restrict__ dbg_counts, uint64_t* lb) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_word;
    const uint32_t P1v = h->P1_valid;
    const uint32_t tile = next_tile(&h->tile_ctr[2], &s_word);
    const uint32_t base = tile * kBlock;
    if (base >= P1v) return;
    const uint32_t i = base + threadIdx.x;
    const v3 gmin = mk(h->bmin[0], h->bmin[1], h->bmin[2]);
    const v3 cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    const uint32_t rx = (uint32_t)h->res[0], rxy = (uint32_t)h->res[0] * (uint32_t)h->res[1];
    const uint32_t C2 = h->C2;
    LeafSetup s;
    uint32_t cnt = 0, tri = 0;
    if (i < P1v) {
        tri = top_vals[i];
        // a top level that overflowed its buffers may hold unwritten slots:
        // never follow an id outside the scene
        const uint32_t key = top_keys[i];
        int c = (tri < h->n_tris && key < h->C1) ? leaf_range(t, key, tri, cells, gmin, cs, rx, rxy, s) : 0;
        if (c < 0) {
            atomicOr(&h->status, ST_NEG_EXTENT);
            c = 0;
        }
        cnt = (uint32_t)c;
        if (dbg_counts) dbg_counts[i] = cnt;
    }
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(cnt, &agg, s_scan);
    const uint32_t prefix = lb_tile_prefix(lb, tile, agg, &h->status, &s_word);
    if (base + kBlock >= P1v && threadIdx.x == 0) {
        const uint32_t P2 = prefix + agg;
        h->need_P2 = P2;
        if (P2 > a.cap_P2) {
            atomicOr(&h->status, ST_CAP_P2);
            h->P2 = a.cap_P2;
        } else {
            h->P2 = P2;
        }
    }
    if (cnt == 0) return;
    uint32_t slot = prefix + local;
    if ((uint64_t)slot + cnt > a.cap_P2) return;
    const uint32_t crx = f2u(s.cres.x), crxy = f2u(s.cres.x * s.cres.y);
    if (s.hp[0] - s.lo[0] > 1 && s.hp[1] - s.lo[1] > 1 && s.hp[2] - s.lo[2] > 1) {
        const v3 n_ = normalize(cross(s.v1 - s.v0, s.v2 - s.v0));
        const v3 scs = cs / s.cres;
        const v3 c0 = (mk((float)s.lo[0], (float)s.lo[1], (float)s.lo[2]) * scs + s.org) + scl(scs, 0.5f);
        v3 cc;
        cc.z = c0.z - scs.z;
        for (int z = s.lo[2]; z < s.hp[2]; ++z) {
            cc.z += scs.z;
            cc.y = c0.y - scs.y;
            for (int y = s.lo[1]; y < s.hp[1]; ++y) {
                cc.y += scs.y;
                cc.x = c0.x - scs.x;
                for (int x = s.lo[0]; x < s.hp[0]; ++x, ++slot) {
                    cc.x += scs.x;
                    const v3 dv = scl(n_, dot(cc - s.v0, n_));
                    uint32_t key;
                    if (fabsf(dv.x) <= scs.x * 0.5f + kEps && fabsf(dv.y) <= scs.y * 0.5f + kEps &&
                        fabsf(dv.z) <= scs.z * 0.5f + kEps)
                        key = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + s.px;
                    else
                        key = C2;
                    keys[slot] = key;
                    vals[slot] = tri;
                }
            }
        }
    } else {
        for (int z = s.lo[2]; z < s.hp[2]; ++z)
            for (int y = s.lo[1]; y < s.hp[1]; ++y)
                for (int x = s.lo[0]; x < s.hp[0]; ++x, ++slot) {
                    keys[slot] = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + s.px;
                    vals[slot] = tri;
                }
    }
}

// ---------------------------------------------------------------- launchers
void launch_init(GridHeader* h, uint32_t n_tris, const ZeroSegs& segs, hipStream_t st) {
    size_t total = 0;
    for (int i = 0; i < segs.count; ++i) total += segs.bytes[i] >> 4;
    uint32_t blocks = (uint32_t)((total + kBlock - 1) / kBlock);
    blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
    hipLaunchKernelGGL(k_init, dim3(blocks), dim3(kBlock), 0, st, h, n_tris, segs);
}

void launch_bounds(const TriSoA& t, uint32_t n, GridHeader* h, hipStream_t st) {
    uint32_t blocks = (n + kBlock * 8 - 1) / (kBlock * 8);
    blocks = blocks < 1 ? 1 : (blocks > 2048 ? 2048 : blocks);
    hipLaunchKernelGGL(k_bounds, dim3(blocks), dim3(kBlock), 0, st, t, n, h);
}

void launch_top_emit(const TriSoA& t, uint32_t n, GridHeader* h, const BuildArgs& a, uint32_t* keys,
                     uint32_t* vals, uint32_t* dbg, uint64_t* lb, hipStream_t st) {
    const uint32_t blocks = (n + kBlock - 1) / kBlock;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_top_emit, dim3(blocks), dim3(kBlock), 0, st, t, n, h, a, keys, vals, dbg, lb);
}

void launch_ranges(const uint32_t* keys, const uint32_t* d_n, uint32_t cap, const uint32_t* d_C,
                   uint32_t* cells, uint32_t* d_pvalid, hipStream_t st) {
    uint32_t blocks = (cap + kBlock * 4 - 1) / (kBlock * 4);
    blocks = blocks < 1 ? 1 : (blocks > 4096 ? 4096 : blocks);
    hipLaunchKernelGGL(k_ranges, dim3(blocks), dim3(kBlock), 0, st, keys, d_n, cap, d_C, cells, d_pvalid);
}

void launch_leaf_cells(GridHeader* h, const BuildArgs& a, const uint32_t* ranges, uint32_t* cells,
                       uint32_t* dbg, uint64_t* lb, hipStream_t st) {
    const uint32_t blocks = (a.cap_C1 + kBlock - 1) / kBlock;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_leaf_cells, dim3(blocks), dim3(kBlock), 0, st, h, a, ranges, cells, dbg, lb);
}

void launch_leaf_emit(const TriSoA& t, GridHeader* h, const BuildArgs& a, const uint32_t* top_keys,
                      const uint32_t* top_vals, const uint32_t* cells, uint32_t* keys, uint32_t* vals,
                      uint32_t* dbg, uint64_t* lb, hipStream_t st) {
    const uint32_t blocks = (a.cap_P1 + kBlock - 1) / kBlock;
    if (blocks == 0) return;
    hipLaunchKernelGGL(k_leaf_emit, dim3(blocks), dim3(kBlock), 0, st, t, h, a, top_keys, top_vals, cells, keys,
                       vals, dbg, lb);
}

}  // namespace sg
