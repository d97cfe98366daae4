// sg_trace.hip -- closest-hit 3D-DDA traversal of the uniform and two-level
// grids and the diffuse bounce generator, for gfx950.
//
// One ray per lane.  Rays are SoA (coalesced loads), grid cells / leaves /
// indices are plain buffers (the reference's texture fetches replaced by
// explicit clamp-to-edge and zero-out-of-range reads), and each triangle is
// stored as a 48-byte traversal record {v0, e1 = v1-v0, e2 = v2-v0}: the
// edge subtractions are the same binary32 operations the intersector would
// perform, so results are unchanged and every Moller-Trumbore test needs two
// 16-byte loads plus one 4-byte load after the index fetch.
#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

// ---------------------------------------------------------- triangle records
__global__ __launch_bounds__(kBlock) void k_prep_tris(TriSoA t, uint32_t n, float4* __restrict__ rec) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const v3 v0 = mk(t.v[0][i], t.v[1][i], t.v[2][i]);
    const v3 v1 = mk(t.v[3][i], t.v[4][i], t.v[5][i]);
    const v3 v2 = mk(t.v[6][i], t.v[7][i], t.v[8][i]);
    const v3 e1 = v1 - v0, e2 = v2 - v0;
    rec[3 * (size_t)i + 0] = make_float4(v0.x, v0.y, v0.z, e1.x);
    rec[3 * (size_t)i + 1] = make_float4(e1.y, e1.z, e2.x, e2.y);
    rec[3 * (size_t)i + 2] = make_float4(e2.z, 0.f, 0.f, 0.f);
}

void launch_prep_tris(const TriSoA& t, uint32_t n, float4* rec, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_prep_tris, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, t, n, rec);
}

// ---------------------------------------------------------------- traversal
struct Ray {
    v3 o, D, rcp;
    float t;
    uint32_t best;
    uint32_t tests;
};

// SimpleIntersector::MollerTrumboreIntersectionTest<.., true>,
// src/Intersector/SimpleIntersector.hpp:50-89.  D is 1/rcp, the recomputed
// direction (:58); strict < keeps the first of equal-t candidates.
SG_D void intersect(const uint32_t* __restrict__ indices, const float4* __restrict__ rec, uint32_t b,
                    uint32_t e, Ray& r) {
    for (uint32_t it = b; it != e; ++it) {
        const uint32_t tri = indices[it];
        const float4 q0 = rec[3 * (size_t)tri];
        const float4 q1 = rec[3 * (size_t)tri + 1];
        const float q2 = reinterpret_cast<const float*>(rec + 3 * (size_t)tri + 2)[0];
        const v3 v0 = mk(q0.x, q0.y, q0.z);
        const v3 e1 = mk(q0.w, q1.x, q1.y);
        const v3 e2 = mk(q1.z, q1.w, q2);
        const v3 pv = cross(r.D, e2);
        const float inv = 1.f / dot(e1, pv);
        const v3 s = r.o - v0;
        const float alpha = inv * dot(s, pv);
        const v3 q = cross(s, e1);
        const float beta = inv * dot(q, r.D);
        const float dist = inv * dot(e2, q);
        if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > kEps && dist < r.t) {
            r.t = dist;
            r.best = it;
        }
    }
    r.tests += e - b;
}

// MIN_DIMENSION, src/Control/Algorithm/TwoLevelGridTraverser.h:633-634
SG_D int min_dim(v3 t) { return (t.x < t.y) ? ((t.x < t.z) ? 0 : 2) : ((t.y < t.z) ? 1 : 2); }

// tex3D clamp addressing on integral coordinates
SG_D uint32_t clamp_idx(float c, int res) { return (uint32_t)fminf(fmaxf(c, 0.f), (float)(res - 1)); }

struct GParams {
    v3 bmin, bmax, cs, crcp;
    int res[3];
    uint32_t C2;
};

SG_D GParams load_gparams(const GridHeader* h) {
    GParams p;
    p.bmin = mk(h->bmin[0], h->bmin[1], h->bmin[2]);
    p.bmax = mk(h->bmax[0], h->bmax[1], h->bmax[2]);
    p.cs = mk(h->cs[0], h->cs[1], h->cs[2]);
    p.crcp = mk(h->cell_rcp[0], h->cell_rcp[1], h->cell_rcp[2]);
    p.res[0] = h->res[0];
    p.res[1] = h->res[1];
    p.res[2] = h->res[2];
    p.C2 = h->C2;
    return p;
}

// traverseInit of both traversers: BBox::fastClip (src/Primitive/BBox.hpp:60-75)
// then the entry cell and tMax.  two_level clamps tEntry at 0 and always
// steps EPS inside (TwoLevelGridTraverser.h:586-626); the uniform grid steps
// only for tEntry >= 0 and uses '/' (SimpleGridTraverser.h:325-369).
template <bool TWO>
SG_D bool traverse_init(const GParams& p, const Ray& r, float& tEntry, v3& cell, v3& tMax) {
    const v3 t1 = (p.bmin - r.o) * r.rcp, t2 = (p.bmax - r.o) * r.rcp;
    const v3 tn = vmin(t1, t2), tx = vmax(t1, t2);
    tEntry = fmaxf(fmaxf(tn.x, tn.y), tn.z);
    const float tExit = fminf(fminf(tx.x, tx.y), tx.z);
    if (tExit <= tEntry || tExit < 0.f) return false;
    v3 pt;
    if (TWO) {
        tEntry = fmaxf(0.f, tEntry);
        const float te = tEntry + kEps;
        pt = r.o + mk(te, te, te) / r.rcp;
    } else {
        const float te = tEntry + kEps;
        pt = (tEntry >= 0.f) ? r.o + mk(te, te, te) / r.rcp : r.o;
    }
    cell = (pt - p.bmin) * p.crcp;
    cell = mk(floorf(cell.x), floorf(cell.y), floorf(cell.z));
    const v3 b = mk(r.rcp.x > 0.f ? 1.f : 0.f, r.rcp.y > 0.f ? 1.f : 0.f, r.rcp.z > 0.f ? 1.f : 0.f);
    tMax = (((cell + b) * p.cs + p.bmin) - r.o) * r.rcp;
    return (cell.x != (float)(r.rcp.x > 0.f ? p.res[0] : -1)) &&
           (cell.y != (float)(r.rcp.y > 0.f ? p.res[1] : -1)) &&
           (cell.z != (float)(r.rcp.z > 0.f ? p.res[2] : -1));
}

// UniformGridTraverser::traverse, src/Control/Algorithm/SimpleGridTraverser.h:274-401
SG_D void trace_uniform(const GridView& g, const GParams& p, Ray& r, uint32_t& steps) {
    float tEntry;
    v3 cell, tMax;
    bool flag = (r.t > 0.f) && traverse_init<false>(p, r, tEntry, cell, tMax);
    // a ray crosses at most res.x+res.y+res.z cells; the cap only ends walks
    // that can never reach the exit test (NaN / inf rays, entry cell past
    // the far boundary), which the reference would loop on forever
    const uint32_t cap = 2u * (uint32_t)(p.res[0] + p.res[1] + p.res[2]) + 16u;
    uint32_t guard = 0;
    while (flag && guard++ < cap) {
        const uint32_t k = clamp_idx(cell.x, p.res[0]) + clamp_idx(cell.y, p.res[1]) * (uint32_t)p.res[0] +
                           clamp_idx(cell.z, p.res[2]) * (uint32_t)p.res[0] * (uint32_t)p.res[1];
        const uint2 rg = *reinterpret_cast<const uint2*>(g.cells + 2 * (size_t)k);
        intersect(g.indices, g.rec, rg.x, rg.y, r);
        flag = r.t > tMax.x || r.t > tMax.y || r.t > tMax.z;
        if (!flag) break;
        ++steps;
        const int d = min_dim(tMax);
        const float rc = comp(r.rcp, d);
        setc(cell, d, comp(cell, d) + (rc > 0.f ? 1.f : -1.f));
        setc(tMax, d, comp(tMax, d) + comp(p.cs, d) * fabsf(rc));
        const float cd = comp(cell, d);
        flag = !(fabsf(cd - (float)p.res[d]) < 0.1f || fabsf(cd + 1.f) < 0.1f);
    }
}

// TwoLevelGridTraverser::traverse, src/Control/Algorithm/TwoLevelGridTraverser.h:445-671
SG_D void trace_twolevel(const GridView& g, const GParams& p, Ray& r, uint32_t& lsteps, uint32_t& tsteps,
                         uint32_t& entered) {
    float tEntry;
    v3 cell, tMax;
    bool flag1 = r.t > 0.f;
    if (!traverse_init<true>(p, r, tEntry, cell, tMax)) flag1 = false;
    const uint32_t cap = 2u * (uint32_t)(p.res[0] + p.res[1] + p.res[2]) + 16u;
    uint32_t guard = 0;
    while (flag1 && guard++ < cap) {
        const uint32_t k = clamp_idx(cell.x, p.res[0]) + clamp_idx(cell.y, p.res[1]) * (uint32_t)p.res[0] +
                           clamp_idx(cell.z, p.res[2]) * (uint32_t)p.res[0] * (uint32_t)p.res[1];
        const uint2 top = *reinterpret_cast<const uint2*>(g.cells + 2 * (size_t)k);
        if (top.y & kNotEmpty) {
            ++entered;
            const v3 ep = r.o + mk(tEntry, tEntry, tEntry) / r.rcp;
            const uint32_t c0 = top.y & 0xFFu, c1 = (top.y >> 8) & 0xFFu, c2 = (top.y >> 16) & 0xFFu;
            const v3 cres = mk((float)c0, (float)c1, (float)c2);
            const v3 srcp = p.crcp * cres;
            const v3 mb = cell * p.cs + p.bmin;
            v3 sc = mk(floorf((ep.x - mb.x) * srcp.x), floorf((ep.y - mb.y) * srcp.y), floorf((ep.z - mb.z) * srcp.z));
            const v3 scs = mk(1.f / srcp.x, 1.f / srcp.y, 1.f / srcp.z);
            sc = vmin(cres - mk(1.f, 1.f, 1.f), vmax(sc, mk(0.f, 0.f, 0.f)));
            v3 stMax;
            stMax.x = ((sc.x + (float)(r.rcp.x > 0.f ? 1 : 0)) * scs.x + mb.x - r.o.x) * r.rcp.x;
            stMax.y = ((sc.y + (float)(r.rcp.y > 0.f ? 1 : 0)) * scs.y + mb.y - r.o.y) * r.rcp.y;
            stMax.z = ((sc.z + (float)(r.rcp.z > 0.f ? 1 : 0)) * scs.z + mb.z - r.o.z) * r.rcp.z;
            bool flag2 = true;
            const uint32_t icap = 2u * (c0 + c1 + c2) + 16u;
            uint32_t iguard = 0;
            while (flag2 && iguard++ < icap) {
                const float lf = sc.x + (float)c0 * sc.y + (float)c0 * (float)c1 * sc.z;
                const uint32_t li = top.x + f2u(lf);
                uint2 rg = make_uint2(0u, 0u);
                if (li < p.C2) rg = *reinterpret_cast<const uint2*>(g.leaves + 2 * (size_t)li);
                intersect(g.indices, g.rec, rg.x, rg.y, r);
                flag2 = r.t > stMax.x || r.t > stMax.y || r.t > stMax.z;
                if (!flag2) break;
                ++lsteps;
                const int d = min_dim(stMax);
                const float rc = comp(r.rcp, d);
                setc(sc, d, comp(sc, d) + (rc > 0.f ? 1.f : -1.f));
                setc(stMax, d, comp(stMax, d) + comp(scs, d) * fabsf(rc));
                const float sd = comp(sc, d);
                flag2 = !(fabsf(sd - comp(cres, d)) < 0.1f || fabsf(sd + 1.f) < 0.1f);
            }
        }
        flag1 = r.t > tMax.x || r.t > tMax.y || r.t > tMax.z;
        if (!flag1) break;
        ++tsteps;
        const int d = min_dim(tMax);
        const float rc = comp(r.rcp, d);
        tEntry = comp(tMax, d);
        setc(cell, d, comp(cell, d) + (rc > 0.f ? 1.f : -1.f));
        setc(tMax, d, comp(tMax, d) + comp(p.cs, d) * fabsf(rc));
        const float cd = comp(cell, d);
        flag1 = !(fabsf(cd - (float)p.res[d]) < 0.1f || fabsf(cd + 1.f) < 0.1f);
    }
}

// the trace kernel of src/Engine/RayTracingKernels.h:37-136 for SoA rays:
// rcp(dir) (:111-113), traverse, store prim = indices[best] (miss: ~0u) and t
template <int KIND, bool STATS>
__global__ __launch_bounds__(kBlock) void k_trace(GridView g, RaySoA rays, uint32_t cap, const uint32_t* d_n,
                                                  uint32_t* __restrict__ prim, float* __restrict__ tout,
                                                  uint32_t* __restrict__ stats) {
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    // a failed or overflowed build is never traversed: every ray misses
    if (g.hdr->status & ~(ST_LEAF_RES | ST_NEG_EXTENT)) {
        prim[i] = kMiss;
        tout[i] = rays.a[6][i];
        if (STATS)
            for (int k = 0; k < 4; ++k) stats[4 * (size_t)i + k] = 0;
        return;
    }
    const GParams p = load_gparams(g.hdr);
    Ray r;
    r.o = mk(rays.a[0][i], rays.a[1][i], rays.a[2][i]);
    const v3 d = mk(rays.a[3][i], rays.a[4][i], rays.a[5][i]);
    r.rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    r.D = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z);
    r.t = rays.a[6][i];
    r.best = kMiss;
    r.tests = 0;
    uint32_t s1 = 0, s2 = 0, s3 = 0;
    if (KIND == 0) trace_uniform(g, p, r, s1);
    else trace_twolevel(g, p, r, s1, s2, s3);
    prim[i] = r.best == kMiss ? kMiss : g.indices[r.best];
    tout[i] = r.t;
    if (STATS) {
        stats[4 * (size_t)i + 0] = r.tests;
        stats[4 * (size_t)i + 1] = s1;
        stats[4 * (size_t)i + 2] = s2;
        stats[4 * (size_t)i + 3] = s3;
    }
}

void launch_trace(const GridView& g, const RaySoA& r, uint32_t n, const uint32_t* d_n, uint32_t* prim, float* t,
                  uint32_t* stats, hipStream_t st) {
    if (n == 0) return;
    const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    if (g.kind == 0) {
        if (stats) hipLaunchKernelGGL((k_trace<0, true>), grid, block, 0, st, g, r, n, d_n, prim, t, stats);
        else hipLaunchKernelGGL((k_trace<0, false>), grid, block, 0, st, g, r, n, d_n, prim, t, stats);
    } else {
        if (stats) hipLaunchKernelGGL((k_trace<1, true>), grid, block, 0, st, g, r, n, d_n, prim, t, stats);
        else hipLaunchKernelGGL((k_trace<1, false>), grid, block, 0, st, g, r, n, d_n, prim, t, stats);
    }
}

// ------------------------------------------------------------------ bounce
// One diffuse bounce per hit (src/Integrator/SimplePathTracer.hpp:115-166
// without Russian roulette): frame from the hit triangle's normalised edge1
// and n x edge1, n flipped against the ray, origin offset 0.001 n, cosine
// sample (src/Utils/HemisphereSamplers.hpp:54-72) from a counter hash of the
// primary ray id.  Compacted in primary order with a decoupled look-back.
__global__ __launch_bounds__(kBlock) void k_bounce(TriSoA tris, RaySoA rays, const uint32_t* __restrict__ prim,
                                                   const float* __restrict__ tin, uint32_t n, uint64_t seed,
                                                   RaySoAOut out, uint32_t* __restrict__ src, uint32_t* d_count,
                                                   uint64_t* scratch) {
    __shared__ uint32_t s_scan[kBlock / 64];
    __shared__ uint32_t s_word;
    uint32_t* ctr = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* status = ctr + 1;
    uint64_t* lb = scratch + 2;
    const uint32_t tile = next_tile(ctr, &s_word);
    const uint32_t i = tile * kBlock + threadIdx.x;
    const bool hit = i < n && prim[i] != kMiss;
    uint32_t agg;
    const uint32_t local = block_excl_scan<kBlock / 64>(hit ? 1u : 0u, &agg, s_scan);
    const uint32_t prefix = lb_tile_prefix(lb, tile, agg, status, &s_word);
    if (tile == (n + kBlock - 1) / kBlock - 1 && threadIdx.x == 0) *d_count = prefix + agg;
    if (!hit) return;
    const uint32_t m = prefix + local;
    const uint32_t tri = prim[i];
    const v3 o = mk(rays.a[0][i], rays.a[1][i], rays.a[2][i]);
    const v3 d = mk(rays.a[3][i], rays.a[4][i], rays.a[5][i]);
    const v3 v0 = mk(tris.v[0][tri], tris.v[1][tri], tris.v[2][tri]);
    const v3 v1 = mk(tris.v[3][tri], tris.v[4][tri], tris.v[5][tri]);
    const v3 v2 = mk(tris.v[6][tri], tris.v[7][tri], tris.v[8][tri]);
    v3 ta = v1 - v0, tb = v2 - v0;
    v3 nn = normalize(cross(ta, tb));
    ta = normalize(ta);
    tb = cross(nn, ta);
    if (dot(nn, d) > 0.f) nn = mk(-nn.x, -nn.y, -nn.z);
    const v3 no = (o + scl(d, tin[i])) + scl(nn, 0.001f);
    const float u1 = u01(hash3(seed, i, 0)), u2 = u01(hash3(seed, i, 1));
    float sn, cn;
    sincos_2pi(u1, &sn, &cn);
    const float rr = sqrtf(1.f - u2);
    const v3 rd = mk(sn * rr, cn * rr, sqrtf(u2));
    const v3 nd = (scl(ta, rd.x) + scl(tb, rd.y)) + scl(nn, rd.z);
    out.a[0][m] = no.x;
    out.a[1][m] = no.y;
    out.a[2][m] = no.z;
    out.a[3][m] = nd.x;
    out.a[4][m] = nd.y;
    out.a[5][m] = nd.z;
    out.a[6][m] = 3.402823466e+38f;
    if (src) src[m] = i;
}

__global__ void k_bounce_init(uint64_t* p, size_t n, uint32_t* d_count) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_count = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0;
}

size_t bounce_scratch_bytes(uint32_t n) {
    const size_t tiles = ((size_t)n + kBlock - 1) / kBlock;
    return ((tiles + 2) * sizeof(uint64_t) + 15) & ~(size_t)15;
}

void launch_bounce(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                   uint64_t seed, const RaySoAOut& out, uint32_t* src, uint32_t* d_count, void* scratch,
                   hipStream_t st) {
    const uint32_t tiles = (n + kBlock - 1) / kBlock;
    uint64_t* s = reinterpret_cast<uint64_t*>(scratch);
    hipLaunchKernelGGL(k_bounce_init, dim3((tiles + 2 + 255) / 256), dim3(256), 0, st, s, (size_t)tiles + 2,
                       d_count);
    if (n == 0) return;
    hipLaunchKernelGGL(k_bounce, dim3(tiles), dim3(kBlock), 0, st, tris, r, prim, t, n, seed, out, src, d_count, s);
}

}  // namespace sg
