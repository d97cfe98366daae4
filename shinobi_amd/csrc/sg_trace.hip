// sg_trace.hip -- closest-hit 3D-DDA traversal of the uniform and two-level
// grids and the diffuse bounce generator, for gfx950.
//
// One ray per lane.  Rays are SoA (coalesced loads), grid cells / leaves
// are plain buffers (the reference's texture fetches replaced by explicit
// clamp-to-edge and zero-out-of-range reads), and reference i of a cell is
// tested from its 36-byte reference record (GridView::refs, written by the
// build in reference order): {v0, e1 = v1-v0, e2 = v2-v0}.  The edge
// subtractions are the binary32 operations the intersector would perform,
// so results are unchanged; a test is two 16-byte loads and one 4-byte load
// at an address known from the cell range alone (no index fetch between the
// range and the triangle).
#include <algorithm>

#include "sg_device.hpp"
#include "sg_kernels.hpp"

namespace sg {

// ---------------------------------------------------------- triangle records
__global__ __launch_bounds__(kBlock) void k_prep_tris(TriSoA t, uint32_t n) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    store_tri_record(t.rec, i, mk(t.v[0][i], t.v[1][i], t.v[2][i]), mk(t.v[3][i], t.v[4][i], t.v[5][i]),
                     mk(t.v[6][i], t.v[7][i], t.v[8][i]));
}

void launch_prep_tris(const TriSoA& t, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_prep_tris, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, t, n);
}

// ---------------------------------------------------------------- traversal
// The traversal reads the grid and the rays through views whose pointers are
// typed as global memory.  The multi-batch kernel reads its pointers from a
// table in memory, and the compiler treats such pointers as generic (flat
// loads, a 64-bit address per lane) unless their type says global; with
// these views every kernel gets global loads with a scalar base.
template <class T>
using gptr = __attribute__((address_space(1))) T*;

struct GView {
    gptr<const GridHeader> hdr;
    gptr<const uint32_t> cells, leaves, indices;
    gptr<const float> refs;
};

struct GRays {
    gptr<const float> a[7];
};

// 16- and 8-byte loads through global pointers (the HIP vector types cannot
// be copied from a global-typed reference, clang's vector types can)
typedef uint32_t u2v __attribute__((ext_vector_type(2)));
SG_D float4 ld4u(gptr<const float> p) {   // 16 bytes at 4-byte alignment
    const f4u v = *(gptr<const f4u>)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
SG_D uint2 ld2(gptr<const uint32_t> p) {
    const u2v v = *(gptr<const u2v>)p;
    return make_uint2(v.x, v.y);
}

SG_D GView gview(const GridView& g) {
    GView v;
    v.hdr = (gptr<const GridHeader>)g.hdr;
    v.cells = (gptr<const uint32_t>)g.cells;
    v.leaves = (gptr<const uint32_t>)g.leaves;
    v.indices = (gptr<const uint32_t>)g.indices;
    v.refs = (gptr<const float>)g.refs;
    return v;
}

SG_D GRays grays(const RaySoA& r) {
    GRays v;
    for (int k = 0; k < 7; ++k) v.a[k] = (gptr<const float>)r.a[k];
    return v;
}

struct Ray {
    v3 o, D, rcp;   // D = 1/rcp, the correctly rounded reciprocal of each rcp component
    float t;
    uint32_t best;
    uint32_t tests;
    bool fast;      // every rcp component is a div_rn divisor (D serves as its reciprocal)
};

// a / rcp per component, as the reference's divisions by rayDirRCP
SG_D v3 div_rcp(float a, const Ray& r) {
    return mk(div_rn(a, r.rcp.x, r.D.x, r.fast), div_rn(a, r.rcp.y, r.D.y, r.fast),
              div_rn(a, r.rcp.z, r.D.z, r.fast));
}

// SimpleIntersector::MollerTrumboreIntersectionTest<.., true>,
// src/Intersector/SimpleIntersector.hpp:50-89, for reference `it`: D is
// 1/rcp, the recomputed direction (:58); strict < keeps the first of
// equal-t candidates.
SG_D void mt_test(const float4 q0, const float4 q1, const float q2, uint32_t it, Ray& r) {
    const v3 v0 = mk(q0.x, q0.y, q0.z);
    const v3 e1 = mk(q0.w, q1.x, q1.y);
    const v3 e2 = mk(q1.z, q1.w, q2);
    const v3 pv = cross(r.D, e2);
    const float inv = rcp_rn(dot(e1, pv));
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

constexpr int kTopBatch = 16;       // lanes waiting for a top step that make the wave take it
constexpr uint32_t kTopWaitMax = 3; // iterations a lane waits for it at most

// The wave tests the references [b, e) of lane L's cell against lane L's ray:
// every active lane takes a strided share and the wave keeps the minimum
// (t, reference), which is what the ordered loop with strict < returns (the
// smallest distance, the first reference among equal ones).  Returns the
// key (distance bits << 32 | reference), ~0 when nothing closer than t hit.
__device__ __attribute__((always_inline)) uint64_t coop_tests(gptr<const float> refs, uint32_t lb, uint32_t le,
                                                         float ox, float oy, float oz, float Dx, float Dy, float Dz,
                                                         float t0, uint64_t act) {
    const uint32_t nact = (uint32_t)__popcll(act);
    const uint32_t rank = (uint32_t)__popcll(act & lanemask_lt());
    Ray r;
    r.o = mk(ox, oy, oz);
    r.D = mk(Dx, Dy, Dz);
    r.t = t0;
    r.best = kMiss;
    for (uint32_t it = lb + rank; it < le; it += nact) {
        const gptr<const float> q = refs + 9 * (size_t)it;
        mt_test(ld4u(q), ld4u(q + 4), q[8], it, r);
    }
    // positive distances order as unsigned integers: the smallest key is the
    // smallest t, then the first reference
    const uint64_t key = r.best == kMiss ? ~0ull : ((uint64_t)__float_as_uint(r.t) << 32) | r.best;
    uint64_t m = ~0ull;
    for (uint64_t a = act & __ballot(key != ~0ull); a; a &= a - 1) {   // lanes that found a closer hit
        const int j = __ffsll((long long)a) - 1;
        const uint64_t kj = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)(key >> 32), j) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)key, j);
        m = kj < m ? kj : m;
    }
    return m;
}

SG_D float lane_f(float x, int L) { return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), L)); }


// One DDA step along MIN_DIMENSION (src/Control/Algorithm/TwoLevelGridTraverser.h:633-634,
// SimpleGridTraverser.h:378-398): the axis d with the smallest tMax (first of
// x, y, z on ties, z when a comparison is unordered) advances cell[d] by
// sign(rcp[d]) (-1 for rcp <= 0 or NaN) and tMax[d] by size[d]*|rcp[d]|; the
// walk ends when cell[d] reaches res[d] or -1 (|difference| < 0.1).  The
// axis is kept as two lane masks and every update is a select, so the step
// has no branches; values are those of the per-axis scalar updates.
struct Step {
    bool x, y;     // d == 0, d == 1 (else d == 2)
};

SG_D Step min_axis(v3 t) {
    const bool xy = t.x < t.y;
    return Step{xy && t.x < t.z, !xy && t.y < t.z};
}

SG_D float sel(Step a, float x, float y, float z) { return a.x ? x : (a.y ? y : z); }

// returns false when the step left the grid (reference: cell == res or -1)
SG_D bool dda_step(Step a, v3& cell, v3& tMax, v3 rcp, v3 size, v3 resf) {
    const float rc = sel(a, rcp.x, rcp.y, rcp.z);
    const float inc = rc > 0.f ? 1.f : -1.f;
    const float dt = sel(a, size.x, size.y, size.z) * fabsf(rc);
    const bool z = !a.x && !a.y;
    cell.x = a.x ? cell.x + inc : cell.x;
    cell.y = a.y ? cell.y + inc : cell.y;
    cell.z = z ? cell.z + inc : cell.z;
    tMax.x = a.x ? tMax.x + dt : tMax.x;
    tMax.y = a.y ? tMax.y + dt : tMax.y;
    tMax.z = z ? tMax.z + dt : tMax.z;
    const float cd = sel(a, cell.x, cell.y, cell.z);
    const float rd = sel(a, resf.x, resf.y, resf.z);
    return !(fabsf(cd - rd) < 0.1f || fabsf(cd + 1.f) < 0.1f);
}

// tex3D clamp addressing on integral coordinates
SG_D uint32_t clamp_idx(float c, int res) { return (uint32_t)fminf(fmaxf(c, 0.f), (float)(res - 1)); }

struct GParams {
    v3 bmin, bmax, cs, crcp, resf;
    int res[3];
    uint32_t C2;
};

// header values are the same for every lane; readfirstlane keeps them in
// scalar registers where the compiler would otherwise load some of them
// per lane (it does so in the multi-batch kernel, whose header pointer
// comes from memory)
SG_D float uni(float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); }
SG_D v3 uni3(const float* a) { return mk(uni(a[0]), uni(a[1]), uni(a[2])); }

SG_D GParams load_gparams(gptr<const GridHeader> h) {
    GParams p;
    float f[12];
    for (int k = 0; k < 3; ++k) {
        f[k] = h->bmin[k];
        f[3 + k] = h->bmax[k];
        f[6 + k] = h->cs[k];
        f[9 + k] = h->cell_rcp[k];
    }
    p.bmin = uni3(f);
    p.bmax = uni3(f + 3);
    p.cs = uni3(f + 6);
    p.crcp = uni3(f + 9);
    for (int k = 0; k < 3; ++k) p.res[k] = __builtin_amdgcn_readfirstlane(h->res[k]);
    p.resf = mk((float)p.res[0], (float)p.res[1], (float)p.res[2]);
    p.C2 = (uint32_t)__builtin_amdgcn_readfirstlane(h->C2);
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
        pt = r.o + div_rcp(tEntry + kEps, r);
    } else {
        pt = (tEntry >= 0.f) ? r.o + div_rcp(tEntry + kEps, r) : r.o;
    }
    cell = (pt - p.bmin) * p.crcp;
    cell = mk(floorf(cell.x), floorf(cell.y), floorf(cell.z));
    const v3 b = mk(r.rcp.x > 0.f ? 1.f : 0.f, r.rcp.y > 0.f ? 1.f : 0.f, r.rcp.z > 0.f ? 1.f : 0.f);
    tMax = (((cell + b) * p.cs + p.bmin) - r.o) * r.rcp;
    return (cell.x != (float)(r.rcp.x > 0.f ? p.res[0] : -1)) &&
           (cell.y != (float)(r.rcp.y > 0.f ? p.res[1] : -1)) &&
           (cell.z != (float)(r.rcp.z > 0.f ? p.res[2] : -1));
}

// occlusion threshold of the shadow traversers (TwoLevelGridTraverser.h:718,
// SimpleGridTraverser.h:436)
constexpr float kShadowT = 0.9999f;

// The top-level (uniform: only) walk state of one ray.
struct Walk {
    v3 cell, tMax;
    float tEntry;
    uint32_t guard;
    bool active;
};

// a ray crosses at most res.x+res.y+res.z cells of a level; the cap only
// ends walks that can never reach the exit test (NaN / inf rays, entry cell
// past the far boundary), which the reference would loop on forever
SG_D uint32_t step_cap(const GParams& p) { return 2u * (uint32_t)(p.res[0] + p.res[1] + p.res[2]) + 16u; }

// UniformGridTraverser::traverse, src/Control/Algorithm/SimpleGridTraverser.h:274-401;
// SHADOW: UniformGridShadowTraverser (:409-464) starts only if t >= 0.9999,
// stops once t < 0.9999 and otherwise walks to the grid exit
template <bool SHADOW>
SG_D void walk_init_uniform(const GParams& p, const Ray& r, Walk& w) {
    w.guard = 0;
    w.active = (SHADOW ? r.t >= kShadowT : r.t > 0.f) && traverse_init<false>(p, r, w.tEntry, w.cell, w.tMax);
}

// The walks below are flat loops: in every iteration each lane either
// tests one reference of its current cell or, with the cell done, takes one
// traversal step (t check, DDA step, next cell's range), and the wave keeps
// iterating while any lane is busy.  Per lane this is the reference's
// sequence of operations (nested cell and test loops); for the wave it
// bounds the iteration count by its busiest lane instead of summing, level
// by level, the longest loop of any lane, which is what a nested loop costs
// when the 64 rays of a wave sit in cells of different sizes.  Lanes whose
// current cell holds at least coop_threshold() references get the whole
// wave for it (coop_tests).
struct Refs {
    uint32_t it, e;    // references [it, e) of the current cell not yet tested
};

SG_D void refs_begin(const GView& g, Refs& q, uint2 rg, Ray& r) {
    q.it = rg.x;
    q.e = rg.y;
    r.tests += rg.y - rg.x;
}

// How many references a cell needs before the wave tests it together.  A
// wave lives as long as its longest ray, whose tests run one per iteration,
// each a memory latency; once most lanes are done, the idle lanes can test
// a remaining ray's cell in one round instead.  Sharing costs the other
// active lanes one round, so the threshold falls with the number of active
// lanes n: n/2 + 2, so 34 for a full wave and 2 for a lone ray (measured
// on the height field against n + 1, n/3 + 2, n/4 + 2 and a fixed 64).
SG_D uint32_t coop_threshold(bool active) { return (uint32_t)__popcll(__ballot(active)) / 2u + 2u; }

// The wave takes over the dense cells: every lane with at least `thr`
// references pending has them all tested by the wave (coop_tests).
SG_D void refs_coop(const GView& g, Refs& q, Ray& r, uint32_t thr) {
    uint64_t big = __ballot(q.e - q.it >= thr && q.it < q.e);
    if (big) {
        const uint64_t act = __ballot(true);
        while (big) {
            const int L = __ffsll((long long)big) - 1;
            big &= big - 1;
            const uint64_t m = coop_tests(g.refs, (uint32_t)__builtin_amdgcn_readlane((int)q.it, L),
                                          (uint32_t)__builtin_amdgcn_readlane((int)q.e, L), lane_f(r.o.x, L),
                                          lane_f(r.o.y, L), lane_f(r.o.z, L), lane_f(r.D.x, L), lane_f(r.D.y, L),
                                          lane_f(r.D.z, L), lane_f(r.t, L), act);
            if (lane_id() == L) {
                if (m != ~0ull) {
                    r.t = __uint_as_float((uint32_t)(m >> 32));
                    r.best = (uint32_t)m;
                }
                q.it = q.e;
            }
        }
    }
}

// the record of the reference a lane tests in this iteration
struct RefRec {
    float4 a, b;
    float c;
};

SG_D void ref_load(const GView& g, const Refs& q, RefRec& x) {
    const gptr<const float> rr = g.refs + 9 * (size_t)q.it;
    x.a = ld4u(rr);
    x.b = ld4u(rr + 4);
    x.c = rr[8];
}

// min of the three exit distances: r.t > tmin(m) is r.t > m.x || r.t > m.y
// || r.t > m.z, NaN components included (fminf returns the other operand
// for one NaN, NaN for three)
SG_D float tmin(v3 m) { return fminf(fminf(m.x, m.y), m.z); }

SG_D uint2 load_cell(const GView& g, const GParams& p, v3 c) {
    const uint32_t k = clamp_idx(c.x, p.res[0]) + clamp_idx(c.y, p.res[1]) * (uint32_t)p.res[0] +
                       clamp_idx(c.z, p.res[2]) * (uint32_t)p.res[0] * (uint32_t)p.res[1];
    return ld2(g.cells + 2 * (size_t)k);
}

// one iteration of the uniform walk for every lane of the wave
template <bool SHADOW>
SG_D void uniform_iter(const GView& g, const GParams& p, Ray& r, Walk& w, Refs& q, uint32_t cap, uint32_t& steps) {
    refs_coop(g, q, r, coop_threshold(w.active));
    const bool test = q.it < q.e;
    RefRec x;
    if (test) ref_load(g, q, x);
    // A lane whose cell ends in this iteration (its last test, or none)
    // takes the DDA step before the test: the step does not depend on the
    // test, and the next cell's range then loads while the record is in
    // flight instead of after it.  The t check that decides whether the
    // step stands uses the t after the test.
    const bool fin = w.active && q.it + 1u >= q.e;
    float tm = 0.f;
    bool stepped = false;
    uint2 nr = make_uint2(0u, 0u);
    if (fin) {
        tm = tmin(w.tMax);
        stepped = dda_step(min_axis(w.tMax), w.cell, w.tMax, r.rcp, p.cs, p.resf) && w.guard++ < cap;
        if (stepped) nr = load_cell(g, p, w.cell);
    }
    if (test) {
        mt_test(x.a, x.b, x.c, q.it, r);
        ++q.it;
    }
    if (fin) {
        // the cell is done: t check, then the step taken above
        if (SHADOW ? !(r.t >= kShadowT) : !(r.t > tm)) {
            w.active = false;
        } else {
            ++steps;
            w.active = stepped;
            if (stepped) refs_begin(g, q, nr, r);
        }
    }
}

// the first cell of a walk: the guard counts it, its range starts the tests
SG_D void uniform_enter(const GView& g, const GParams& p, Ray& r, Walk& w, Refs& q, uint32_t cap) {
    if (w.active && !(w.guard++ < cap)) w.active = false;
    if (w.active) refs_begin(g, q, load_cell(g, p, w.cell), r);
}

template <bool SHADOW>
SG_D void walk_uniform(const GView& g, const GParams& p, Ray& r, Walk& w, uint32_t& steps) {
    const uint32_t cap = step_cap(p);
    Refs q;
    q.it = q.e = 0;
    uniform_enter(g, p, r, w, q, cap);
    while (__ballot(w.active)) uniform_iter<SHADOW>(g, p, r, w, q, cap, steps);
}

// TwoLevelGridTraverser::traverse, src/Control/Algorithm/TwoLevelGridTraverser.h:445-671;
// SHADOW: TwoLevelGridShadowTraverser (:677-828) does not start a ray with
// t < 0.9999, visits every leaf cell of an entered top cell and every top
// cell up to the grid exit until t < 0.9999 (no t > tMax tests).  It ignores
// a failed traverseInit and walks from an uninitialised cell; here such a
// ray is not traversed.
template <bool SHADOW>
SG_D void walk_init_twolevel(const GParams& p, const Ray& r, Walk& w) {
    w.guard = 0;
    w.active = SHADOW ? !(r.t < kShadowT) : r.t > 0.f;
    if (!traverse_init<true>(p, r, w.tEntry, w.cell, w.tMax)) w.active = false;
}

// the leaf walk of one top cell (TwoLevelGridTraverser.h:489-528 set-up)
SG_D bool take_step(uint32_t& left) {
    if (left == 0u) return false;
    --left;
    return true;
}

struct Leaf {
    v3 sc, stMax, scs, cres;    // cres: the leaf resolution as floats
    uint32_t base, left;        // left: steps the guard still admits
};

SG_D uint2 load_leaf(const GView& g, const GParams& p, const Leaf& L) {
    const float lf = L.sc.x + L.cres.x * L.sc.y + L.cres.x * L.cres.y * L.sc.z;
    const uint32_t li = L.base + f2u(lf);
    if (li >= p.C2) return make_uint2(0u, 0u);
    return ld2(g.leaves + 2 * (size_t)li);
}

SG_D void leaf_enter(const GParams& p, const GView& g, const Ray& r, const Walk& w, uint2 top, Leaf& L) {
    const v3 ep = r.o + div_rcp(w.tEntry, r);
    const uint32_t c0 = top.y & 0xFFu, c1 = (top.y >> 8) & 0xFFu, c2 = (top.y >> 16) & 0xFFu;
    L.base = top.x;
    L.cres = mk((float)c0, (float)c1, (float)c2);
    const v3 srcp = p.crcp * L.cres;
    const v3 mb = w.cell * p.cs + p.bmin;
    v3 sc = mk(floorf((ep.x - mb.x) * srcp.x), floorf((ep.y - mb.y) * srcp.y), floorf((ep.z - mb.z) * srcp.z));
    L.scs = mk(g.hdr->leaf_size[0][c0], g.hdr->leaf_size[1][c1], g.hdr->leaf_size[2][c2]);
    sc = vmin(L.cres - mk(1.f, 1.f, 1.f), vmax(sc, mk(0.f, 0.f, 0.f)));
    L.sc = sc;
    L.stMax.x = ((sc.x + (float)(r.rcp.x > 0.f ? 1 : 0)) * L.scs.x + mb.x - r.o.x) * r.rcp.x;
    L.stMax.y = ((sc.y + (float)(r.rcp.y > 0.f ? 1 : 0)) * L.scs.y + mb.y - r.o.y) * r.rcp.y;
    L.stMax.z = ((sc.z + (float)(r.rcp.z > 0.f ? 1 : 0)) * L.scs.z + mb.z - r.o.z) * r.rcp.z;
    // the guard admits 2 (c0 + c1 + c2) + 16 leaf cells, the entry cell one of them
    L.left = 2u * (c0 + c1 + c2) + 15u;
}

template <bool SHADOW>
SG_D void walk_twolevel(const GView& g, const GParams& p, Ray& r, Walk& w, uint32_t& lsteps, uint32_t& tsteps,
                        uint32_t& entered) {
    const uint32_t cap = step_cap(p);
    Refs q;
    q.it = q.e = 0;
    Leaf L;
    bool in_leaf = false;
    // enter the top cell at w.cell: a non-empty one starts its leaf walk
    auto top_cell = [&]() {
        if (!(w.guard++ < cap)) {
            w.active = false;
            return;
        }
        const uint2 top = load_cell(g, p, w.cell);
        if (top.y & kNotEmpty) {
            ++entered;
            leaf_enter(p, g, r, w, top, L);
            in_leaf = true;
            refs_begin(g, q, load_leaf(g, p, L), r);
        }
    };
    if (w.active) top_cell();
    // A lane done with its top cell waits (top_wait) until the wave takes its
    // top steps together: at least kTopBatch lanes waiting, a lane waiting
    // kTopWaitMax iterations, or no lane with anything else to do.  The top
    // step and the set-up of the next leaf walk are the longest code of the
    // loop; batching them runs that code in fewer iterations.
    bool top_wait = false;
    uint32_t waited = 0;
    while (__ballot(w.active)) {
        refs_coop(g, q, r, coop_threshold(w.active));
        const bool test = q.it < q.e;
        RefRec x;
        if (test) ref_load(g, q, x);
        // a lane whose leaf cell ends in this iteration takes the leaf DDA
        // step before its test, as in the uniform walk
        const bool fin = w.active && !top_wait && q.it + 1u >= q.e;
        float tm = 0.f;
        bool stepped = false;
        uint2 nr = make_uint2(0u, 0u);
        if (fin && in_leaf) {
            tm = tmin(L.stMax);
            stepped = dda_step(min_axis(L.stMax), L.sc, L.stMax, r.rcp, L.scs, L.cres) && take_step(L.left);
            if (stepped) nr = load_leaf(g, p, L);
        }
        if (test) {
            mt_test(x.a, x.b, x.c, q.it, r);
            ++q.it;
        }
        if (fin) {
            bool top_next = !in_leaf;
            if (in_leaf) {
                // a leaf cell is done: t check, then the leaf step taken above
                const bool go = SHADOW ? !(r.t < kShadowT) : r.t > tm;
                if (go) ++lsteps;
                if (go && stepped) {
                    refs_begin(g, q, nr, r);
                } else {
                    in_leaf = false;
                    top_next = true;
                }
            }
            top_wait = top_next;
        }
        const uint64_t waiting = __ballot(w.active && top_wait);
        if (waiting) {
            const bool go = __popcll(waiting) >= kTopBatch || !__ballot(w.active && !top_wait) ||
                            __ballot(w.active && top_wait && waited >= kTopWaitMax);
            if (w.active && top_wait) {
                if (!go) {
                    ++waited;
                } else {
                    top_wait = false;
                    waited = 0;
                    // the top cell is done: t check, top DDA step, next top cell
                    if (SHADOW ? r.t < kShadowT : !(r.t > w.tMax.x || r.t > w.tMax.y || r.t > w.tMax.z)) {
                        w.active = false;
                    } else {
                        ++tsteps;
                        const Step a = min_axis(w.tMax);
                        w.tEntry = sel(a, w.tMax.x, w.tMax.y, w.tMax.z);
                        w.active = dda_step(a, w.cell, w.tMax, r.rcp, p.cs, p.resf);
                        if (w.active) top_cell();
                    }
                }
            }
        }
    }
}

// A ray's start: rcp(dir) as the trace kernel computes it
// (src/Engine/RayTracingKernels.h:111-113) and the intersector's recomputed
// direction D = 1/rcp (src/Intersector/SimpleIntersector.hpp:58)
SG_D void ray_begin(Ray& r, v3 o, v3 d, float t) {
    r.o = o;
    r.rcp = mk(rcp_rn(d.x), rcp_rn(d.y), rcp_rn(d.z));
    r.D = mk(rcp_rn(r.rcp.x), rcp_rn(r.rcp.y), rcp_rn(r.rcp.z));
    r.fast = div_fast_range(r.rcp.x) && div_fast_range(r.rcp.y) && div_fast_range(r.rcp.z);
    r.t = t;
    r.best = kMiss;
    r.tests = 0;
}

// traverse (both traversers): leaves r.t and r.best (a reference index)
template <int KIND, bool SHADOW>
SG_D void ray_walk(const GView& g, const GParams& p, Ray& r, uint32_t& s1, uint32_t& s2, uint32_t& s3) {
    Walk w;
    if (KIND == 0) {
        walk_init_uniform<SHADOW>(p, r, w);
        walk_uniform<SHADOW>(g, p, r, w, s1);
    } else {
        walk_init_twolevel<SHADOW>(p, r, w);
        walk_twolevel<SHADOW>(g, p, r, w, s1, s2, s3);
    }
}

// a failed or overflowed build is never traversed: every ray misses
SG_D bool grid_failed(const GView& g) { return (g.hdr->status & ~(ST_LEAF_RES | ST_NEG_EXTENT)) != 0u; }

// the trace kernel of src/Engine/RayTracingKernels.h:37-136 for one SoA ray:
// rcp(dir) (:111-113), traverse, store prim = indices[best] (miss: ~0u) and t
// SPARSE: lanes without a ray (valid false) run the walk loop inactive, so
// that they take part in the wave-shared tests; otherwise every lane has one.
template <int KIND, bool STATS, bool SHADOW, bool SPARSE = false>
SG_D void trace_one(const GView& g, const GRays& rays, uint32_t i, gptr<uint32_t> prim, gptr<float> tout,
                    gptr<uint32_t> stats, bool valid = true) {
    if (grid_failed(g)) {
        if (!SPARSE || valid) {
            prim[i] = kMiss;
            tout[i] = rays.a[6][i];
            if (STATS)
                for (int k = 0; k < 4; ++k) stats[4 * (size_t)i + k] = 0;
        }
        return;
    }
    const GParams p = load_gparams(g.hdr);
    const uint32_t j = SPARSE && !valid ? 0u : i;   // a lane without a ray reads ray 0, inactive
    Ray r;
    ray_begin(r, mk(rays.a[0][j], rays.a[1][j], rays.a[2][j]), mk(rays.a[3][j], rays.a[4][j], rays.a[5][j]),
              SPARSE && !valid ? 0.f : rays.a[6][j]);   // t = 0: never traversed
    uint32_t s1 = 0, s2 = 0, s3 = 0;
    ray_walk<KIND, SHADOW>(g, p, r, s1, s2, s3);
    if (SPARSE && !valid) return;
    prim[i] = r.best == kMiss ? kMiss : g.indices[r.best];
    tout[i] = r.t;
    if (STATS) {
        stats[4 * (size_t)i + 0] = r.tests;
        stats[4 * (size_t)i + 1] = s1;
        stats[4 * (size_t)i + 2] = s2;
        stats[4 * (size_t)i + 3] = s3;
    }
}

// Occupancy of the trace kernels: the traversal is bound by memory latency
// and issue, and waves per SIMD measured 3 / 4 / 5 -> 1.56 / 1.80 / 1.99
// Grays/s on the height field.  With the default machine scheduler the
// two-level walk ran best at 6 waves in 80 VGPRs with a 12-byte spill
// (+1-2 % over 5 waves, 15 % slower at 7); with the max-ilp scheduler this
// unit is built with, 5 waves without spills run the height field's bounce
// rays 4 % faster than 6 (0.667 -> 0.641 ms) and the dense-cluster paths
// 3 % faster, the ruins scene's bounce rays 4 % slower; the camera launch
// is unchanged (its kernel fits 6 waves either way).  coop_tests is
// inlined: as a call it held 12 more VGPRs across the call.
#define SG_TRACE_OCCUPANCY __attribute__((amdgpu_waves_per_eu(5)))

// Rays per wave for a batch of n rays: 64, or fewer (down to 8) while the
// chip (`fill` waves at once) would not be full.  A small batch -- a late
// path segment, a small request -- lasts as long as its longest ray, and
// the lanes without a ray then help the others through the wave-shared
// tests.  The sparse form needs at most `fill` waves.
SG_HD uint32_t rays_per_wave(uint32_t n, uint32_t fill) {
    uint32_t R = 64;
    while (R > 8u && (uint64_t)n <= (uint64_t)(R / 2u) * fill) R /= 2u;
    return R;
}

// One batch, the block size a launch parameter.  SPARSE = false: one ray
// per lane, for batches of 64 rays per wave; SPARSE = true: rays_per_wave
// rays per wave.  With a device-side count both are launched and each
// exits unless the count is its case, so the dense kernel carries none of
// the sparse form's registers.
template <int KIND, bool STATS, bool SHADOW, bool SPARSE>
__global__ __launch_bounds__(kBlock) SG_TRACE_OCCUPANCY void k_trace(GridView g, RaySoA rays, uint32_t cap, const uint32_t* d_n,
                                                  uint32_t fill, uint32_t* __restrict__ prim,
                                                  float* __restrict__ tout, uint32_t* __restrict__ stats) {
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t R = rays_per_wave(n, fill);
    if (SPARSE != (R < 64u)) return;
    if (!SPARSE) {
        const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= n) return;
        trace_one<KIND, STATS, SHADOW>(gview(g), grays(rays), i, (gptr<uint32_t>)prim, (gptr<float>)tout,
                                       (gptr<uint32_t>)stats);
    } else {
        const uint32_t wv = blockIdx.x * (blockDim.x / 64u) + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
        if (wv * R >= n) return;   // the whole wave has no ray
        const uint32_t l = (uint32_t)lane_id(), i = wv * R + l;
        trace_one<KIND, STATS, SHADOW, true>(gview(g), grays(rays), i, (gptr<uint32_t>)prim, (gptr<float>)tout,
                                             (gptr<uint32_t>)stats, l < R && i < n);
    }
}

// waves the device holds at once (6 per SIMD, 4 SIMDs per CU), per device
static uint32_t device_fill() {
    static uint32_t cache[64] = {0};
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (dev < 0 || dev >= 64) return 6144u;
    if (!cache[dev]) {
        int cus = 0;
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        cache[dev] = (uint32_t)(cus > 0 ? cus : 256) * 24u;
    }
    return cache[dev];
}

void launch_trace(const GridView& g, const RaySoA& r, uint32_t n, const uint32_t* d_n, uint32_t* prim, float* t,
                  uint32_t* stats, bool shadow, int block, hipStream_t st) {
    if (n == 0) return;
    // one ray per lane keeps the 64 lanes of a wave on neighbouring rays, so
    // coherent rays share cell, leaf and triangle loads
    // one-wave blocks by default: the dispatcher then refills the chip wave
    // by wave as the short waves finish (height field 2.33 -> 2.36, ruins
    // 1.52 -> 1.61 Grays/s against 256-thread blocks)
    const int B = (block == 64 || block == 128 || block == 256) ? block : 64;
    const uint32_t fill = device_fill();
    // dense grid: one lane per ray of the capacity; sparse grid: `fill` waves
    // cover every count that takes the sparse form (n <= 32 fill, >= 8 rays
    // a wave)
    const dim3 dense((n + B - 1) / B), sparse((fill + B / 64 - 1) / (B / 64)), tb(B);
    const bool want_dense = d_n || rays_per_wave(n, fill) == 64u;
    const bool want_sparse = d_n || rays_per_wave(n, fill) < 64u;
#define SG_TRACE(K, S, H)                                                                                         \
    do {                                                                                                          \
        if (want_dense)                                                                                           \
            hipLaunchKernelGGL((k_trace<K, S, H, false>), dense, tb, 0, st, g, r, n, d_n, fill, prim, t, stats); \
        if (want_sparse)                                                                                          \
            hipLaunchKernelGGL((k_trace<K, S, H, true>), sparse, tb, 0, st, g, r, n, d_n, fill, prim, t, stats); \
    } while (0)
    if (g.kind == 0) {
        if (shadow) { if (stats) SG_TRACE(0, true, true); else SG_TRACE(0, false, true); }
        else { if (stats) SG_TRACE(0, true, false); else SG_TRACE(0, false, false); }
    } else {
        if (shadow) { if (stats) SG_TRACE(1, true, true); else SG_TRACE(1, false, true); }
        else { if (stats) SG_TRACE(1, true, false); else SG_TRACE(1, false, false); }
    }
#undef SG_TRACE
}

// ------------------------------------ camera rays and SimpleRayBuffer output
// RegularPrimaryRayGenerator<..>::operator() (src/Engine/RayGenerators.h:40-58)
// for pixel id `id`: RegularPixelSampler<1,1>::getPixelCoords with sample 0
// (src/Primitive/PixelSampler.h:37-52), then Camera::getDirection
// (src/Primitive/Camera.h:67-70), topLeft + x stepX + y stepY
SG_D v3 camera_dir(const CamArgs& c, uint32_t id) {
    float py = (truncf(0.f) + 0.5f) / 1.f;
    float px = 0.f - truncf(0.f) + 0.5f / 1.f;
    const float rx = (float)c.res_x;
    const float tmp = (float)id / rx;
    py += truncf(tmp);
    px += (tmp - truncf(tmp)) * rx;
    return (mk(c.tl[0], c.tl[1], c.tl[2]) + scl(mk(c.sx[0], c.sx[1], c.sx[2]), px)) +
           scl(mk(c.sy[0], c.sy[1], c.sy[2]), py);
}

// Camera launches write their outputs at pixel ids, so the waves of
// neighbouring tiles share output cache lines (2-pixel-wide tiles: 8 waves
// per 64-byte line of each output array).  Blocks are dispatched to the 8
// XCDs round-robin and each XCD has its own L2, so neighbouring blocks
// would each write back a partial line.  This remaps the block index so
// that 8 consecutive tiles run on one XCD: block b = 8 j + x runs on XCD x,
// and takes the (j mod 8)-th tile of group 8 (j / 8) + x.  Blocks beyond the
// last whole 64 keep their index.
constexpr uint32_t kXcds = 8;
SG_D uint32_t xcd_group_block(uint32_t b, uint32_t nb) {
    if (b >= (nb & ~(kXcds * kXcds - 1u))) return b;
    const uint32_t x = b % kXcds, j = b / kXcds;
    return ((j / kXcds) * kXcds + x) * kXcds + j % kXcds;
}

// Closest hit with the reference trace kernel's full contract: rays from SoA
// arrays (CAM = false, output slot = ray index) or generated from the camera
// (CAM = true, ray k of the launch is pixel tile_pixel(k), output slot = the
// pixel id); outputs prim / t and SimpleRayBuffer::store
// (src/Engine/RayBuffers.h:52-77) with the direction the trace kernel
// restores, 1 / rcp (src/Engine/RayTracingKernels.h:128-130), and bestHit =
// the reference index the traverser returns (0xFFFFFFFF on a miss).
template <int KIND, bool CAM>
__global__ __launch_bounds__(kBlock) SG_TRACE_OCCUPANCY void k_trace_gen(GridView g, RaySoA rays, CamArgs cam,
                                                                        uint32_t cap, const uint32_t* d_n,
                                                                        TraceOut out) {
    const uint32_t n = (!CAM && d_n) ? min(*d_n, cap) : cap;
    const uint32_t i = (CAM ? xcd_group_block(blockIdx.x, gridDim.x) : blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const GView gv = gview(g);
    const bool failed = grid_failed(gv);
    const GParams p = load_gparams(gv.hdr);
    uint32_t slot = i;
    Ray r;
    if (CAM) {
        slot = tile_pixel(cam, i);
        const v3 o = mk(cam.pos[0], cam.pos[1], cam.pos[2]);
        const v3 d = camera_dir(cam, slot);
        if (out.rays[0]) {
            const float v[7] = {o.x, o.y, o.z, d.x, d.y, d.z, 3.402823466e+38f};
            for (int k = 0; k < 7; ++k) out.rays[k][slot] = v[k];
        }
        ray_begin(r, o, d, 3.402823466e+38f);
    } else {
        const GRays gr = grays(rays);
        ray_begin(r, mk(gr.a[0][i], gr.a[1][i], gr.a[2][i]), mk(gr.a[3][i], gr.a[4][i], gr.a[5][i]), gr.a[6][i]);
    }
    if (!failed) {
        uint32_t s1 = 0, s2 = 0, s3 = 0;
        ray_walk<KIND, false>(gv, p, r, s1, s2, s3);
    }
    if (out.prim) out.prim[slot] = r.best == kMiss ? kMiss : gv.indices[r.best];
    if (out.t) out.t[slot] = r.t;
    if (out.raybuf) {
        // hitPoint = org + (t - EPS) * dir, then dir, t and bestHit, stride apart
        const float te = r.t - kEps;
        const size_t s = out.stride;
        float* b = out.raybuf + slot;
        b[0] = r.o.x + te * r.D.x;
        b[s] = r.o.y + te * r.D.y;
        b[2 * s] = r.o.z + te * r.D.z;
        b[3 * s] = r.D.x;
        b[4 * s] = r.D.y;
        b[5 * s] = r.D.z;
        b[6 * s] = r.t;
        b[7 * s] = __uint_as_float(r.best);
    }
}

void launch_trace_gen(const GridView& g, const RaySoA& r, const CamArgs* cam, uint32_t n, const uint32_t* d_n,
                      const TraceOut& out, hipStream_t st) {
    if (n == 0) return;
    constexpr int B = 64;   // one-wave blocks, as launch_trace
    const dim3 grid((n + B - 1) / B), tb(B);
    CamArgs c{};
    if (cam) c = *cam;
    if (g.kind == 0) {
        if (cam) hipLaunchKernelGGL((k_trace_gen<0, true>), grid, tb, 0, st, g, r, c, n, d_n, out);
        else hipLaunchKernelGGL((k_trace_gen<0, false>), grid, tb, 0, st, g, r, c, n, d_n, out);
    } else {
        if (cam) hipLaunchKernelGGL((k_trace_gen<1, true>), grid, tb, 0, st, g, r, c, n, d_n, out);
        else hipLaunchKernelGGL((k_trace_gen<1, false>), grid, tb, 0, st, g, r, c, n, d_n, out);
    }
}

// ------------------------------------------------- many batches, one launch
// Traversal of a table of (grid, ray batch) items in one launch: the
// batches are cut into 64-ray chunks numbered across the table, wave w of
// the grid traces chunk w against its item's grid (a binary search on the
// items' first chunks), so batches of any size share one launch and the
// hardware dispatcher balances the load as for a single batch.  A
// persistent form, in which each wave took the next chunk from a device
// counter, ran coherent batches 1.4x slower: consecutive chunks of a wave
// no longer neighbour those of the waves beside it.  One kernel per grid
// kind: a kernel that could run either walk would hold the registers of
// the larger one.
template <int KIND, bool SHADOW>
__global__ __launch_bounds__(kBlock) SG_TRACE_OCCUPANCY void k_trace_items(const TraceItem* __restrict__ items, uint32_t n_items,
                                                        uint32_t total_chunks) {
    const uint32_t c = blockIdx.x * (blockDim.x / 64u) + (uint32_t)__builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    if (c >= total_chunks) return;
    // Last item whose first chunk is <= c.  The table is read with
    // readfirstlane, so the compiler knows every value is wave-uniform and
    // keeps the item's pointers in scalar registers, as kernel arguments
    // are, instead of in vector registers.
    uint32_t lo = 0, hi = n_items;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if ((uint32_t)__builtin_amdgcn_readfirstlane(items[mid].chunk0) <= c) lo = mid;
        else hi = mid;
    }
    TraceItem it;
    {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(items + lo);
        uint32_t* dst = reinterpret_cast<uint32_t*>(&it);
#pragma unroll
        for (int k = 0; k < (int)(sizeof(TraceItem) / 4); ++k) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    }
    // the item's first wave reports the status of the grid it traced
    // against (a failed or overflowed build is traversed as all misses)
    if (it.status && c == it.chunk0 && lane_id() == 0) *it.status = it.g.hdr->status;
    const uint32_t i = (c - it.chunk0) * 64u + (uint32_t)lane_id();
    if (i < it.n) {
        // the table's pointers all address device memory
        trace_one<KIND, false, SHADOW>(gview(it.g), grays(it.rays), i, (gptr<uint32_t>)it.prim,
                                       (gptr<float>)it.t, nullptr);
    }
}

ItemSplit split_items(const TraceItem* in, uint32_t n, TraceItem* out) {
    ItemSplit s;
    for (uint32_t i = 0; i < n; ++i) s.n[in[i].g.kind == 0 ? 0 : 1]++;
    uint32_t at[2] = {0, s.n[0]};
    for (uint32_t i = 0; i < n; ++i) {
        const int k = in[i].g.kind == 0 ? 0 : 1;
        TraceItem it = in[i];
        it.chunk0 = s.chunks[k];
        s.chunks[k] += (it.n + 63u) / 64u;
        out[at[k]++] = it;
    }
    return s;
}

constexpr uint32_t kItemWaves = 1;   // waves per block of the multi-batch launch

void launch_trace_items(const TraceItem* items, const ItemSplit& split, bool shadow, hipStream_t st) {
    for (int k = 0; k < 2; ++k) {
        if (split.n[k] == 0 || split.chunks[k] == 0) continue;
        const TraceItem* part = items + (k ? split.n[0] : 0);
        const uint32_t blocks = (split.chunks[k] + kItemWaves - 1) / kItemWaves;
#define SG_ITEMS(K, H) \
    hipLaunchKernelGGL((k_trace_items<K, H>), dim3(blocks), dim3(64 * kItemWaves), 0, st, part, split.n[k], \
                       split.chunks[k])
        if (k == 0) { if (shadow) SG_ITEMS(0, true); else SG_ITEMS(0, false); }
        else { if (shadow) SG_ITEMS(1, true); else SG_ITEMS(1, false); }
#undef SG_ITEMS
    }
}

// ------------------------------------------------------------------ bounce
// One diffuse bounce per hit (src/Integrator/SimplePathTracer.hpp:115-166
// without Russian roulette): frame from the hit triangle's normalised edge1
// and n x edge1, n flipped against the ray, origin offset 0.001 n, cosine
// sample (src/Utils/HemisphereSamplers.hpp:54-72) from a counter hash of the
// primary ray id.
//
// The bounce rays are compacted and ordered by direction class: 4 bins per
// component of the new direction, 64 classes, and within a class in input
// order (rank r).  The classes take turns in rounds of kBounceRound rays:
// the output is sorted by (r / kBounceRound, class, r % kBounceRound).  A
// wave then traces rays of one class (shared cells and triangles: the
// height field's bounce trace went 1.40 -> 1.24 ms when grouped), and the
// costly classes (grazing directions) spread over the whole launch instead
// of arriving together (grouped 0.82 ms, rounds of 256 0.75 ms).  It is a
// counting sort fused into the compaction: per-tile class counts, a scan of
// each class over the tiles, then each tile ranks its rays with wave
// ballots (stable) and writes them at bounce_slot.
constexpr int kBounceTile = kBlock;
constexpr uint32_t kDirClasses = 64;
constexpr uint32_t kBounceRound = 256;

// output slot of the ray of class c with in-class rank r, given the class
// totals: all rays of earlier rounds, then this round's rays of the classes
// before c
SG_D uint32_t bounce_slot(const uint32_t* __restrict__ tot, uint32_t c, uint32_t r) {
    const uint32_t j0 = (r / kBounceRound) * kBounceRound;
    uint32_t s = r - j0;
    for (uint32_t k = 0; k < kDirClasses; ++k) {
        const uint32_t left = tot[k] > j0 ? tot[k] - j0 : 0u;
        s += min(tot[k], j0) + (k < c ? min(left, kBounceRound) : 0u);
    }
    return s;
}

SG_D void bounce_ray(const TriSoA& tris, const RaySoA& rays, uint32_t i, uint32_t tri, float t, uint64_t seed,
                     v3& no, v3& nd) {
    const v3 o = mk(rays.a[0][i], rays.a[1][i], rays.a[2][i]);
    const v3 d = mk(rays.a[3][i], rays.a[4][i], rays.a[5][i]);
    v3 v0, v1, v2;
    load_tri_record(tris.rec, tri, v0, v1, v2);
    v3 ta = v1 - v0, tb = v2 - v0;
    v3 nn = normalize(cross(ta, tb));
    ta = normalize(ta);
    tb = cross(nn, ta);
    if (dot(nn, d) > 0.f) nn = mk(-nn.x, -nn.y, -nn.z);
    no = (o + scl(d, t)) + scl(nn, 0.001f);
    const float u1 = u01(hash3(seed, i, 0)), u2 = u01(hash3(seed, i, 1));
    float sn, cn;
    sincos_2pi(u1, &sn, &cn);
    const float rr = sqrtf(1.f - u2);
    const v3 rd = mk(sn * rr, cn * rr, sqrtf(u2));
    nd = (scl(ta, rd.x) + scl(tb, rd.y)) + scl(nn, rd.z);
}

// direction class: bin floor(2 (c + 1)) of each component, clamped to 0..3
SG_HD uint32_t dir_class(v3 d) {
    auto bin = [](float c) -> uint32_t {
        const float f = (c + 1.f) * 2.f;
        return f < 1.f ? 0u : (f < 2.f ? 1u : (f < 3.f ? 2u : 3u));
    };
    return bin(d.x) | bin(d.y) << 2 | bin(d.z) << 4;
}

// counts[c * tiles + tile] = bounce rays of class c in the tile
__global__ __launch_bounds__(kBlock) void k_bounce_count(TriSoA tris, RaySoA rays, const uint32_t* __restrict__ prim,
                                                         const float* __restrict__ tin, uint32_t cap,
                                                         const uint32_t* d_n, uint64_t seed,
                                                         uint32_t* __restrict__ counts, uint32_t tiles) {
    __shared__ uint32_t s_c[kDirClasses];
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t i = blockIdx.x * kBounceTile + threadIdx.x;
    if (threadIdx.x < kDirClasses) s_c[threadIdx.x] = 0;
    __syncthreads();
    if (i < n && prim[i] != kMiss) {
        v3 no, nd;
        bounce_ray(tris, rays, i, prim[i], tin[i], seed, no, nd);
        atomicAdd(&s_c[dir_class(nd)], 1u);
    }
    __syncthreads();
    if (threadIdx.x < kDirClasses) counts[(size_t)threadIdx.x * tiles + blockIdx.x] = s_c[threadIdx.x];
}

// block c: exclusive scan of class c's tile counts in place, its total to
// totals[c]
__global__ __launch_bounds__(1024) void k_bounce_scan(uint32_t* counts, uint32_t tiles, uint32_t* totals) {
    __shared__ uint32_t s_scan[16];
    uint32_t* row = counts + (size_t)blockIdx.x * tiles;
    uint32_t run = 0;
    for (uint32_t c0 = 0; c0 < tiles; c0 += 1024 * 4) {
        const uint32_t t0 = c0 + threadIdx.x * 4;
        uint32_t v[4], sum = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            v[k] = t0 + k < tiles ? row[t0 + k] : 0u;
            sum += v[k];
        }
        uint32_t tot;
        uint32_t pre = run + block_excl_scan<16>(sum, &tot, s_scan);
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            if (t0 + k < tiles) row[t0 + k] = pre;
            pre += v[k];
        }
        run += tot;
    }
    if (threadIdx.x == 0) totals[blockIdx.x] = run;
}

__global__ __launch_bounds__(kBlock) void k_bounce(TriSoA tris, RaySoA rays, const uint32_t* __restrict__ prim,
                                                   const float* __restrict__ tin, uint32_t cap, const uint32_t* d_n,
                                                   uint64_t seed, RaySoAOut out, uint32_t* __restrict__ src,
                                                   const uint32_t* __restrict__ counts,
                                                   const uint32_t* __restrict__ totals, uint32_t tiles,
                                                   uint32_t* d_count) {
    constexpr int NW = kBlock / 64;
    __shared__ uint32_t s_base[kDirClasses], s_tot[kDirClasses], s_wc[NW][kDirClasses];
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t i = blockIdx.x * kBounceTile + threadIdx.x;
    if (threadIdx.x < kDirClasses) {
        s_base[threadIdx.x] = counts[(size_t)threadIdx.x * tiles + blockIdx.x];   // the tile's first rank
        s_tot[threadIdx.x] = totals[threadIdx.x];
        if (blockIdx.x == 0 && threadIdx.x == kDirClasses - 1) {
            uint32_t all = 0;
            for (uint32_t c = 0; c < kDirClasses; ++c) all += totals[c];
            *d_count = all;
        }
    }
    for (uint32_t k = threadIdx.x; k < NW * kDirClasses; k += kBlock) (&s_wc[0][0])[k] = 0;
    const bool hit = i < n && prim[i] != kMiss;
    v3 no = mk(0.f, 0.f, 0.f), nd = no;
    uint32_t c = 0;
    if (hit) {
        bounce_ray(tris, rays, i, prim[i], tin[i], seed, no, nd);
        c = dir_class(nd);
    }
    const uint64_t peers = digit_peers(hit, c, 6);
    const uint32_t rank = (uint32_t)__popcll(peers & lanemask_lt());
    __syncthreads();
    if (hit && rank == 0) s_wc[wave_id()][c] = (uint32_t)__popcll(peers);
    __syncthreads();
    if (!hit) return;
    uint32_t r = s_base[c] + rank;   // rank within the class
    for (int w = 0; w < wave_id(); ++w) r += s_wc[w][c];
    const uint32_t m = bounce_slot(s_tot, c, r);
    out.a[0][m] = no.x;
    out.a[1][m] = no.y;
    out.a[2][m] = no.z;
    out.a[3][m] = nd.x;
    out.a[4][m] = nd.y;
    out.a[5][m] = nd.z;
    out.a[6][m] = 3.402823466e+38f;
    if (src) src[m] = i;
}

__global__ void k_set_u32(uint32_t* p, uint32_t v);

// ------------------------------------------------------ ordered bounce rays
// Bounce rays in an order given by a per-ray 16-bit key: compaction in input
// order (k_bounce_hits / k_bounce_scan), the rays and their keys written at
// their compacted slots (k_bounce_keyed), a stable two-pass radix sort of
// (key, slot) pairs, then a gather into the output (k_bounce_gather).
// Keys (smaller first, ties in input order):
//   BOUNCE_EXIT_FIRST: 65535 - q, q = min(65535, (uint)(max(e, 0) * scale)),
//     e the exit distance of the ray from the box [bmin, bmax] (per axis the
//     far slab distance fmaxf((bmin - o) / d, (bmax - o) / d), the smallest
//     over x, y, z) and scale = 65535 / the box diagonal: longest first.
//   BOUNCE_AXIS_FIRST: min(65535, (uint)(|d . axis| * 65535)).
// A launch lasts about as long as its longest rays; starting them first (and
// keeping rays of similar length, hence similar paths, in one wave) shortens
// it.  Sixteen bits order the rays as well as the full distance does
// (measured on three scenes) in half the sort passes.
SG_D uint32_t quant16(float f) {   // NaN and f <= 0 -> 0
    if (!(f > 0.f)) return 0u;
    return f >= 65535.f ? 65535u : (uint32_t)f;
}

SG_D uint32_t bounce_key(const BounceOrder& q, v3 o, v3 d) {
    if (q.mode == BOUNCE_AXIS_FIRST)
        return quant16(fabsf(dot(d, mk(q.axis[0], q.axis[1], q.axis[2]))) * 65535.f);
    const float fx = fmaxf((q.bmin[0] - o.x) / d.x, (q.bmax[0] - o.x) / d.x);
    const float fy = fmaxf((q.bmin[1] - o.y) / d.y, (q.bmax[1] - o.y) / d.y);
    const float fz = fmaxf((q.bmin[2] - o.z) / d.z, (q.bmax[2] - o.z) / d.z);
    return 65535u - quant16(fmaxf(fminf(fminf(fx, fy), fz), 0.f) * q.scale);
}

// hits per tile: counts[tile] (one class)
__global__ __launch_bounds__(kBlock) void k_bounce_hits(const uint32_t* __restrict__ prim, uint32_t cap,
                                                        const uint32_t* d_n, uint32_t* __restrict__ counts) {
    __shared__ uint32_t s_scan[kBlock / 64];
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t i = blockIdx.x * kBounceTile + threadIdx.x;
    const uint32_t hit = (i < n && prim[i] != kMiss) ? 1u : 0u;
    const uint32_t c = block_sum<kBlock / 64>(hit, s_scan);
    if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// The compacted rays wait for the sort as one 32-byte record each,
// {o.x, o.y, o.z, d.x} {d.y, d.z, source index, 0}, so the gather reads one
// half cache line per ray instead of a line from each of seven arrays.
__global__ __launch_bounds__(kBlock) void k_bounce_keyed(TriSoA tris, RaySoA rays, const uint32_t* __restrict__ prim,
                                                         const float* __restrict__ tin, uint32_t cap,
                                                         const uint32_t* d_n, uint64_t seed, BounceOrder q,
                                                         float4* __restrict__ rec, uint32_t* __restrict__ keys,
                                                         uint32_t* __restrict__ vals,
                                                         const uint32_t* __restrict__ counts,
                                                         const uint32_t* __restrict__ totals, uint32_t* d_count) {
    __shared__ uint32_t s_scan[kBlock / 64];
    const uint32_t n = d_n ? min(*d_n, cap) : cap;
    const uint32_t i = blockIdx.x * kBounceTile + threadIdx.x;
    if (blockIdx.x == 0 && threadIdx.x == 0) *d_count = totals[0];
    const bool hit = i < n && prim[i] != kMiss;
    uint32_t agg;
    const uint32_t m = counts[blockIdx.x] + block_excl_scan<kBlock / 64>(hit ? 1u : 0u, &agg, s_scan);
    if (!hit) return;
    v3 no, nd;
    bounce_ray(tris, rays, i, prim[i], tin[i], seed, no, nd);
    rec[2 * (size_t)m] = make_float4(no.x, no.y, no.z, nd.x);
    rec[2 * (size_t)m + 1] = make_float4(nd.y, nd.z, __uint_as_float(i), 0.f);
    keys[m] = bounce_key(q, no, nd);
    vals[m] = m;
}

__global__ __launch_bounds__(kBlock) void k_bounce_gather(const float4* __restrict__ rec,
                                                          const uint32_t* __restrict__ vals, uint32_t cap,
                                                          const uint32_t* d_count, RaySoAOut out,
                                                          uint32_t* __restrict__ src) {
    const uint32_t n = min(*d_count, cap);
    for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < n; j += gridDim.x * kBlock) {
        const uint32_t m = vals[j];
        const float4 a = rec[2 * (size_t)m], b = rec[2 * (size_t)m + 1];
        out.a[0][j] = a.x;
        out.a[1][j] = a.y;
        out.a[2][j] = a.z;
        out.a[3][j] = a.w;
        out.a[4][j] = b.x;
        out.a[5][j] = b.y;
        out.a[6][j] = 3.402823466e+38f;
        if (src) src[j] = __float_as_uint(b.z);
    }
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

size_t bounce_ordered_scratch_bytes(uint32_t n) {
    const size_t tiles = ((size_t)n + kBounceTile - 1) / kBounceTile;
    const size_t w = align256((size_t)n * 4);
    return align256((tiles + 16) * 4) + 8 * w + 4 * w + sort_scratch_words(n, 16, 8) * 4 + 256;
}

void launch_bounce_ordered(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                           const uint32_t* d_n, uint64_t seed, const BounceOrder& q, const RaySoAOut& out,
                           uint32_t* src, uint32_t* d_count, void* scratch, hipStream_t st) {
    if (n == 0) {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, d_count, 0u);
        return;
    }
    const uint32_t tiles = (n + kBounceTile - 1) / kBounceTile;
    const size_t w = align256((size_t)n * 4);
    char* p = reinterpret_cast<char*>(scratch);
    uint32_t* counts = reinterpret_cast<uint32_t*>(p);
    uint32_t* totals = counts + tiles;
    p += align256(((size_t)tiles + 16) * 4);
    float4* rec = reinterpret_cast<float4*>(p);   // 32 bytes per ray: 8 words
    p += 8 * w;
    uint32_t* k0 = reinterpret_cast<uint32_t*>(p);
    uint32_t* v0 = reinterpret_cast<uint32_t*>(p + w);
    uint32_t* k1 = reinterpret_cast<uint32_t*>(p + 2 * w);
    uint32_t* v1 = reinterpret_cast<uint32_t*>(p + 3 * w);
    uint32_t* sort_scratch = reinterpret_cast<uint32_t*>(p + 4 * w);
    hipLaunchKernelGGL(k_bounce_hits, dim3(tiles), dim3(kBlock), 0, st, prim, n, d_n, counts);
    hipLaunchKernelGGL(k_bounce_scan, dim3(1), dim3(1024), 0, st, counts, tiles, totals);
    hipLaunchKernelGGL(k_bounce_keyed, dim3(tiles), dim3(kBlock), 0, st, tris, r, prim, t, n, d_n, seed, q, rec, k0,
                       v0, counts, totals, d_count);
    launch_sort_pairs(k0, v0, k1, v1, d_count, n, 16, 8, sort_scratch, st);
    const uint32_t* vs = (sort_passes(16, 8) & 1) ? v1 : v0;
    uint32_t gb = (n + kBlock - 1) / kBlock;
    gb = gb > 4096 ? 4096 : gb;
    hipLaunchKernelGGL(k_bounce_gather, dim3(gb), dim3(kBlock), 0, st, rec, vs, n, d_count, out, src);
}

__global__ void k_set_u32(uint32_t* p, uint32_t v) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

size_t bounce_scratch_bytes(uint32_t n) {
    const size_t tiles = ((size_t)n + kBounceTile - 1) / kBounceTile;
    return ((kDirClasses * (tiles + 1) + 4) * sizeof(uint32_t) + 15) & ~(size_t)15;
}

void launch_bounce(const TriSoA& tris, const RaySoA& r, const uint32_t* prim, const float* t, uint32_t n,
                   const uint32_t* d_n, uint64_t seed, const RaySoAOut& out, uint32_t* src, uint32_t* d_count,
                   void* scratch, hipStream_t st) {
    if (n == 0) {
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, d_count, 0u);
        return;
    }
    const uint32_t tiles = (n + kBounceTile - 1) / kBounceTile;
    uint32_t* counts = reinterpret_cast<uint32_t*>(scratch);
    uint32_t* totals = counts + (size_t)kDirClasses * tiles;
    hipLaunchKernelGGL(k_bounce_count, dim3(tiles), dim3(kBlock), 0, st, tris, r, prim, t, n, d_n, seed, counts,
                       tiles);
    hipLaunchKernelGGL(k_bounce_scan, dim3(kDirClasses), dim3(1024), 0, st, counts, tiles, totals);
    hipLaunchKernelGGL(k_bounce, dim3(tiles), dim3(kBlock), 0, st, tris, r, prim, t, n, d_n, seed, out, src, counts,
                       totals, tiles, d_count);
}

}  // namespace sg
