/*
 * sg_oracle.c -- CPU oracle (test infrastructure only) for the Shinobi grid
 * acceleration hot path.  See sg_oracle.h for the contract.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference repository root).  Build: -O2 -ffp-contract=off, no
 * -ffast-math (oracle/Makefile); x86-64 SSE arithmetic, so binary32 ops round
 * exactly once.
 */
#define _GNU_SOURCE
#include "sg_oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define OR_EPS 0.00001f               /* src/CUDAStdAfx.h:96 */
#define NOTEMPTY 0x40000000u          /* src/Control/Structure/TwoLevelGrid.hpp:66 */

/* ------------------------------------------------------------------------- */
/* small helpers                                                             */
/* ------------------------------------------------------------------------- */

typedef struct { float x, y, z; } v3;

static inline v3 mk(float x, float y, float z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 mul(v3 a, v3 b) { return mk(a.x * b.x, a.y * b.y, a.z * b.z); }
static inline v3 divv(v3 a, v3 b) { return mk(a.x / b.x, a.y / b.y, a.z / b.z); }
static inline v3 scl(v3 a, float s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline v3 vmin(v3 a, v3 b) { return mk(fminf(a.x, b.x), fminf(a.y, b.y), fminf(a.z, b.z)); }
static inline v3 vmax(v3 a, v3 b) { return mk(fmaxf(a.x, b.x), fmaxf(a.y, b.y), fmaxf(a.z, b.z)); }
/* vec3f::dot, src/Core/Algebra.hpp:485-488 (left to right) */
static inline float dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
/* vec3f::cross, src/Core/Algebra.hpp:503-510 */
static inline v3 cross(v3 a, v3 b)
{
    return mk(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
/* operator~ = v * lenRCP(), lenRCP restated as 1/sqrtf (Algebra.hpp:570-577,606-609) */
static inline v3 normalize(v3 a) { float r = 1.0f / sqrtf(dot(a, a)); return scl(a, r); }
static inline float comp(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void setc(v3 *a, int i, float f) { if (i == 0) a->x = f; else if (i == 1) a->y = f; else a->z = f; }

/* float -> unsigned / signed conversion with the saturating semantics of the
 * GPU conversion instructions (NaN -> 0); identical to C casts in range. */
static inline uint32_t f2u(float f)
{
    if (!(f > 0.0f)) return 0u;
    if (f >= 4294967296.0f) return 0xFFFFFFFFu;
    return (uint32_t)f;
}
static inline int32_t f2i(float f)
{
    if (f != f) return 0;
    if (f >= 2147483648.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

static inline void tri_load(const float *const v[9], uint32_t i, v3 *a, v3 *b, v3 *c)
{
    *a = mk(v[0][i], v[1][i], v[2][i]);
    *b = mk(v[3][i], v[4][i], v[5][i]);
    *c = mk(v[6][i], v[7][i], v[8][i]);
}

/* BBox::empty + 3x extend, src/Primitive/BBox.hpp:78-82,109-119 */
static inline void tri_bounds(v3 a, v3 b, v3 c, v3 *mn, v3 *mx)
{
    v3 lo = mk(FLT_MAX, FLT_MAX, FLT_MAX), hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    lo = vmin(lo, a); hi = vmax(hi, a);
    lo = vmin(lo, b); hi = vmax(hi, b);
    lo = vmin(lo, c); hi = vmax(hi, c);
    *mn = lo; *mx = hi;
}

/* ------------------------------------------------------------------------- */
/* threads                                                                   */
/* ------------------------------------------------------------------------- */

typedef void (*range_fn)(void *ctx, size_t begin, size_t end, int tid);
typedef struct { range_fn fn; void *ctx; size_t b, e; int tid; } job_t;

static void *job_main(void *arg)
{
    job_t *j = (job_t *)arg;
    j->fn(j->ctx, j->b, j->e, j->tid);
    return NULL;
}

/* contiguous static split of [0,n) over nthreads; chunk t always gets the
 * same range, so results never depend on scheduling */
static void parallel_for(size_t n, int nthreads, range_fn fn, void *ctx)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads == 1 || n < 2048) { fn(ctx, 0, n, 0); return; }
    if ((size_t)nthreads > n) nthreads = (int)n;
    pthread_t th[256];
    job_t jobs[256];
    if (nthreads > 256) nthreads = 256;
    for (int t = 0; t < nthreads; ++t) {
        jobs[t].fn = fn; jobs[t].ctx = ctx; jobs[t].tid = t;
        jobs[t].b = n * (size_t)t / (size_t)nthreads;
        jobs[t].e = n * (size_t)(t + 1) / (size_t)nthreads;
    }
    for (int t = 1; t < nthreads; ++t) pthread_create(&th[t], NULL, job_main, &jobs[t]);
    job_main(&jobs[0]);
    for (int t = 1; t < nthreads; ++t) pthread_join(th[t], NULL);
}

/* ------------------------------------------------------------------------- */
/* deterministic math                                                        */
/* ------------------------------------------------------------------------- */

static inline uint64_t d2u(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static inline double u2d(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }

/* x^y for x > 0 from IEEE double +,-,*,/ only (atanh series for log, Taylor
 * for exp, exact exponent handling), so any IEEE machine computes the same
 * bits.  It stands in for powf(x, 0.3333333f) of
 * src/Control/Algorithm/GridBuilder.h:49-50 and
 * src/Control/Algorithm/TwoLevelBuildKernels.h:57-58. */
double or_det_pow(double x, double y)
{
    static const double atanh_c[14] = {
        0x1.0000000000000p+0, 0x1.5555555555555p-2, 0x1.999999999999ap-3, 0x1.2492492492492p-3,
        0x1.c71c71c71c71cp-4, 0x1.745d1745d1746p-4, 0x1.3b13b13b13b14p-4, 0x1.1111111111111p-4,
        0x1.e1e1e1e1e1e1ep-5, 0x1.af286bca1af28p-5, 0x1.8618618618618p-5, 0x1.642c8590b2164p-5,
        0x1.47ae147ae147bp-5, 0x1.2f684bda12f68p-5};
    static const double exp_c[18] = {
        0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
        0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
        0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
        0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33, 0x1.93974a8c07c9dp-37, 0x1.ae7f3e733b81fp-41,
        0x1.ae7f3e733b81fp-45, 0x1.952c77030ad4ap-49};
    const double ln2_hi = 0x1.62e42fef00000p-1, ln2_lo = 0x1.473de00000000p-34;
    const double inv_ln2 = 0x1.71547652b82fep+0;

    if (x != x || y != y) return NAN;
    if (x <= 0.0) return x == 0.0 ? 0.0 : NAN;
    if (x == INFINITY) return INFINITY;
    uint64_t b = d2u(x);
    int e = (int)((b >> 52) & 0x7ff);
    if (e == 0) { x = x * 0x1p54; b = d2u(x); e = (int)((b >> 52) & 0x7ff) - 54; }
    e -= 1023;
    double m = u2d((b & 0x000FFFFFFFFFFFFFull) | 0x3FF0000000000000ull);   /* [1,2) */
    if (m > 0x1.6a09e667f3bcdp+0) { m = m * 0.5; e += 1; }
    double z = (m - 1.0) / (m + 1.0);
    double z2 = z * z;
    double s = atanh_c[13];
    for (int k = 12; k >= 0; --k) s = s * z2 + atanh_c[k];
    double lnm = (2.0 * z) * s;
    double de = (double)e;
    double lnx = (de * ln2_hi + lnm) + de * ln2_lo;
    double w = y * lnx;
    if (w > 709.0) return INFINITY;
    if (w < -708.0) return 0.0;
    double n = floor(w * inv_ln2 + 0.5);
    double r = (w - n * ln2_hi) - n * ln2_lo;
    double p = exp_c[17];
    for (int k = 16; k >= 0; --k) p = p * r + exp_c[k];
    int ni = (int)n;
    return p * u2d((uint64_t)(ni + 1023) << 52);
}

float or_cbrtpow(float x) { return (float)or_det_pow((double)x, (double)0.3333333f); }

/* sin/cos of 2*pi*u from double Taylor series after an exact quarter-turn
 * reduction; stands in for sinCos(2*M_PI*u) in
 * src/Utils/HemisphereSamplers.hpp:64. */
void or_sincos_2pi(float u, float *s_out, float *c_out)
{
    static const double sc[10] = {
        0x1.0000000000000p+0, -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
        0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41,
        0x1.952c77030ad4ap-49, -0x1.2f49b46814157p-57};
    static const double cc[10] = {
        0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
        0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37,
        0x1.ae7f3e733b81fp-45, -0x1.6827863b97d97p-53};
    const double two_pi = 0x1.921fb54442d18p+2;
    double a = (double)u;
    double q = floor(a * 4.0 + 0.5);          /* exact */
    double f = a - q * 0.25;                  /* exact, |f| <= 1/8 turn */
    double th = f * two_pi;
    double t2 = th * th;
    double ps = sc[9], pc = cc[9];
    for (int k = 8; k >= 0; --k) { ps = ps * t2 + sc[k]; pc = pc * t2 + cc[k]; }
    double sn = th * ps, cs = pc;
    int qi = ((int)q) & 3;
    double so, co;
    if (qi == 0) { so = sn; co = cs; }
    else if (qi == 1) { so = cs; co = -sn; }
    else if (qi == 2) { so = -sn; co = -cs; }
    else { so = -cs; co = sn; }
    *s_out = (float)so;
    *c_out = (float)co;
}

/* SplitMix64 counter hash */
static inline uint64_t sm64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
uint64_t or_hash3(uint64_t seed, uint64_t a, uint64_t b) { return sm64(sm64(sm64(seed) + a) + b); }
float or_u01(uint64_t h) { return (float)(uint32_t)(h >> 40) * 0x1p-24f; }

/* ------------------------------------------------------------------------- */
/* scan / sort / ranges (src/Utils/Scan.h:35-68, src/Utils/Sort.h:33-48)     */
/* ------------------------------------------------------------------------- */

void or_exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n)
{
    uint32_t acc = 0;
    for (size_t i = 0; i < n; ++i) { uint32_t x = in[i]; out[i] = acc; acc += x; }
}

void or_inclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n)
{
    uint32_t acc = 0;
    for (size_t i = 0; i < n; ++i) { acc += in[i]; out[i] = acc; }
}

typedef struct {
    const uint32_t *ks, *vs; uint32_t *kd, *vd;
    int shift; size_t *hist; /* [nthreads][256] */ int nthreads;
} sort_ctx;

static void sort_hist(void *c_, size_t b, size_t e, int tid)
{
    sort_ctx *c = (sort_ctx *)c_;
    size_t *h = c->hist + (size_t)tid * 256;
    memset(h, 0, 256 * sizeof(size_t));
    for (size_t i = b; i < e; ++i) h[(c->ks[i] >> c->shift) & 255u]++;
}

static void sort_scatter(void *c_, size_t b, size_t e, int tid)
{
    sort_ctx *c = (sort_ctx *)c_;
    size_t *h = c->hist + (size_t)tid * 256;   /* holds start offsets now */
    for (size_t i = b; i < e; ++i) {
        uint32_t k = c->ks[i];
        size_t d = h[(k >> c->shift) & 255u]++;
        c->kd[d] = k;
        c->vd[d] = c->vs[i];
    }
}

/* Stable LSD radix sort by key (8-bit digits), the semantics of
 * chag::pp::sort as driven by src/Utils/Sort.cu:28-56. */
int or_sort_pairs(uint32_t *keys, uint32_t *vals, size_t n, int nthreads)
{
    if (n < 2) return OR_OK;
    if (nthreads < 1) nthreads = 1;
    if (n < 2048) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    uint32_t maxk = 0;
    for (size_t i = 0; i < n; ++i) if (keys[i] > maxk) maxk = keys[i];
    int bits = 0;
    while (bits < 32 && (maxk >> bits) != 0) ++bits;
    uint32_t *k2 = (uint32_t *)malloc(n * 4), *v2 = (uint32_t *)malloc(n * 4);
    size_t *hist = (size_t *)malloc((size_t)nthreads * 256 * sizeof(size_t));
    if (!k2 || !v2 || !hist) { free(k2); free(v2); free(hist); return OR_ERR_NOMEM; }
    uint32_t *ks = keys, *vs = vals, *kd = k2, *vd = v2;
    for (int shift = 0; shift < bits; shift += 8) {
        sort_ctx c = {ks, vs, kd, vd, shift, hist, nthreads};
        parallel_for(n, nthreads, sort_hist, &c);
        int nt = (nthreads == 1 || n < 2048) ? 1 : (nthreads > (int)n ? (int)n : nthreads);
        size_t acc = 0;
        for (int d = 0; d < 256; ++d)
            for (int t = 0; t < nt; ++t) { size_t x = hist[(size_t)t * 256 + d]; hist[(size_t)t * 256 + d] = acc; acc += x; }
        parallel_for(n, nthreads, sort_scatter, &c);
        uint32_t *t1 = ks; ks = kd; kd = t1;
        uint32_t *t2 = vs; vs = vd; vd = t2;
    }
    if (ks != keys) { memcpy(keys, ks, n * 4); memcpy(vals, vs, n * 4); }
    free(k2); free(v2); free(hist);
    return OR_OK;
}

/* prepareCellRanges / prepareLeafCellRanges semantics,
 * src/Control/Algorithm/FastBuildKernels.h:218-313 and
 * src/Control/Algorithm/TwoLevelBuildKernels.h:458-536: cell[k] = {first i
 * with key k, last i with key k + 1} for every k < ncells; the dummy key
 * (>= ncells) is not stored (its reference write is out of bounds). */
void or_ranges(const uint32_t *k, size_t n, uint32_t *cells, uint32_t ncells)
{
    memset(cells, 0, (size_t)ncells * 2 * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) {
        uint32_t key = k[i];
        if (key >= ncells) continue;
        if (i == 0 || k[i - 1] != key) cells[2 * (size_t)key] = (uint32_t)i;
        if (i + 1 == n || k[i + 1] != key) cells[2 * (size_t)key + 1] = (uint32_t)(i + 1);
    }
}

/* ------------------------------------------------------------------------- */
/* bounds and resolution                                                     */
/* ------------------------------------------------------------------------- */

/* host bounds loop, src/Control/Algorithm/FastTwoLevelGridBuilder.h:75-81 */
void or_bounds(const float *const v[9], uint32_t n, float bmin[3], float bmax[3])
{
    v3 lo = mk(FLT_MAX, FLT_MAX, FLT_MAX), hi = mk(-FLT_MAX, -FLT_MAX, -FLT_MAX);
    for (uint32_t i = 0; i < n; ++i) {
        v3 a, b, c;
        tri_load(v, i, &a, &b, &c);
        lo = vmin(lo, a); hi = vmax(hi, a);
        lo = vmin(lo, b); hi = vmax(hi, b);
        lo = vmin(lo, c); hi = vmax(hi, c);
    }
    bmin[0] = lo.x; bmin[1] = lo.y; bmin[2] = lo.z;
    bmax[0] = hi.x; bmax[1] = hi.y; bmax[2] = hi.z;
}

/* GridBuilder::computeResolution, src/Control/Algorithm/GridBuilder.h:41-62;
 * two-level top (divisor 6): FastTwoLevelGridBuilder.h:87-102. */
int or_resolution(const float bmin[3], const float bmax[3], uint32_t n,
                  float lambda, int divisor, int res[3])
{
    v3 d = mk(bmax[0] - bmin[0], bmax[1] - bmin[1], bmax[2] - bmin[2]);
    if (n == 0 || !(d.x > 0.0f) || !(d.y > 0.0f) || !(d.z > 0.0f)) return OR_ERR_INVALID;
    float volume = d.x * d.y * d.z;
    float m = or_cbrtpow(lambda * (float)n / volume);
    d = scl(d, m);
    res[0] = f2i(d.x); res[1] = f2i(d.y); res[2] = f2i(d.z);
    if (divisor > 1) {
        for (int i = 0; i < 3; ++i) { int r = res[i] / divisor; res[i] = r > 2 ? r : 2; }
    }
    if (res[0] < 1 || res[1] < 1 || res[2] < 1) return OR_ERR_INVALID;
    return OR_OK;
}

/* ------------------------------------------------------------------------- */
/* top level: count / emit / sort / ranges                                   */
/* ------------------------------------------------------------------------- */

typedef struct {
    const float *const *v;
    uint32_t n;
    v3 gmin, cs, rcp, resf;
    uint32_t rx, rxy, C;
    uint32_t *counts, *offs, *keys, *vals;
} top_ctx;

/* countRefs cell range, src/Control/Algorithm/FastBuildKernels.h:47-69 */
static inline uint32_t top_range(const top_ctx *c, v3 mn, v3 mx, uint32_t lo[3], uint32_t hi[3])
{
    v3 lof = vmax(mk(0.f, 0.f, 0.f), add(mul(sub(mn, c->gmin), c->rcp), mk(-OR_EPS, -OR_EPS, -OR_EPS)));
    v3 hif = vmin(sub(c->resf, mk(1.f, 1.f, 1.f)), add(mul(sub(mx, c->gmin), c->rcp), mk(OR_EPS, OR_EPS, OR_EPS)));
    lo[0] = f2u(lof.x); lo[1] = f2u(lof.y); lo[2] = f2u(lof.z);
    hi[0] = f2u(hif.x); hi[1] = f2u(hif.y); hi[2] = f2u(hif.z);
    /* uint wrap in the reference gives factor 0 for hi == lo-1; never emit
     * for an inverted range */
    if (hi[0] < lo[0] || hi[1] < lo[1] || hi[2] < lo[2]) return 0u;
    return (hi[0] - lo[0] + 1u) * (hi[1] - lo[1] + 1u) * (hi[2] - lo[2] + 1u);
}

static void top_count_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    top_ctx *c = (top_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        v3 a, bb, cc, mn, mx;
        uint32_t lo[3], hi[3];
        tri_load(c->v, (uint32_t)i, &a, &bb, &cc);
        tri_bounds(a, bb, cc, &mn, &mx);
        c->counts[i] = top_range(c, mn, mx, lo, hi);
    }
}

/* buildUnsortedGrid, src/Control/Algorithm/FastBuildKernels.h:88-215, with
 * the slot of triangle i = exclusive scan of the counts (canonical order) */
static void top_emit_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    top_ctx *c = (top_ctx *)c_;
    const uint32_t dummy = f2u(c->resf.x * c->resf.y * c->resf.z);
    for (size_t i = b; i < e; ++i) {
        v3 v0, v1, v2, mn, mx;
        uint32_t lo[3], hi[3];
        tri_load(c->v, (uint32_t)i, &v0, &v1, &v2);
        tri_bounds(v0, v1, v2, &mn, &mx);
        uint32_t cnt = top_range(c, mn, mx, lo, hi);
        if (cnt == 0) continue;
        uint32_t slot = c->offs[i];
        if (hi[0] - lo[0] > 0u && hi[1] - lo[1] > 0u && hi[2] - lo[2] > 0u) {
            v3 n = normalize(cross(sub(v1, v0), sub(v2, v0)));
            v3 half = scl(c->cs, 0.5f);
            v3 c0 = add(add(mul(mk((float)lo[0], (float)lo[1], (float)lo[2]), c->cs), c->gmin), half);
            v3 cc;
            cc.z = c0.z - c->cs.z;
            for (uint32_t z = lo[2]; z <= hi[2]; ++z) {
                cc.z += c->cs.z;
                cc.y = c0.y - c->cs.y;
                for (uint32_t y = lo[1]; y <= hi[1]; ++y) {
                    cc.y += c->cs.y;
                    cc.x = c0.x - c->cs.x;
                    for (uint32_t x = lo[0]; x <= hi[0]; ++x, ++slot) {
                        cc.x += c->cs.x;
                        v3 dv = scl(n, dot(sub(cc, v0), n));
                        if (fabsf(dv.x) <= half.x + OR_EPS && fabsf(dv.y) <= half.y + OR_EPS &&
                            fabsf(dv.z) <= half.z + OR_EPS)
                            c->keys[slot] = x + y * c->rx + z * c->rxy;
                        else
                            c->keys[slot] = dummy;
                        c->vals[slot] = (uint32_t)i;
                    }
                }
            }
        } else {
            for (uint32_t z = lo[2]; z <= hi[2]; ++z)
                for (uint32_t y = lo[1]; y <= hi[1]; ++y)
                    for (uint32_t x = lo[0]; x <= hi[0]; ++x, ++slot) {
                        c->keys[slot] = x + y * c->rx + z * c->rxy;
                        c->vals[slot] = (uint32_t)i;
                    }
        }
    }
}

static or_grid *grid_alloc(void)
{
    or_grid *g = (or_grid *)calloc(1, sizeof(or_grid));
    return g;
}

void or_grid_free(or_grid *g)
{
    if (!g) return;
    free(g->tri_counts); free(g->top_keys); free(g->top_vals); free(g->top_ranges);
    free(g->top_cells); free(g->leaf_counts); free(g->pair_counts); free(g->leaf_keys);
    free(g->leaf_vals); free(g->leaves); free(g->indices);
    free(g);
}

/* shared top-level phase: bounds, resolution, parameters, count, scan, emit,
 * sort, ranges.  FastSimpleGridBuilder.h:59-327 and
 * FastTwoLevelGridBuilder.h:61-220 */
static int build_top(or_grid *g, const float *const v[9], uint32_t n, const or_params *p, int divisor)
{
    g->n_tris = n;
    or_bounds(v, n, g->bmin, g->bmax);
    int st;
    if (p->res_override[0] > 0 && p->res_override[1] > 0 && p->res_override[2] > 0) {
        g->res[0] = p->res_override[0]; g->res[1] = p->res_override[1]; g->res[2] = p->res_override[2];
        st = (n == 0 || !(g->bmax[0] > g->bmin[0]) || !(g->bmax[1] > g->bmin[1]) || !(g->bmax[2] > g->bmin[2]))
                 ? OR_ERR_INVALID : OR_OK;
    } else {
        st = or_resolution(g->bmin, g->bmax, n, p->lambda, divisor, g->res);
    }
    if (st != OR_OK) return st;
    v3 resf = mk((float)g->res[0], (float)g->res[1], (float)g->res[2]);
    v3 mn = mk(g->bmin[0], g->bmin[1], g->bmin[2]), mx = mk(g->bmax[0], g->bmax[1], g->bmax[2]);
    v3 cs = divv(sub(mx, mn), resf);                      /* FastSimpleGridBuilder.h:114-115 */
    v3 rcp = divv(mk(1.f, 1.f, 1.f), cs);                  /* FastSimpleGridBuilder.h:117 */
    v3 crcp = divv(resf, sub(mx, mn));                     /* SimpleGrid.h:104-107 */
    g->cs[0] = cs.x; g->cs[1] = cs.y; g->cs[2] = cs.z;
    g->rcp_build[0] = rcp.x; g->rcp_build[1] = rcp.y; g->rcp_build[2] = rcp.z;
    g->cell_rcp[0] = crcp.x; g->cell_rcp[1] = crcp.y; g->cell_rcp[2] = crcp.z;
    uint64_t C = (uint64_t)g->res[0] * (uint64_t)g->res[1] * (uint64_t)g->res[2];
    if (C >= (1ull << 24)) return OR_ERR_TOO_MANY_CELLS;
    g->C1 = (uint32_t)C;

    top_ctx c;
    memset(&c, 0, sizeof(c));
    c.v = v; c.n = n; c.gmin = mn; c.cs = cs; c.rcp = rcp; c.resf = resf;
    c.rx = (uint32_t)resf.x; c.rxy = f2u(resf.x * resf.y); c.C = g->C1;
    g->tri_counts = (uint32_t *)malloc((size_t)n * 4 + 4);
    uint32_t *offs = (uint32_t *)malloc((size_t)n * 4 + 4);
    if (!g->tri_counts || !offs) { free(offs); return OR_ERR_NOMEM; }
    c.counts = g->tri_counts;
    parallel_for(n, p->nthreads, top_count_fn, &c);
    or_exclusive_scan_u32(g->tri_counts, offs, n);
    uint64_t P = 0;
    for (uint32_t i = 0; i < n; ++i) P += g->tri_counts[i];
    if (P >= 0xFFFFFFFFull) { free(offs); return OR_ERR_NOMEM; }
    g->P1 = (uint32_t)P;
    g->top_keys = (uint32_t *)malloc((size_t)P * 4 + 4);
    g->top_vals = (uint32_t *)malloc((size_t)P * 4 + 4);
    if (!g->top_keys || !g->top_vals) { free(offs); return OR_ERR_NOMEM; }
    c.offs = offs; c.keys = g->top_keys; c.vals = g->top_vals;
    parallel_for(n, p->nthreads, top_emit_fn, &c);
    free(offs);
    st = or_sort_pairs(g->top_keys, g->top_vals, P, p->nthreads);
    if (st != OR_OK) return st;
    g->P1_valid = 0;
    while (g->P1_valid < g->P1 && g->top_keys[g->P1_valid] < g->C1) g->P1_valid++;
    g->top_ranges = (uint32_t *)malloc((size_t)g->C1 * 8);
    if (!g->top_ranges) return OR_ERR_NOMEM;
    or_ranges(g->top_keys, P, g->top_ranges, g->C1);
    return OR_OK;
}

or_grid *or_build_uniform(const float *const v[9], uint32_t n, const or_params *p)
{
    or_grid *g = grid_alloc();
    if (!g) return NULL;
    g->kind = 0;
    g->status = build_top(g, v, n, p, 1);
    if (g->status != OR_OK) return g;
    /* prepareCellRanges compaction: indices[i] = pair[i].y for all i */
    g->indices = (uint32_t *)malloc((size_t)g->P1 * 4 + 4);
    if (!g->indices) { g->status = OR_ERR_NOMEM; return g; }
    memcpy(g->indices, g->top_vals, (size_t)g->P1 * 4);
    g->n_indices = g->P1;
    return g;
}

/* ------------------------------------------------------------------------- */
/* leaf level                                                                */
/* ------------------------------------------------------------------------- */

typedef struct {
    const float *const *v;
    const or_grid *g;
    v3 gmin, cs;
    uint32_t rx, rxy;
    const uint32_t *top_keys, *top_vals, *cells;   /* packed top cells */
    uint32_t *pcount, *poffs, *lkeys, *lvals;
    uint32_t C2;
    int *neg;
} leaf_ctx;

/* countLeafLevelRefs / writeLeafLevelRefs sub-cell range,
 * src/Control/Algorithm/TwoLevelBuildKernels.h:175-222 and :327-372 */
static inline int leaf_range(const leaf_ctx *c, uint32_t key, uint32_t tri,
                             v3 *v0o, v3 *v1o, v3 *v2o, v3 *orgo, v3 *creso,
                             int lo[3], int hp[3], uint32_t *packed_x)
{
    uint32_t tx = key % c->rx;
    uint32_t ty = (key % c->rxy) / c->rx;
    uint32_t tz = key / c->rxy;
    v3 v0, v1, v2, mn, mx;
    tri_load(c->v, tri, &v0, &v1, &v2);
    tri_bounds(v0, v1, v2, &mn, &mx);
    uint32_t cx = c->cells[2 * (size_t)key], cy = c->cells[2 * (size_t)key + 1];
    v3 cres = mk((float)(cy & 0xFFu), (float)((cy >> 8) & 0xFFu), (float)((cy >> 16) & 0xFFu));
    v3 org = mk((float)tx * c->cs.x + c->gmin.x, (float)ty * c->cs.y + c->gmin.y, (float)tz * c->cs.z + c->gmin.z);
    v3 srcp = divv(cres, c->cs);
    v3 lof = add(mul(sub(mn, org), srcp), mk(-0.001f, -0.001f, -0.001f));
    v3 hif = add(mul(sub(mx, org), srcp), mk(1.001f, 1.001f, 1.001f));
    int a;
    a = f2i(lof.x); lo[0] = a > 0 ? a : 0;
    a = f2i(lof.y); lo[1] = a > 0 ? a : 0;
    a = f2i(lof.z); lo[2] = a > 0 ? a : 0;
    int rc;
    rc = (int)cres.x; a = f2i(hif.x); hp[0] = rc < a ? rc : a;
    rc = (int)cres.y; a = f2i(hif.y); hp[1] = rc < a ? rc : a;
    rc = (int)cres.z; a = f2i(hif.z); hp[2] = rc < a ? rc : a;
    *v0o = v0; *v1o = v1; *v2o = v2; *orgo = org; *creso = cres; *packed_x = cx;
    int ex = hp[0] - lo[0], ey = hp[1] - lo[1], ez = hp[2] - lo[2];
    if (ex < 0 || ey < 0 || ez < 0) return -1;
    return ex * ey * ez;
}

static void leaf_count_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    leaf_ctx *c = (leaf_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        v3 v0, v1, v2, org, cres;
        int lo[3], hp[3];
        uint32_t px;
        int cnt = leaf_range(c, c->top_keys[i], c->top_vals[i], &v0, &v1, &v2, &org, &cres, lo, hp, &px);
        if (cnt < 0) { *c->neg = 1; cnt = 0; }
        c->pcount[i] = (uint32_t)cnt;
    }
}

/* writeLeafLevelRefs, src/Control/Algorithm/TwoLevelBuildKernels.h:279-455,
 * slot of top pair i = exclusive scan of the per-pair counts */
static void leaf_emit_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    leaf_ctx *c = (leaf_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        if (c->pcount[i] == 0) continue;
        v3 v0, v1, v2, org, cres;
        int lo[3], hp[3];
        uint32_t px;
        uint32_t tri = c->top_vals[i];
        leaf_range(c, c->top_keys[i], tri, &v0, &v1, &v2, &org, &cres, lo, hp, &px);
        uint32_t slot = c->poffs[i];
        uint32_t crx = f2u(cres.x), crxy = f2u(cres.x * cres.y);
        if (hp[0] - lo[0] > 1 && hp[1] - lo[1] > 1 && hp[2] - lo[2] > 1) {
            v3 n = normalize(cross(sub(v1, v0), sub(v2, v0)));
            v3 scs = divv(c->cs, cres);
            v3 c0 = add(add(mul(mk((float)lo[0], (float)lo[1], (float)lo[2]), scs), org), scl(scs, 0.5f));
            v3 cc;
            cc.z = c0.z - scs.z;
            for (int z = lo[2]; z < hp[2]; ++z) {
                cc.z += scs.z;
                cc.y = c0.y - scs.y;
                for (int y = lo[1]; y < hp[1]; ++y) {
                    cc.y += scs.y;
                    cc.x = c0.x - scs.x;
                    for (int x = lo[0]; x < hp[0]; ++x, ++slot) {
                        cc.x += scs.x;
                        v3 dv = scl(n, dot(sub(cc, v0), n));
                        if (fabsf(dv.x) <= scs.x * 0.5f + OR_EPS && fabsf(dv.y) <= scs.y * 0.5f + OR_EPS &&
                            fabsf(dv.z) <= scs.z * 0.5f + OR_EPS)
                            c->lkeys[slot] = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + px;
                        else
                            c->lkeys[slot] = c->C2;
                        c->lvals[slot] = tri;
                    }
                }
            }
        } else {
            for (int z = lo[2]; z < hp[2]; ++z)
                for (int y = lo[1]; y < hp[1]; ++y)
                    for (int x = lo[0]; x < hp[0]; ++x, ++slot) {
                        c->lkeys[slot] = (uint32_t)x + (uint32_t)y * crx + (uint32_t)z * crxy + px;
                        c->lvals[slot] = tri;
                    }
        }
    }
}

or_grid *or_build_twolevel(const float *const v[9], uint32_t n, const or_params *p)
{
    or_grid *g = grid_alloc();
    if (!g) return NULL;
    g->kind = 1;
    g->status = build_top(g, v, n, p, p->top_divisor > 0 ? p->top_divisor : 6);
    if (g->status != OR_OK) return g;
    const uint32_t C1 = g->C1;
    v3 cs = mk(g->cs[0], g->cs[1], g->cs[2]);
    v3 gmin = mk(g->bmin[0], g->bmin[1], g->bmin[2]);

    /* countLeafLevelCells<false>, src/Control/Algorithm/TwoLevelBuildKernels.h:40-103 */
    g->leaf_counts = (uint32_t *)malloc((size_t)C1 * 4);
    g->top_cells = (uint32_t *)malloc((size_t)C1 * 8);
    uint32_t *lofs = (uint32_t *)malloc(((size_t)C1 + 1) * 4);
    if (!g->leaf_counts || !g->top_cells || !lofs) { free(lofs); g->status = OR_ERR_NOMEM; return g; }
    const float vol = cs.x * cs.y * cs.z;
    int limit = 0;
    for (uint32_t k = 0; k < C1; ++k) {
        uint32_t r = g->top_ranges[2 * (size_t)k + 1] - g->top_ranges[2 * (size_t)k];
        float m = or_cbrtpow(p->lambda_leaf * (float)r / vol);
        v3 rf = scl(cs, m);
        int isLeaf = r <= p->leaf_min_refs;
        uint32_t rx = isLeaf ? 1u : f2u(rf.x);
        uint32_t ry = isLeaf ? 1u : f2u(rf.y);
        uint32_t rz = isLeaf ? 1u : f2u(rf.z);
        if (rx < 1 || ry < 1 || rz < 1 || rx > 255 || ry > 255 || rz > 255) limit = 1;
        g->leaf_counts[k] = rx * ry * rz;
        uint32_t bits = NOTEMPTY;
        if (r == 0) bits &= 0x3FFFFFFFu;
        bits |= (rx & 0xFFu) | ((ry & 0xFFu) << 8) | ((rz & 0xFFu) << 16);
        g->top_cells[2 * (size_t)k + 1] = bits;
    }
    /* ExclusiveScan + prepareTopLevelCellRanges, FastTwoLevelGridBuilder.h:288-294 */
    uint64_t C2 = 0;
    for (uint32_t k = 0; k < C1; ++k) { lofs[k] = (uint32_t)C2; C2 += g->leaf_counts[k]; }
    if (C2 >= 0xFFFFFFFFull) { free(lofs); g->status = OR_ERR_NOMEM; return g; }
    for (uint32_t k = 0; k < C1; ++k) g->top_cells[2 * (size_t)k] = lofs[k];
    free(lofs);
    g->C2 = (uint32_t)C2;

    /* countLeafLevelRefs + InclusiveScan, FastTwoLevelGridBuilder.h:345-434 */
    leaf_ctx c;
    memset(&c, 0, sizeof(c));
    int neg = 0;
    c.v = v; c.g = g; c.gmin = gmin; c.cs = cs;
    c.rx = (uint32_t)g->res[0]; c.rxy = (uint32_t)g->res[0] * (uint32_t)g->res[1];
    c.top_keys = g->top_keys; c.top_vals = g->top_vals; c.cells = g->top_cells;
    c.C2 = g->C2; c.neg = &neg;
    g->pair_counts = (uint32_t *)calloc((size_t)g->P1 + 1, 4);
    uint32_t *poffs = (uint32_t *)malloc(((size_t)g->P1 + 1) * 4);
    if (!g->pair_counts || !poffs) { free(poffs); g->status = OR_ERR_NOMEM; return g; }
    c.pcount = g->pair_counts;
    parallel_for(g->P1_valid, p->nthreads, leaf_count_fn, &c);
    uint64_t P2 = 0;
    for (uint32_t i = 0; i < g->P1; ++i) { poffs[i] = (uint32_t)P2; P2 += g->pair_counts[i]; }
    if (P2 >= 0xFFFFFFFFull) { free(poffs); g->status = OR_ERR_NOMEM; return g; }
    g->P2 = (uint32_t)P2;
    g->leaf_keys = (uint32_t *)malloc((size_t)P2 * 4 + 4);
    g->leaf_vals = (uint32_t *)malloc((size_t)P2 * 4 + 4);
    if (!g->leaf_keys || !g->leaf_vals) { free(poffs); g->status = OR_ERR_NOMEM; return g; }
    c.poffs = poffs; c.lkeys = g->leaf_keys; c.lvals = g->leaf_vals;
    parallel_for(g->P1_valid, p->nthreads, leaf_emit_fn, &c);
    free(poffs);

    /* Sort + prepareLeafCellRanges, FastTwoLevelGridBuilder.h:480-529 */
    int st = or_sort_pairs(g->leaf_keys, g->leaf_vals, P2, p->nthreads);
    if (st != OR_OK) { g->status = st; return g; }
    g->P2_valid = 0;
    while (g->P2_valid < g->P2 && g->leaf_keys[g->P2_valid] < g->C2) g->P2_valid++;
    g->leaves = (uint32_t *)malloc((size_t)g->C2 * 8 + 8);
    g->indices = (uint32_t *)malloc((size_t)P2 * 4 + 4);
    if (!g->leaves || !g->indices) { g->status = OR_ERR_NOMEM; return g; }
    or_ranges(g->leaf_keys, P2, g->leaves, g->C2);
    memcpy(g->indices, g->leaf_vals, (size_t)P2 * 4);
    g->n_indices = g->P2;
    if (neg) g->status = OR_ERR_NEG_EXTENT;
    else if (limit) g->status = OR_ERR_LEAF_RES;
    return g;
}

/* ------------------------------------------------------------------------- */
/* traversal                                                                 */
/* ------------------------------------------------------------------------- */

typedef struct {
    v3 o, D, rcp;
    float t;
    uint32_t best;          /* reference index (position in indices), or MISS */
    uint32_t tests;
} ray_state;

/* SimpleIntersector::MollerTrumboreIntersectionTest<.., true>,
 * src/Intersector/SimpleIntersector.hpp:50-89 (strict < keeps the first of
 * equal-t candidates) */
static inline void intersect(const float *const v[9], const uint32_t *indices,
                             uint32_t begin, uint32_t end, ray_state *r)
{
    for (uint32_t it = begin; it != end; ++it) {
        v3 v0, v1, v2;
        tri_load(v, indices[it], &v0, &v1, &v2);
        v3 e1 = sub(v1, v0), e2 = sub(v2, v0);
        v3 pv = cross(r->D, e2);
        float inv = 1.f / dot(e1, pv);
        v3 s = sub(r->o, v0);
        float alpha = inv * dot(s, pv);
        v3 q = cross(s, e1);
        float beta = inv * dot(q, r->D);
        float dist = inv * dot(e2, q);
        if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > OR_EPS && dist < r->t) {
            r->t = dist;
            r->best = it;
        }
    }
    r->tests += end - begin;
}

/* MIN_DIMENSION macro, src/Control/Algorithm/TwoLevelGridTraverser.h:633-634 */
static inline int min_dim(v3 t) { return (t.x < t.y) ? ((t.x < t.z) ? 0 : 2) : ((t.y < t.z) ? 1 : 2); }

/* tex3D clamp addressing on integral coordinates */
static inline uint32_t clamp_idx(float c, int res)
{
    float f = fminf(fmaxf(c, 0.f), (float)(res - 1));
    return (uint32_t)f;
}

typedef struct {
    const or_grid *g;
    const float *const *v;
    const float *const *rays;
    uint32_t *prim; float *t; uint32_t *stats;
    int shadow;             /* any-hit: stop once t < OR_SHADOW_T */
    float *raybuf;          /* optional SimpleRayBuffer, 8 words per ray at stride nrays */
    uint32_t nrays;
} trace_ctx;

/* SimpleRayBuffer::store, src/Engine/RayBuffers.h:52-77: hitPoint = org +
 * (t - EPS) * dir, dir, t, bestHit, each at stride numRays; dir is the
 * direction the trace kernel restores, 1 / rcp (RayTracingKernels.h:128-130),
 * and bestHit the traverser's reference index (OR_MISS on a miss, where the
 * reference leaves it uninitialised) */
static void store_raybuf(const trace_ctx *c, uint32_t i, const ray_state *r)
{
    float *b = c->raybuf + i;
    const size_t s = c->nrays;
    const float te = r->t - OR_EPS;
    b[0] = r->o.x + te * r->D.x;
    b[s] = r->o.y + te * r->D.y;
    b[2 * s] = r->o.z + te * r->D.z;
    b[3 * s] = r->D.x;
    b[4 * s] = r->D.y;
    b[5 * s] = r->D.z;
    b[6 * s] = r->t;
    memcpy(&b[7 * s], &r->best, 4);
}

/* occlusion threshold of the shadow traversers, TwoLevelGridTraverser.h:718,
 * SimpleGridTraverser.h:436 */
#define OR_SHADOW_T 0.9999f

/* UniformGridTraverser, src/Control/Algorithm/SimpleGridTraverser.h:274-401;
 * shadow: UniformGridShadowTraverser :409-464 (starts only if t >= 0.9999,
 * walks to the grid exit and stops once t < 0.9999; no t > tMax test) */
static void trace_uniform_one(const trace_ctx *c, uint32_t i)
{
    const or_grid *g = c->g;
    ray_state r;
    r.o = mk(c->rays[0][i], c->rays[1][i], c->rays[2][i]);
    v3 d = mk(c->rays[3][i], c->rays[4][i], c->rays[5][i]);
    r.rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);     /* RayTracingKernels.h:111-113 */
    r.D = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z); /* SimpleIntersector.hpp:58 */
    r.t = c->rays[6][i];
    r.best = OR_MISS; r.tests = 0;
    uint32_t steps = 0;
    v3 bmin = mk(g->bmin[0], g->bmin[1], g->bmin[2]), bmax = mk(g->bmax[0], g->bmax[1], g->bmax[2]);
    v3 cs = mk(g->cs[0], g->cs[1], g->cs[2]), crcp = mk(g->cell_rcp[0], g->cell_rcp[1], g->cell_rcp[2]);
    int flag = 0;
    v3 cell = mk(0, 0, 0), tMax = mk(0, 0, 0);
    if (c->shadow ? r.t >= OR_SHADOW_T : r.t > 0.f) {
        /* traverseInit :325-369 with BBox::fastClip, BBox.hpp:60-75 */
        v3 t1 = mul(sub(bmin, r.o), r.rcp), t2 = mul(sub(bmax, r.o), r.rcp);
        v3 tn = vmin(t1, t2), tx = vmax(t1, t2);
        float tEntry = fmaxf(fmaxf(tn.x, tn.y), tn.z);
        float tExit = fminf(fminf(tx.x, tx.y), tx.z);
        if (!(tExit <= tEntry || tExit < 0.f)) {
            v3 p = (tEntry >= 0.f) ? add(r.o, divv(mk(tEntry + OR_EPS, tEntry + OR_EPS, tEntry + OR_EPS), r.rcp)) : r.o;
            cell = mul(sub(p, bmin), crcp);
            cell = mk(floorf(cell.x), floorf(cell.y), floorf(cell.z));
            v3 b = mk(r.rcp.x > 0.f ? 1.f : 0.f, r.rcp.y > 0.f ? 1.f : 0.f, r.rcp.z > 0.f ? 1.f : 0.f);
            tMax = mul(sub(add(mul(add(cell, b), cs), bmin), r.o), r.rcp);
            flag = (cell.x != (float)(r.rcp.x > 0.f ? g->res[0] : -1)) &&
                   (cell.y != (float)(r.rcp.y > 0.f ? g->res[1] : -1)) &&
                   (cell.z != (float)(r.rcp.z > 0.f ? g->res[2] : -1));
        }
    }
    /* step cap: ends only walks that never reach the exit test (NaN / inf
     * rays, entry cell past the far boundary) */
    const uint32_t cap = 2u * (uint32_t)(g->res[0] + g->res[1] + g->res[2]) + 16u;
    uint32_t guard = 0;
    while (flag && guard++ < cap) {
        uint32_t k = clamp_idx(cell.x, g->res[0]) +
                     clamp_idx(cell.y, g->res[1]) * (uint32_t)g->res[0] +
                     clamp_idx(cell.z, g->res[2]) * (uint32_t)g->res[0] * (uint32_t)g->res[1];
        intersect(c->v, g->indices, g->top_ranges[2 * (size_t)k], g->top_ranges[2 * (size_t)k + 1], &r);
        if (c->shadow) {
            if (!(r.t >= OR_SHADOW_T)) break;
        } else {
            flag = r.t > tMax.x || r.t > tMax.y || r.t > tMax.z;
            if (!flag) break;
        }
        /* traverseCell :371-401 */
        ++steps;
        int dd = min_dim(tMax);
        float rc = comp(r.rcp, dd);
        setc(&cell, dd, comp(cell, dd) + (rc > 0.f ? 1.f : -1.f));
        setc(&tMax, dd, comp(tMax, dd) + comp(cs, dd) * fabsf(rc));
        float cd = comp(cell, dd);
        flag = !(fabsf(cd - (float)g->res[dd]) < 0.1f || fabsf(cd + 1.f) < 0.1f);
    }
    c->prim[i] = r.best == OR_MISS ? OR_MISS : g->indices[r.best];
    c->t[i] = r.t;
    if (c->raybuf) store_raybuf(c, i, &r);
    if (c->stats) {
        c->stats[4 * (size_t)i] = r.tests; c->stats[4 * (size_t)i + 1] = steps;
        c->stats[4 * (size_t)i + 2] = 0; c->stats[4 * (size_t)i + 3] = 0;
    }
}

/* TwoLevelGridTraverser, src/Control/Algorithm/TwoLevelGridTraverser.h:445-671;
 * shadow: TwoLevelGridShadowTraverser :677-828 (a ray with t < 0.9999 does
 * not start; every leaf cell of an entered top cell and every top cell up to
 * the grid exit is visited until t < 0.9999; no t > tMax tests).  The
 * reference ignores a failed traverseInit there and walks from an
 * uninitialised cell; here such a ray is not traversed. */
static void trace_twolevel_one(const trace_ctx *c, uint32_t i)
{
    const or_grid *g = c->g;
    ray_state r;
    r.o = mk(c->rays[0][i], c->rays[1][i], c->rays[2][i]);
    v3 d = mk(c->rays[3][i], c->rays[4][i], c->rays[5][i]);
    r.rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
    r.D = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z);
    r.t = c->rays[6][i];
    r.best = OR_MISS; r.tests = 0;
    uint32_t lsteps = 0, tsteps = 0, entered = 0;
    v3 bmin = mk(g->bmin[0], g->bmin[1], g->bmin[2]), bmax = mk(g->bmax[0], g->bmax[1], g->bmax[2]);
    v3 cs = mk(g->cs[0], g->cs[1], g->cs[2]), crcp = mk(g->cell_rcp[0], g->cell_rcp[1], g->cell_rcp[2]);
    int flag1 = c->shadow ? !(r.t < OR_SHADOW_T) : r.t > 0.f;
    v3 cell = mk(0, 0, 0), tMax = mk(0, 0, 0);
    float tEntry, tExit;
    {
        /* traverseInit :586-626 */
        v3 t1 = mul(sub(bmin, r.o), r.rcp), t2 = mul(sub(bmax, r.o), r.rcp);
        v3 tn = vmin(t1, t2), tx = vmax(t1, t2);
        tEntry = fmaxf(fmaxf(tn.x, tn.y), tn.z);
        tExit = fminf(fminf(tx.x, tx.y), tx.z);
        int ok;
        if (tExit <= tEntry || tExit < 0.f) ok = 0;
        else {
            tEntry = fmaxf(0.f, tEntry);
            float te = tEntry + OR_EPS;
            v3 p = add(r.o, divv(mk(te, te, te), r.rcp));
            cell = mul(sub(p, bmin), crcp);
            cell = mk(floorf(cell.x), floorf(cell.y), floorf(cell.z));
            v3 b = mk(r.rcp.x > 0.f ? 1.f : 0.f, r.rcp.y > 0.f ? 1.f : 0.f, r.rcp.z > 0.f ? 1.f : 0.f);
            tMax = mul(sub(add(mul(add(cell, b), cs), bmin), r.o), r.rcp);
            ok = (cell.x != (float)(r.rcp.x > 0.f ? g->res[0] : -1)) &&
                 (cell.y != (float)(r.rcp.y > 0.f ? g->res[1] : -1)) &&
                 (cell.z != (float)(r.rcp.z > 0.f ? g->res[2] : -1));
        }
        if (!ok) flag1 = 0;
    }
    const uint32_t cap = 2u * (uint32_t)(g->res[0] + g->res[1] + g->res[2]) + 16u;
    uint32_t guard = 0;
    while (flag1 && guard++ < cap) {
        uint32_t k = clamp_idx(cell.x, g->res[0]) +
                     clamp_idx(cell.y, g->res[1]) * (uint32_t)g->res[0] +
                     clamp_idx(cell.z, g->res[2]) * (uint32_t)g->res[0] * (uint32_t)g->res[1];
        uint32_t cx = g->top_cells[2 * (size_t)k], cy = g->top_cells[2 * (size_t)k + 1];
        if (cy & NOTEMPTY) {
            ++entered;
            /* sub-grid set-up :490-528 */
            v3 ep = add(r.o, divv(mk(tEntry, tEntry, tEntry), r.rcp));
            uint32_t c0 = cy & 0xFFu, c1 = (cy >> 8) & 0xFFu, c2 = (cy >> 16) & 0xFFu;
            v3 cres = mk((float)c0, (float)c1, (float)c2);
            v3 srcp = mul(crcp, cres);
            v3 mb = add(mul(cell, cs), bmin);
            v3 sc;
            sc.x = floorf((ep.x - mb.x) * srcp.x);
            sc.y = floorf((ep.y - mb.y) * srcp.y);
            sc.z = floorf((ep.z - mb.z) * srcp.z);
            v3 scs = mk(1.f / srcp.x, 1.f / srcp.y, 1.f / srcp.z);
            sc = vmin(sub(cres, mk(1.f, 1.f, 1.f)), vmax(sc, mk(0.f, 0.f, 0.f)));
            v3 stMax;
            stMax.x = ((sc.x + (float)(r.rcp.x > 0.f ? 1 : 0)) * scs.x + mb.x - r.o.x) * r.rcp.x;
            stMax.y = ((sc.y + (float)(r.rcp.y > 0.f ? 1 : 0)) * scs.y + mb.y - r.o.y) * r.rcp.y;
            stMax.z = ((sc.z + (float)(r.rcp.z > 0.f ? 1 : 0)) * scs.z + mb.z - r.o.z) * r.rcp.z;
            int flag2 = 1;
            const uint32_t icap = 2u * (c0 + c1 + c2) + 16u;
            uint32_t iguard = 0;
            while (flag2 && iguard++ < icap) {
                /* leaf fetch :537-539; tex1Dfetch returns 0 out of range */
                float lf = sc.x + (float)c0 * sc.y + (float)c0 * (float)c1 * sc.z;
                uint32_t li = cx + f2u(lf);
                uint32_t lb = 0, le = 0;
                if (li < g->C2) { lb = g->leaves[2 * (size_t)li]; le = g->leaves[2 * (size_t)li + 1]; }
                intersect(c->v, g->indices, lb, le, &r);
                if (c->shadow) {
                    if (r.t < OR_SHADOW_T) break;
                } else {
                    flag2 = r.t > stMax.x || r.t > stMax.y || r.t > stMax.z;
                    if (!flag2) break;
                }
                /* traverseCell :651-671 */
                ++lsteps;
                int dd = min_dim(stMax);
                float rc = comp(r.rcp, dd);
                setc(&sc, dd, comp(sc, dd) + (rc > 0.f ? 1.f : -1.f));
                setc(&stMax, dd, comp(stMax, dd) + comp(scs, dd) * fabsf(rc));
                float sd = comp(sc, dd);
                flag2 = !(fabsf(sd - comp(cres, dd)) < 0.1f || fabsf(sd + 1.f) < 0.1f);
            }
        }
        if (c->shadow) {
            if (r.t < OR_SHADOW_T) break;
        } else {
            flag1 = r.t > tMax.x || r.t > tMax.y || r.t > tMax.z;
            if (!flag1) break;
        }
        /* traverseMacroCell :628-649 */
        ++tsteps;
        int dd = min_dim(tMax);
        float rc = comp(r.rcp, dd);
        tEntry = comp(tMax, dd);
        setc(&cell, dd, comp(cell, dd) + (rc > 0.f ? 1.f : -1.f));
        setc(&tMax, dd, comp(tMax, dd) + comp(cs, dd) * fabsf(rc));
        float cd = comp(cell, dd);
        flag1 = !(fabsf(cd - (float)g->res[dd]) < 0.1f || fabsf(cd + 1.f) < 0.1f);
    }
    c->prim[i] = r.best == OR_MISS ? OR_MISS : g->indices[r.best];
    c->t[i] = r.t;
    if (c->raybuf) store_raybuf(c, i, &r);
    if (c->stats) {
        c->stats[4 * (size_t)i] = r.tests; c->stats[4 * (size_t)i + 1] = lsteps;
        c->stats[4 * (size_t)i + 2] = tsteps; c->stats[4 * (size_t)i + 3] = entered;
    }
}

static void trace_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    trace_ctx *c = (trace_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        if (c->g->kind == 0) trace_uniform_one(c, (uint32_t)i);
        else trace_twolevel_one(c, (uint32_t)i);
    }
}

void or_trace(const or_grid *g, const float *const v[9], const float *const rays[7], uint32_t nrays,
              uint32_t *prim, float *t, uint32_t *stats, int nthreads)
{
    trace_ctx c = {g, v, rays, prim, t, stats, 0, NULL, nrays};
    parallel_for(nrays, nthreads, trace_fn, &c);
}

void or_trace_store(const or_grid *g, const float *const v[9], const float *const rays[7], uint32_t nrays,
                    uint32_t *prim, float *t, float *raybuf, int nthreads)
{
    trace_ctx c = {g, v, rays, prim, t, NULL, 0, raybuf, nrays};
    parallel_for(nrays, nthreads, trace_fn, &c);
}

void or_trace_shadow(const or_grid *g, const float *const v[9], const float *const rays[7], uint32_t nrays,
                     uint32_t *prim, float *t, uint32_t *stats, int nthreads)
{
    trace_ctx c = {g, v, rays, prim, t, stats, 1, NULL, nrays};
    parallel_for(nrays, nthreads, trace_fn, &c);
}

/* ------------------------------------------------------------------------- */
/* brute-force referees                                                      */
/* ------------------------------------------------------------------------- */

typedef struct {
    const float *const *v; uint32_t n;
    const float *const *rays;
    uint32_t *prim; float *t; double *td;
} brute_ctx;

static void brute32_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    brute_ctx *c = (brute_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        ray_state r;
        r.o = mk(c->rays[0][i], c->rays[1][i], c->rays[2][i]);
        v3 d = mk(c->rays[3][i], c->rays[4][i], c->rays[5][i]);
        r.rcp = mk(1.f / d.x, 1.f / d.y, 1.f / d.z);
        r.D = mk(1.f / r.rcp.x, 1.f / r.rcp.y, 1.f / r.rcp.z);
        r.t = c->rays[6][i];
        r.best = OR_MISS; r.tests = 0;
        if (r.t > 0.f) {
            for (uint32_t k = 0; k < c->n; ++k) {
                v3 v0, v1, v2;
                tri_load(c->v, k, &v0, &v1, &v2);
                v3 e1 = sub(v1, v0), e2 = sub(v2, v0);
                v3 pv = cross(r.D, e2);
                float inv = 1.f / dot(e1, pv);
                v3 s = sub(r.o, v0);
                float alpha = inv * dot(s, pv);
                v3 q = cross(s, e1);
                float beta = inv * dot(q, r.D);
                float dist = inv * dot(e2, q);
                if (alpha >= 0.f && beta >= 0.f && alpha + beta <= 1.f && dist > OR_EPS && dist < r.t) {
                    r.t = dist; r.best = k;
                }
            }
        }
        c->prim[i] = r.best;
        c->t[i] = r.t;
    }
}

void or_brute_f32(const float *const v[9], uint32_t n, const float *const rays[7], uint32_t nrays,
                  uint32_t *prim, float *t, int nthreads)
{
    brute_ctx c = {v, n, rays, prim, t, NULL};
    parallel_for(nrays, nthreads, brute32_fn, &c);
}

static void brute64_fn(void *c_, size_t b, size_t e, int tid)
{
    (void)tid;
    brute_ctx *c = (brute_ctx *)c_;
    for (size_t i = b; i < e; ++i) {
        double o[3] = {c->rays[0][i], c->rays[1][i], c->rays[2][i]};
        double d[3] = {c->rays[3][i], c->rays[4][i], c->rays[5][i]};
        double t = c->rays[6][i];
        uint32_t best = OR_MISS;
        if (t > 0.0) {
            for (uint32_t k = 0; k < c->n; ++k) {
                double v0[3] = {c->v[0][k], c->v[1][k], c->v[2][k]};
                double e1[3] = {(double)c->v[3][k] - v0[0], (double)c->v[4][k] - v0[1], (double)c->v[5][k] - v0[2]};
                double e2[3] = {(double)c->v[6][k] - v0[0], (double)c->v[7][k] - v0[1], (double)c->v[8][k] - v0[2]};
                double pv[3] = {d[1] * e2[2] - d[2] * e2[1], d[2] * e2[0] - d[0] * e2[2], d[0] * e2[1] - d[1] * e2[0]};
                double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
                if (det == 0.0) continue;
                double inv = 1.0 / det;
                double s[3] = {o[0] - v0[0], o[1] - v0[1], o[2] - v0[2]};
                double alpha = inv * (s[0] * pv[0] + s[1] * pv[1] + s[2] * pv[2]);
                double q[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
                double beta = inv * (q[0] * d[0] + q[1] * d[1] + q[2] * d[2]);
                double dist = inv * (e2[0] * q[0] + e2[1] * q[1] + e2[2] * q[2]);
                if (alpha >= 0.0 && beta >= 0.0 && alpha + beta <= 1.0 && dist > (double)OR_EPS && dist < t) {
                    t = dist; best = k;
                }
            }
        }
        c->prim[i] = best;
        c->td[i] = t;
    }
}

void or_brute_f64(const float *const v[9], uint32_t n, const float *const rays[7], uint32_t nrays,
                  uint32_t *prim, double *t, int nthreads)
{
    brute_ctx c = {v, n, rays, prim, NULL, t};
    parallel_for(nrays, nthreads, brute64_fn, &c);
}

/* ------------------------------------------------------------------------- */
/* bounce rays                                                               */
/* ------------------------------------------------------------------------- */

/* src/Integrator/SimplePathTracer.hpp:115-166 (one bounce per hit, no
 * Russian roulette) with the cosine sampler of
 * src/Utils/HemisphereSamplers.hpp:54-72 */
/* direction class of a bounce ray: bin floor(2 (c + 1)) of each component,
 * clamped to 0..3; class = bx | by << 2 | bz << 4 */
static uint32_t dir_bin(float c)
{
    const float f = (c + 1.f) * 2.f;
    return f < 1.f ? 0u : (f < 2.f ? 1u : (f < 3.f ? 2u : 3u));
}

/* the bounce rays of every hit in input order: 7 floats per ray into tmp,
 * the input ray index into tsrc; returns their number */
static uint32_t gen_bounces(const float *const v[9], const float *const rays[7], const uint32_t *prim,
                            const float *t, uint32_t nrays, uint64_t seed, float *tmp, uint32_t *tsrc)
{
    uint32_t m = 0;
    for (uint32_t i = 0; i < nrays; ++i) {
        if (prim[i] == OR_MISS) continue;
        v3 o = mk(rays[0][i], rays[1][i], rays[2][i]);
        v3 d = mk(rays[3][i], rays[4][i], rays[5][i]);
        v3 v0, v1, v2;
        tri_load(v, prim[i], &v0, &v1, &v2);
        v3 ta = sub(v1, v0), tb = sub(v2, v0);
        v3 n = normalize(cross(ta, tb));
        ta = normalize(ta);
        tb = cross(n, ta);
        if (dot(n, d) > 0.f) n = mk(-n.x, -n.y, -n.z);
        v3 no = add(add(o, scl(d, t[i])), scl(n, 0.001f));
        float u1 = or_u01(or_hash3(seed, i, 0)), u2 = or_u01(or_hash3(seed, i, 1));
        float sn, cn;
        or_sincos_2pi(u1, &sn, &cn);
        float rr = sqrtf(1.f - u2);
        v3 rd = mk(sn * rr, cn * rr, sqrtf(u2));
        v3 nd = add(add(scl(ta, rd.x), scl(tb, rd.y)), scl(n, rd.z));
        float *r = tmp + (size_t)7 * m;
        r[0] = no.x; r[1] = no.y; r[2] = no.z;
        r[3] = nd.x; r[4] = nd.y; r[5] = nd.z;
        r[6] = FLT_MAX;
        tsrc[m] = i;
        ++m;
    }
    return m;
}

uint32_t or_bounce(const float *const v[9], const float *const rays[7], const uint32_t *prim,
                   const float *t, uint32_t nrays, uint64_t seed, float *const out[7], uint32_t *src)
{
    /* the bounce rays in input order, then ordered by (r / 256, class, r % 256),
     * r the ray's rank within its direction class in input order */
    float *tmp = (float *)malloc((size_t)7 * (nrays ? nrays : 1) * sizeof(float));
    uint32_t *tsrc = (uint32_t *)malloc((size_t)(nrays ? nrays : 1) * sizeof(uint32_t));
    uint32_t *cls = (uint32_t *)malloc((size_t)(nrays ? nrays : 1) * sizeof(uint32_t));
    const uint32_t m = gen_bounces(v, rays, prim, t, nrays, seed, tmp, tsrc);
    for (uint32_t k = 0; k < m; ++k) {
        const float *r = tmp + (size_t)7 * k;
        cls[k] = dir_bin(r[3]) | dir_bin(r[4]) << 2 | dir_bin(r[5]) << 4;
    }
    uint32_t tot[64] = {0}, seen[64] = {0};
    for (uint32_t k = 0; k < m; ++k) ++tot[cls[k]];
    for (uint32_t k = 0; k < m; ++k) {
        const uint32_t c = cls[k], r = seen[c]++;
        const uint32_t j0 = r / 256u * 256u;
        uint32_t p = r - j0;   /* rays of earlier rounds, then this round's of earlier classes */
        for (uint32_t q = 0; q < 64; ++q) {
            p += tot[q] < j0 ? tot[q] : j0;
            if (q < c) {
                const uint32_t left = tot[q] > j0 ? tot[q] - j0 : 0u;
                p += left < 256u ? left : 256u;
            }
        }
        for (int a = 0; a < 7; ++a) out[a][p] = tmp[(size_t)7 * k + a];
        src[p] = tsrc[k];
    }
    free(tmp);
    free(tsrc);
    free(cls);
    return m;
}

/* f -> [0, 65535]: 0 for NaN and f <= 0, else min(65535, (uint32)f) */
static uint32_t quant16(float f)
{
    if (!(f > 0.f)) return 0u;
    return f >= 65535.f ? 65535u : (uint32_t)f;
}

uint32_t or_bounce_ordered(const float *const v[9], const float *const rays[7], const uint32_t *prim,
                           const float *t, uint32_t nrays, uint64_t seed, int order, const float box_min[3],
                           const float box_max[3], const float axis[3], float *const out[7], uint32_t *src)
{
    float *tmp = (float *)malloc((size_t)7 * (nrays ? nrays : 1) * sizeof(float));
    uint32_t *tsrc = (uint32_t *)malloc((size_t)(nrays ? nrays : 1) * sizeof(uint32_t));
    uint32_t *key = (uint32_t *)malloc((size_t)(nrays ? nrays : 1) * sizeof(uint32_t));
    uint32_t *idx = (uint32_t *)malloc((size_t)(nrays ? nrays : 1) * sizeof(uint32_t));
    const uint32_t m = gen_bounces(v, rays, prim, t, nrays, seed, tmp, tsrc);
    const float ddx = box_max[0] - box_min[0], ddy = box_max[1] - box_min[1], ddz = box_max[2] - box_min[2];
    const float scale = 65535.f / sqrtf((ddx * ddx + ddy * ddy) + ddz * ddz);
    for (uint32_t k = 0; k < m; ++k) {
        const float *r = tmp + (size_t)7 * k;
        const v3 o = mk(r[0], r[1], r[2]), d = mk(r[3], r[4], r[5]);
        if (order == 2) {
            key[k] = quant16(fabsf(dot(d, mk(axis[0], axis[1], axis[2]))) * 65535.f);
        } else {
            /* far slab distance per axis, the smallest of the three: the
             * distance at which the ray leaves the box, in steps of the
             * diagonal / 65535, longest first */
            const float fx = fmaxf((box_min[0] - o.x) / d.x, (box_max[0] - o.x) / d.x);
            const float fy = fmaxf((box_min[1] - o.y) / d.y, (box_max[1] - o.y) / d.y);
            const float fz = fmaxf((box_min[2] - o.z) / d.z, (box_max[2] - o.z) / d.z);
            key[k] = 65535u - quant16(fmaxf(fminf(fminf(fx, fy), fz), 0.f) * scale);
        }
        idx[k] = k;
    }
    or_sort_pairs(key, idx, m, 4);   /* stable: ties keep input order */
    for (uint32_t j = 0; j < m; ++j) {
        for (int a = 0; a < 7; ++a) out[a][j] = tmp[(size_t)7 * idx[j] + a];
        src[j] = tsrc[idx[j]];
    }
    free(tmp);
    free(tsrc);
    free(key);
    free(idx);
    return m;
}
