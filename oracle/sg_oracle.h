/*
 * sg_oracle.h -- CPU oracle for the grid-acceleration hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the Shinobi
 * reference builders and traversers (FastSimpleGridBuilder /
 * FastTwoLevelGridBuilder + UniformGridTraverser / TwoLevelGridTraverser +
 * SimpleIntersector).  Only tests/, the smoke check and the cpu_baseline leg
 * of bench.py may load it, and only as the checker or as the timed CPU
 * baseline.  The product path (shinobi_amd) never links or calls it.
 *
 * Semantics: IEEE-754 binary32, round to nearest, no contraction, evaluated
 * in the order the reference source writes it (see sg_oracle.c for the
 * file:line each function follows).  Where the reference is nondeterministic
 * (atomic emission slots) the oracle produces the canonical order: triangles
 * ascending, then the reference's z/y/x loop order, made final by a stable
 * sort on the cell key.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   bounds / resolution ........ closed forms on hand-built scenes
 *   per-triangle counts/emission  brute-force enumeration + invariants
 *   scan ........................ reference ScanTest known answer
 *   stable sort ................. textbook stable sort (numpy kind='stable')
 *   ranges ...................... searchsorted lower/upper bound
 *   traversal + MT .............. FP64 and FP32 brute force, IntersectorTest
 *                                 tie scene, axis-parallel and inside rays
 */
#ifndef SG_ORACLE_H
#define SG_ORACLE_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MISS 0xFFFFFFFFu

enum {
    OR_OK = 0,
    OR_ERR_INVALID = 1,      /* empty scene, zero-extent bounds, res < 1     */
    OR_ERR_LEAF_RES = 2,     /* a leaf resolution outside [1,255] (D5)       */
    OR_ERR_TOO_MANY_CELLS = 3,/* C >= 2^24: float key products inexact (D6)  */
    OR_ERR_NEG_EXTENT = 4,   /* leaf sub-cell extent < 0 (reference corrupts) */
    OR_ERR_NOMEM = 5
};

typedef struct {
    float lambda;            /* top-level density (5 in the reference)            */
    int   top_divisor;       /* two-level: 6 (FastTwoLevelGridBuilder.h:100-102)  */
    float lambda_leaf;       /* leaf density (1.2 in TwoLevelBuildKernels.h:49)   */
    uint32_t leaf_min_refs;  /* a top cell with <= this many refs is one leaf (16)*/
    int   res_override[3];   /* all > 0: fixed resolution (rebuild semantics)     */
    int   nthreads;          /* worker threads (>= 1)                             */
} or_params;

typedef struct {
    int kind;                /* 0 uniform, 1 two-level */
    int status;              /* OR_OK or an OR_ERR_* code */
    int res[3];
    float bmin[3], bmax[3];
    float cs[3];             /* (max-min)/res       build + traversal cell size   */
    float rcp_build[3];      /* 1/cs                build (FastSimpleGridBuilder.h:117) */
    float cell_rcp[3];       /* res/(max-min)       traversal (SimpleGrid.h:104-107)    */
    uint32_t n_tris;
    uint32_t C1, P1, P1_valid;
    uint32_t C2, P2, P2_valid;
    uint32_t *tri_counts;    /* N   per-triangle top-level pair counts            */
    uint32_t *top_keys;      /* P1  sorted top-level pairs (key)                   */
    uint32_t *top_vals;      /* P1  sorted top-level pairs (triangle id)           */
    uint32_t *top_ranges;    /* 2*C1 [begin,end) into the sorted top pairs         */
    uint32_t *top_cells;     /* 2*C1 packed TwoLevelGridCell {leafBegin, bits}     */
    uint32_t *leaf_counts;   /* C1  leaf cells per top cell (unscanned)            */
    uint32_t *pair_counts;   /* P1  leaf refs per top pair                         */
    uint32_t *leaf_keys;     /* P2  sorted leaf pairs (key)                        */
    uint32_t *leaf_vals;     /* P2  sorted leaf pairs (triangle id)                */
    uint32_t *leaves;        /* 2*C2 leaf ranges                                   */
    uint32_t *indices;       /* P (uniform P1, two-level P2) reference -> triangle */
    uint32_t n_indices;
} or_grid;

/* ---- deterministic scalar math shared by builder definitions ---------- */
double or_det_pow(double x, double y);
float  or_cbrtpow(float x);          /* (float)det_pow((double)x, (double)0.3333333f) */
void   or_sincos_2pi(float u, float *s, float *c);

/* ---- data-parallel primitives (reference L2) ---------------------------- */
void or_exclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n);
void or_inclusive_scan_u32(const uint32_t *in, uint32_t *out, size_t n);
/* stable sort of (key,val) pairs by key, in place */
int  or_sort_pairs(uint32_t *keys, uint32_t *vals, size_t n, int nthreads);
/* cell ranges [first,last+1) of every key < ncells; cells zeroed first */
void or_ranges(const uint32_t *sorted_keys, size_t n, uint32_t *cells, uint32_t ncells);

/* ---- scene-level quantities --------------------------------------------- */
void or_bounds(const float *const v[9], uint32_t n, float bmin[3], float bmax[3]);
int  or_resolution(const float bmin[3], const float bmax[3], uint32_t n,
                   float lambda, int divisor, int res[3]);

/* ---- builders ------------------------------------------------------------ */
or_grid *or_build_uniform(const float *const v[9], uint32_t n, const or_params *p);
or_grid *or_build_twolevel(const float *const v[9], uint32_t n, const or_params *p);
void     or_grid_free(or_grid *g);

/* ---- traversal ------------------------------------------------------------
 * rays: ox,oy,oz,dx,dy,dz,tmax (7 SoA arrays).  prim = triangle id or
 * OR_MISS, t = final ray t (tmax on a miss).  stats (optional, 4 per ray):
 * triangle tests, leaf/uniform DDA steps, top-level DDA steps, non-empty top
 * cells entered. */
void or_trace(const or_grid *g, const float *const v[9],
              const float *const rays[7], uint32_t nrays,
              uint32_t *prim, float *t, uint32_t *stats, int nthreads);
/* or_trace that also writes SimpleRayBuffer::store's 8 words per ray
 * (src/Engine/RayBuffers.h:52-77) into raybuf at stride nrays: hit point
 * org + (t - EPS) dir, dir = 1 / (1 / d), t, bestHit (the reference index,
 * OR_MISS on a miss) */
void or_trace_store(const or_grid *g, const float *const v[9], const float *const rays[7], uint32_t nrays,
                    uint32_t *prim, float *t, float *raybuf, int nthreads);
/* any-hit (shadow) traversal of the same grid: the reference's shadow
 * traversers, which stop a ray once its t drops below 0.9999 (occluded) and
 * otherwise walk it to the grid exit; same outputs as or_trace. */
void or_trace_shadow(const or_grid *g, const float *const v[9],
                     const float *const rays[7], uint32_t nrays,
                     uint32_t *prim, float *t, uint32_t *stats, int nthreads);

/* brute force over every triangle, same MT formula, reciprocal-direction form */
void or_brute_f32(const float *const v[9], uint32_t n,
                  const float *const rays[7], uint32_t nrays,
                  uint32_t *prim, float *t, int nthreads);
/* FP64 referee: MT in double with the original direction */
void or_brute_f64(const float *const v[9], uint32_t n,
                  const float *const rays[7], uint32_t nrays,
                  uint32_t *prim, double *t, int nthreads);

/* ---- secondary rays --------------------------------------------------------
 * One cosine-weighted diffuse bounce per hit (SimplePathTracer.hpp:115-166),
 * compacted and ordered by direction class (bin floor(2 (c + 1)) of each
 * component of the new direction, clamped to 0..3; class = bx | by<<2 |
 * bz<<4) in rounds of 256: sorted by (r / 256, class, r % 256), r the ray's
 * rank within its class in input order -- the order the library's generator
 * writes.  Returns the number of
 * rays written; src[i] is the primary ray each bounce came from. */
uint32_t or_bounce(const float *const v[9], const float *const rays[7],
                   const uint32_t *prim, const float *t, uint32_t nrays,
                   uint64_t seed, float *const out[7], uint32_t *src);

/* The same bounce rays ordered by a 16-bit key, smallest first, ties in
 * input order (the library's sg_bounce_ordered): order 1 = 65535 -
 * min(65535, (uint)(max(e, 0) * 65535 / diagonal)), e the exit distance from
 * the box [box_min, box_max] (per axis fmaxf((min - o) / d, (max - o) / d),
 * the smallest of the three), diagonal = sqrtf((dx*dx + dy*dy) + dz*dz);
 * order 2 = min(65535, (uint)(|d . axis| * 65535)). */
uint32_t or_bounce_ordered(const float *const v[9], const float *const rays[7], const uint32_t *prim,
                           const float *t, uint32_t nrays, uint64_t seed, int order, const float box_min[3],
                           const float box_max[3], const float axis[3], float *const out[7], uint32_t *src);

uint64_t or_hash3(uint64_t seed, uint64_t a, uint64_t b);
float    or_u01(uint64_t h);

#ifdef __cplusplus
}
#endif
#endif
