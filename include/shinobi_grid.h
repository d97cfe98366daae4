/*
 * shinobi_grid.h -- C ABI of the MI355X (gfx950) grid-acceleration hot path.
 *
 * The reference wires its builder, primitive storage, intersector and
 * traverser together as compile-time C++ concepts
 * (src/Engine/RayTracingTypes.h:84-128, src/Shinobi.h:48-120); there is no
 * FFI.  This header is the C-ABI replacement of that wiring: plain pointers,
 * sizes and status codes, no HIP or torch types in the signatures (streams
 * are passed as void* and are hipStream_t values; NULL = the null stream).
 * include/shinobi/grid.hpp restores the reference's C++ concepts on top.
 *
 * Threading: calls on different sg_ctx / sg_grid objects may run
 * concurrently; one object must not be used from two threads at once.
 * Errors: every entry point returns an sg_status; nothing prints or exits
 * (the reference's CUDA_SAFE_CALL / CUT_CHECK_ERROR print and exit,
 * src/CUDAStdAfx.h:38).
 */
#ifndef SHINOBI_GRID_H
#define SHINOBI_GRID_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int sg_status;
enum {
    SG_OK = 0,
    SG_ERR_INVALID_ARG = 1,     /* null pointer, bad enum, size overflow          */
    SG_ERR_INVALID_SCENE = 2,   /* empty scene, zero-extent bounds, res < 1       */
    SG_ERR_OUT_OF_MEMORY = 3,   /* device allocation failed                        */
    SG_ERR_HIP = 4,             /* a HIP runtime call failed                       */
    SG_ERR_CAPACITY = 5,        /* a rebuild outgrew the grid's buffers            */
    SG_ERR_LEAF_RES = 6,        /* a leaf resolution left [1,255] (reference D5)  */
    SG_ERR_TOO_MANY_CELLS = 7,  /* >= 2^24 cells: float key products inexact      */
    SG_ERR_NEG_EXTENT = 8,      /* a negative leaf extent (reference corrupts)     */
    SG_ERR_TIMEOUT = 9,         /* a device-side wait exceeded its bound           */
    SG_ERR_BUDGET = 10,         /* request exceeds the serving HBM budget          */
    SG_ERR_NO_DEVICE = 11,      /* no HIP device                                   */
    SG_ERR_IO = 12,             /* a scene or camera file cannot be read           */
    SG_ERR_SCRATCH_LIMIT = 13   /* input larger than an allocation-free entry
                                   point's fixed scratch admits                   */
};

const char *sg_status_str(sg_status s);

typedef struct sg_ctx sg_ctx;
typedef struct sg_scene sg_scene;
typedef struct sg_grid sg_grid;

enum { SG_GRID_UNIFORM = 0, SG_GRID_TWO_LEVEL = 1 };

/* Build parameters.  Defaults (sg_build_params_default) are the reference's:
 * lambda 5 (GridBuilder.h:48, FastTwoLevelGridBuilder.h:90), top divisor 6
 * (FastTwoLevelGridBuilder.h:100-102), leaf lambda 1.2 and the 16-reference
 * leaf threshold (TwoLevelBuildKernels.h:49,73). */
typedef struct {
    int kind;                 /* SG_GRID_UNIFORM or SG_GRID_TWO_LEVEL          */
    float lambda;             /* top-level (or uniform) density                */
    int top_divisor;          /* two-level top resolution divisor              */
    float lambda_leaf;        /* leaf density                                  */
    uint32_t leaf_min_refs;   /* <= this many refs: the top cell is one leaf   */
    int res_override[3];      /* all > 0: use this resolution                  */
    int radix_bits;           /* 4, 6 or 8 (0 = 8)                             */
    float capacity_slack;     /* buffer headroom for rebuilds (0 = 1.25)       */
    int keep_debug;           /* keep per-triangle / per-pair counts           */
    int trace_block;          /* threads per block of this grid's traversal
                                 launches: 64, 128 or 256 (0 = 64)            */
} sg_build_params;

void sg_build_params_default(sg_build_params *p, int kind);

typedef struct {
    int kind;
    sg_status status;
    int res[3];
    float bmin[3], bmax[3];
    float cell_size[3];       /* (max-min)/res                                  */
    float cell_size_rcp[3];   /* res/(max-min)  (traversal, TwoLevelGrid.h:84-87) */
    float build_rcp[3];       /* 1/cell_size    (build, FastTwoLevelGridBuilder.h:123) */
    uint32_t n_tris;
    uint32_t C1, P1, P1_valid;  /* top cells, top pairs (incl. rejected), accepted */
    uint32_t C2, P2, P2_valid;  /* leaf cells, leaf pairs, accepted leaf pairs     */
    uint32_t cap_C1, cap_P1, cap_C2, cap_P2;
    uint64_t device_bytes;
} sg_grid_info;

/* arrays retrievable with sg_grid_download (all uint32) */
enum {
    SG_ARRAY_TOP_KEYS = 0,     /* P1  sorted top-level pair keys               */
    SG_ARRAY_TOP_VALS = 1,     /* P1  sorted top-level pair triangle ids       */
    SG_ARRAY_TOP_RANGES = 2,   /* 2*C1 top-level [begin,end) (uniform: cells)  */
    SG_ARRAY_TOP_CELLS = 3,    /* 2*C1 packed TwoLevelGridCell                 */
    SG_ARRAY_LEAF_KEYS = 4,    /* P2  sorted leaf pair keys (keep_debug)       */
    SG_ARRAY_LEAF_VALS = 5,    /* P2                                           */
    SG_ARRAY_LEAVES = 6,       /* 2*C2 leaf [begin,end)                        */
    SG_ARRAY_INDICES = 7,      /* P   reference -> triangle (FaceSoup::indices) */
    SG_ARRAY_TRI_COUNTS = 8,   /* N   per-triangle top counts (keep_debug)     */
    SG_ARRAY_LEAF_COUNTS = 9,  /* C1  leaf cells per top cell (keep_debug)     */
    SG_ARRAY_PAIR_COUNTS = 10  /* P1  leaf refs per top pair (keep_debug)      */
};

/* rays: seven device arrays ox,oy,oz,dx,dy,dz,tmax (SoA, fp32) */
typedef struct {
    const float *o[3];
    const float *d[3];
    const float *tmax;
} sg_rays;

typedef struct {
    float *o[3];
    float *d[3];
    float *tmax;
} sg_rays_out;

/* ------------------------------------------------------------ context --- */
/* replaces the implicit single CUDA context of the reference */
sg_status sg_ctx_create(int device, sg_ctx **out);
sg_status sg_ctx_destroy(sg_ctx *ctx);
sg_status sg_device_count(int *out);

/* -------------------------------------------------------------- scene --- */
/* Primitive storage: replaces FaceSoup::upload (src/Control/Structure/
 * FaceSoup.h:171-247).  soa9 = v0.x,v0.y,v0.z,v1.x,...,v2.z, n floats each,
 * in host memory (src_on_device = 0) or device memory (1).  The library keeps
 * a device copy and a 36-byte record per triangle, written when vertices are
 * uploaded (here and by sg_scene_update) and read by builds and traversal. */
sg_status sg_scene_create(sg_ctx *ctx, const float *const soa9[9], uint32_t n, int src_on_device,
                          void *stream, sg_scene **out);
/* new vertex positions, same triangle count: replaces FaceSoup::uploadVertices
 * (src/Control/Structure/FaceSoup.h:249-339) */
sg_status sg_scene_update(sg_scene *scene, const float *const soa9[9], int src_on_device, void *stream);
sg_status sg_scene_destroy(sg_scene *scene);
uint32_t sg_scene_size(const sg_scene *scene);
/* device pointers of the scene's SoA arrays (9), read-only: new vertices go
 * through sg_scene_update (src_on_device = 1 for device data), which also
 * rewrites the triangle records */
sg_status sg_scene_device_soa(const sg_scene *scene, const float *out9[9]);

/* -------------------------------------------------------------- build --- */
/* Builder::init + build (FastSimpleGridBuilder.h:59-327,
 * FastTwoLevelGridBuilder.h:61-570; called by StaticRTEngine::build,
 * src/Engine/StaticRTEngine.h:225-237).  Sizes the buffers (it may wait on
 * the stream); the build itself is enqueued on `stream`. */
sg_status sg_build(sg_ctx *ctx, sg_scene *scene, const sg_build_params *p, void *stream, sg_grid **out);
/* Builder::rebuild (FastTwoLevelGridBuilder.h:573-740): same resolution,
 * bounds and everything else recomputed on the device from the scene's
 * current vertices, into the grid's existing buffers.  Never synchronizes,
 * never allocates, captures into a hipGraph.  A rebuild that outgrows the
 * buffers is reported by sg_grid_get_info as SG_ERR_CAPACITY. */
sg_status sg_rebuild(sg_ctx *ctx, sg_scene *scene, sg_grid *grid, void *stream);
/* like sg_rebuild but lets the device recompute the resolution as well */
sg_status sg_rebuild_free_res(sg_ctx *ctx, sg_scene *scene, sg_grid *grid, void *stream);
/* sg_rebuild of n (scene, grid) pairs, the rebuilds spread over up to
 * `branches` (1..SG_REBUILD_BRANCHES_MAX, 0 = the maximum) streams the
 * context owns: the branches start after the work already enqueued on
 * `stream` and the work enqueued on `stream` afterwards waits for all of
 * them.  Largest scene first, each to the branch with the least work so far.
 * A small scene's rebuild is a dozen short launches that leave most of the
 * device idle; concurrent branches overlap them.  Captures into a hipGraph
 * as parallel branches.  Grids must be distinct; each grid's result equals
 * its sg_rebuild.  No counterpart in the reference, whose builders run one
 * scene at a time (FastTwoLevelGridBuilder.h:755-781). */
#define SG_REBUILD_BRANCHES_MAX 16
sg_status sg_rebuild_many(sg_ctx *ctx, uint32_t n, sg_scene *const *scenes, sg_grid *const *grids,
                          uint32_t branches, void *stream);
/* sg_rebuild that also records hipEvent_t events[i] at the phase boundaries,
 * the analogue of the builders' phase timers (FastTwoLevelGridBuilder.h:
 * 755-781): [0] start, [1] bounds, [2] top emission, [3] top sort, [4] top
 * ranges, and for the two-level grid [5] leaf cells, [6] leaf emission,
 * [7] leaf sort, [8] leaf ranges.  Null entries are skipped. */
sg_status sg_rebuild_phased(sg_ctx *ctx, sg_scene *scene, sg_grid *grid, void *stream, void *const *events,
                            int n_events);
/* waits for the grid's stream, reads the device header */
sg_status sg_grid_get_info(sg_grid *grid, sg_grid_info *out);
sg_status sg_grid_download(sg_grid *grid, int which, void *host, size_t bytes);
/* Device pointers of a built grid and its scene, for kernels written against
 * the reference's device concepts (include/shinobi/device.hpp: the storage
 * FaceSoup's indices / operator[] / operator(), SimpleIntersector, the
 * grid traversers).  In the reference they read the grid through bound
 * textures and __constant__ GridParameters (src/Control/Structure/TwoLevelGrid.h:
 * 36-37, 263-301, src/Control/Structure/FaceSoup.h:36-39); here they read
 * these arrays.  Waits for the grid's stream (it reads the device header).
 * The pointers stay valid until the grid (or the scene) is destroyed; a
 * rebuild rewrites the arrays in place. */
typedef struct {
    int kind;                 /* SG_GRID_UNIFORM or SG_GRID_TWO_LEVEL           */
    sg_status status;         /* the build's status, as sg_grid_get_info        */
    int res[3];
    float bmin[3], bmax[3];
    float cell_size[3];       /* (max-min)/res                                  */
    float cell_size_rcp[3];   /* res/(max-min)                                  */
    const uint32_t *cells;    /* 2*n_cells: uniform [begin,end) ranges; two-level
                                 packed TwoLevelGridCell                        */
    const uint32_t *leaves;   /* two-level: 2*n_leaves leaf ranges (else NULL)  */
    uint32_t n_cells, n_leaves;
    const uint32_t *indices;  /* n_refs: reference -> triangle id (FaceSoup::indices) */
    uint32_t n_refs;
    const float *tri[9];      /* the scene's SoA vertices v0.xyz v1.xyz v2.xyz  */
    uint32_t n_tris;
} sg_grid_device_view;
sg_status sg_grid_get_device_view(sg_grid *grid, sg_grid_device_view *out);
sg_status sg_grid_destroy(sg_grid *grid);

/* ---------------------------------------------------------- traversal --- */
/* The trace kernel contract (src/Engine/RayTracingKernels.h:37-136) with
 * SimpleIntersector (src/Intersector/SimpleIntersector.hpp:312-337) and the
 * grid's traverser (UniformGridTraverser / TwoLevelGridTraverser):
 * rays in -> prim (triangle id, 0xFFFFFFFF on a miss) and final t.
 * d_count: optional device pointer holding the ray count (<= n). */
sg_status sg_trace_closest(sg_ctx *ctx, sg_grid *grid, const sg_rays *rays, uint32_t n,
                           const uint32_t *d_count, uint32_t *d_prim, float *d_t, uint32_t *d_stats,
                           void *stream);
/* Any-hit (shadow) traversal: TwoLevelGridShadowTraverser
 * (src/Control/Algorithm/TwoLevelGridTraverser.h:677-828) or
 * UniformGridShadowTraverser (src/Control/Algorithm/SimpleGridTraverser.h:409-464)
 * as bound by DirectIlluminationIntegrator (src/Integrator/DirectIlluminationIntegrator.h:188-189).
 * Rays are segments (direction = light - origin, tmax = 1); a ray is
 * occluded when the returned t < 0.9999 (OcclusionRayBuffer::store,
 * src/Engine/RayBuffers.h:156-167).  prim is the occluder found first
 * (0xFFFFFFFF when unoccluded).  Same arguments as sg_trace_closest. */
sg_status sg_trace_shadow(sg_ctx *ctx, sg_grid *grid, const sg_rays *rays, uint32_t n,
                          const uint32_t *d_count, uint32_t *d_prim, float *d_t, uint32_t *d_stats,
                          void *stream);
/* The trace kernel's output buffer, SimpleRayBuffer::store
 * (src/Engine/RayBuffers.h:52-77): 8 words per ray, SoA at stride n --
 * hit point org + (t - EPS) * dir (x, y, z), dir (x, y, z), t, bestHit (the
 * uint32 bits of the reference index the traverser returns: FaceSoup::indices
 * maps it to the triangle; 0xFFFFFFFF on a miss, where the reference leaves
 * it uninitialised).  dir is 1 / (1 / d) per component, the direction the
 * trace kernel restores from rcp(dir) (src/Engine/RayTracingKernels.h:128-130).
 * d_raybuf: 8 * n floats; d_prim / d_t (optional) as sg_trace_closest. */
sg_status sg_trace_closest_store(sg_ctx *ctx, sg_grid *grid, const sg_rays *rays, uint32_t n,
                                 const uint32_t *d_count, float *d_raybuf, uint32_t *d_prim, float *d_t,
                                 void *stream);

/* A pinhole camera after Camera::init (src/Primitive/Camera.h:41-60) with
 * the resolution of its RegularPixelSampler<1,1> (src/Primitive/PixelSampler.h:34-54). */
typedef struct {
    float position[3];
    float top_left[3];    /* mTopLeft: forward - row / 2 - col / 2 */
    float step_x[3];      /* mStepX = row / resX */
    float step_y[3];      /* mStepY = col / resY */
    uint32_t res_x, res_y;
} sg_pinhole;
/* Camera::init on the host: eye, forward, up, vertical opening angle in degrees */
sg_status sg_pinhole_init(const float eye[3], const float fwd[3], const float up[3], float fov_deg, uint32_t res_x,
                          uint32_t res_y, sg_pinhole *out);
/* Closest hit of the camera's primary rays, generated inside the trace
 * kernel as RegularPrimaryRayGenerator does (src/Engine/RayGenerators.h:36-61,
 * called per ray by trace, src/Engine/RayTracingKernels.h:104-106): ray of
 * pixel id p = x + y * res_x has org = position, dir = topLeft + (frac(p /
 * resX) resX + 0.5) stepX + (trunc(p / resX) + 0.5) stepY, t = FLT_MAX.
 * The launch visits the pixels in tiles of tile_w x tile_h (tiles in
 * row-major order, row-major inside a tile; 1 x 1 is the reference's
 * row-major order) so that a wave holds a tile's rays; every output is
 * written at the pixel id, so results do not depend on the tiles.
 * Outputs, each optional: d_prim / d_t (res_x * res_y each, as
 * sg_trace_closest), d_raybuf (SimpleRayBuffer, 8 * res_x * res_y floats,
 * as sg_trace_closest_store), rays_out (the generated rays, for a bounce). */
sg_status sg_trace_camera(sg_ctx *ctx, sg_grid *grid, const sg_pinhole *cam, uint32_t tile_w, uint32_t tile_h,
                          uint32_t *d_prim, float *d_t, float *d_raybuf, const sg_rays_out *rays_out, void *stream);

/* one cosine-weighted diffuse bounce per hit (SimplePathTracer.hpp:115-166),
 * compacted and ordered by direction class (4 bins of each component of the
 * new direction, class = bx | by<<2 | bz<<4) in rounds of 256 rays: sorted by
 * (r / 256, class, r % 256), r the ray's rank within its class in input
 * order -- waves of one class, the classes spread over the launch; *d_count = number
 * of bounce rays; d_src (opt.) = input ray index of each bounce ray.
 * n = capacity of the input; d_n
 * (optional device pointer) = the number of input rays (<= n), so bounces
 * chain path segments without reading counts back to the host. */
sg_status sg_bounce(sg_ctx *ctx, sg_scene *scene, const sg_rays *rays, const uint32_t *d_prim,
                    const float *d_t, uint32_t n, const uint32_t *d_n, uint64_t seed, const sg_rays_out *out,
                    uint32_t *d_src, uint32_t *d_count, void *stream);
/* Orders of the bounce rays of sg_bounce_ordered.  Ray order changes no
 * result (each ray's hit is its own); it decides which rays share a wave and
 * when the long ones start, and a launch lasts about as long as its longest
 * rays.  No counterpart in the reference, whose path tracer keeps pixel order.
 *   SG_BOUNCE_CLASS_ROUNDS  sg_bounce's order (direction classes in rounds)
 *   SG_BOUNCE_EXIT_FIRST    longest first: descending exit distance e of
 *                           the ray from the box [box_min, box_max] (pass the
 *                           grid's bounds; per axis the far slab distance
 *                           fmaxf((min - o) / d, (max - o) / d), the smallest
 *                           of the three) in 65536 steps of the box diagonal:
 *                           key 65535 - min(65535, (uint)(max(e, 0) * 65535 /
 *                           diagonal)), diagonal = sqrtf((dx*dx + dy*dy) + dz*dz)
 *   SG_BOUNCE_AXIS_FIRST    ascending |d . axis| in 65536 steps: key
 *                           min(65535, (uint)(|d . axis| * 65535)); rays
 *                           running across `axis` first (height fields: up)
 *   Keys ascending, ties in input order; all arithmetic binary32. */
enum { SG_BOUNCE_CLASS_ROUNDS = 0, SG_BOUNCE_EXIT_FIRST = 1, SG_BOUNCE_AXIS_FIRST = 2 };
typedef struct {
    int order;
    float box_min[3], box_max[3];
    float axis[3];
} sg_bounce_order;
/* sg_bounce with the rays in the given order (same rays, same src mapping) */
sg_status sg_bounce_ordered(sg_ctx *ctx, sg_scene *scene, const sg_rays *rays, const uint32_t *d_prim,
                            const float *d_t, uint32_t n, const uint32_t *d_n, uint64_t seed,
                            const sg_bounce_order *order, const sg_rays_out *out, uint32_t *d_src,
                            uint32_t *d_count, void *stream);
/* sg_bounce that allocates nothing: no counterpart in the reference.  For
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same arguments and
 * results as sg_bounce; its scratch is a fixed device array of the library
 * (one per GPU, about 16 MiB, held until the process ends) instead of a
 * per-stream buffer, so n (the input capacity) is limited to
 * SG_BOUNCE_NOALLOC_MAX_RAYS; a larger n returns SG_ERR_SCRATCH_LIMIT and
 * enqueues nothing.  Every call on one GPU uses that GPU's one array: calls
 * whose device work overlaps (different streams, or graphs replayed
 * concurrently) race and must be ordered by the caller. */
#define SG_BOUNCE_NOALLOC_MAX_RAYS 16777216u
sg_status sg_bounce_noalloc(sg_ctx *ctx, sg_scene *scene, const sg_rays *rays, const uint32_t *d_prim,
                            const float *d_t, uint32_t n, const uint32_t *d_n, uint64_t seed,
                            const sg_rays_out *out, uint32_t *d_src, uint32_t *d_count, void *stream);

/* ---------------------------------------------------------- primitives --- */
/* Scratch of sg_bounce, sg_scan_u32 and sg_sort_pairs: the context keeps one
 * device buffer per stream, so calls on different streams never share
 * memory and calls on one stream are ordered by it.  A call that needs a
 * larger buffer than its stream has allocates one; under graph capture it
 * returns SG_ERR_INVALID_ARG instead (make the same call once on that
 * stream before capturing).  An outgrown buffer stays allocated until
 * sg_ctx_destroy, so a captured graph keeps valid memory; two replays of
 * graphs captured on one stream share that stream's buffer and must not
 * run at the same time. */
/* ExclusiveScan / InclusiveScan (src/Utils/Scan.h:35-68) on device arrays */
sg_status sg_scan_u32(sg_ctx *ctx, const uint32_t *d_in, uint32_t *d_out, uint32_t n, int inclusive,
                      void *stream);
/* Sort (src/Utils/Sort.h:33-48): stable sort of n (key,value) pairs by the
 * low key_bits bits of the key, in place on the device arrays */
sg_status sg_sort_pairs(sg_ctx *ctx, uint32_t *d_keys, uint32_t *d_vals, uint32_t n, int key_bits,
                        int radix_bits, void *stream);

/* -------------------------------------------------- device plumbing --- */
/* Memory, copies, streams and events for hosts that have no HIP binding of
 * their own (the Python mirror, other languages over the C ABI).  The
 * reference calls the CUDA runtime directly for these (for instance
 * cudaMalloc in src/Engine/StaticRTEngine.h:131,192); a C++ host may as
 * well use HIP itself and pass its pointers and streams here. */
sg_status sg_malloc(sg_ctx *ctx, size_t bytes, void **out);
sg_status sg_free(sg_ctx *ctx, void *p);
/* pinned host memory (page-locked): copies to and from it are DMA transfers
 * that overlap device work; serving clients keep request data in it */
sg_status sg_host_alloc(sg_ctx *ctx, size_t bytes, void **out);
sg_status sg_host_free(sg_ctx *ctx, void *p);
sg_status sg_memcpy_htod(sg_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
sg_status sg_memcpy_dtoh(sg_ctx *ctx, void *dst, const void *src, size_t bytes, void *stream);
sg_status sg_stream_create(sg_ctx *ctx, void **out);
sg_status sg_stream_destroy(sg_ctx *ctx, void *stream);
sg_status sg_stream_sync(sg_ctx *ctx, void *stream);
sg_status sg_event_create(sg_ctx *ctx, void **out);
sg_status sg_event_destroy(sg_ctx *ctx, void *ev);
sg_status sg_event_record(sg_ctx *ctx, void *ev, void *stream);
sg_status sg_event_elapsed_ms(sg_ctx *ctx, void *ev_start, void *ev_end, float *ms);
/* work enqueued on `stream` after this call waits for `ev` (a fork / join
 * between streams; under graph capture it becomes a graph dependency).  No
 * counterpart in the reference, which runs everything on one stream. */
sg_status sg_stream_wait_event(sg_ctx *ctx, void *stream, void *ev);

/* ------------------------------------------------------------ batching --- */
/* Many (grid, ray batch) items traced by one launch per grid kind: the
 * reference's trace kernel (src/Engine/RayTracingKernels.h:37-136, a
 * persistent pool of ray ids from a global counter, :76) generalised from one
 * ray buffer to a table of batches against different grids.  Wave w traces
 * the w-th 64 rays of the concatenated batches; grids of both kinds mix in
 * one table.  Results equal sg_trace_closest / sg_trace_shadow per item. */
typedef struct sg_batch sg_batch;
sg_status sg_batch_create(sg_ctx *ctx, uint32_t max_items, sg_batch **out);
sg_status sg_batch_destroy(sg_batch *batch);
/* drop every item, the uploaded table and the rebuild plan: until the next
 * sg_batch_commit, sg_batch_launch and sg_batch_rebuild do nothing */
sg_status sg_batch_clear(sg_batch *batch);
/* append one item: device SoA rays, n rays, device outputs (prim, t).  The
 * batch keeps the pointers: the grid, rays and outputs must outlive the
 * item (sg_batch_clear / sg_batch_destroy end it). */
sg_status sg_batch_add(sg_batch *batch, sg_grid *grid, const sg_rays *rays, uint32_t n, uint32_t *d_prim,
                       float *d_t);
/* upload the item table (stream-ordered; the host side may be changed after)
 * and the rebuild plan of the items' grids (sg_batch_rebuild) */
sg_status sg_batch_commit(sg_batch *batch, void *stream);
/* Rebuild every distinct grid of the committed items (sg_rebuild of each,
 * from the scene it was last built or rebuilt from and at its resolution,
 * as of sg_batch_commit).  The uniform grids are rebuilt together, one
 * launch per build phase for all of them: each launch covers the blocks of
 * every grid's own launch of that phase, so a few dozen launches replace a
 * dozen per grid.  Two-level grids are rebuilt on up to branches - 1
 * concurrent streams beside them, as sg_rebuild_many (branches 0 = the
 * maximum).  Every grid ends equal to its sg_rebuild.  Enqueues kernels only
 * and can be captured.  SG_ERR_INVALID_ARG when a grid was rebuilt from
 * another scene after the commit (commit again first).  No counterpart in
 * the reference, whose builders run
 * one scene at a time (FastBuildKernels.h:31-313, called from
 * FastSimpleGridBuilder.h:43-486). */
sg_status sg_batch_rebuild(sg_batch *batch, uint32_t branches, void *stream);
/* trace every committed item, closest hit (shadow = 0) or any hit (1).
 * Enqueues kernels only, so it can be captured with sg_graph_begin/end.
 * Each launch also records, per item, the status bits of the item's grid
 * as the traversal found them (a rebuild replayed in a graph that outgrew
 * its buffers is traversed as all misses). */
sg_status sg_batch_launch(sg_batch *batch, int shadow, void *stream);
/* after the launch (or graph replay) on `stream` has run: waits for the
 * stream and writes, for the first n committed items in sg_batch_add order,
 * SG_OK or the status of a grid that was traversed as all misses
 * (SG_ERR_CAPACITY for a rebuild that outgrew the grid's buffers).  No
 * counterpart in the reference, whose builders exit on errors. */
sg_status sg_batch_get_status(sg_batch *batch, void *stream, sg_status *out, uint32_t n);

/* --------------------------------------------------------------- graphs --- */
/* Stream capture into a replayable HIP graph: rebuilds (sg_rebuild),
 * single traces and sg_batch_launch enqueue kernels only and may be captured
 * (sg_build may not: it sizes the grid on the host). */
typedef struct sg_graph sg_graph;
sg_status sg_graph_begin(sg_ctx *ctx, void *stream);
sg_status sg_graph_end(sg_ctx *ctx, void *stream, sg_graph **out);
sg_status sg_graph_launch(sg_graph *graph, void *stream);
sg_status sg_graph_destroy(sg_graph *graph);

/* ------------------------------------------------------------ autotuner --- */
/* Picks grid density, sort radix width and traversal block size for one
 * scene by measurement: each candidate is built (rebuild timed on the
 * device) and traces a ray sample; the cost is build_ms + trace_ms scaled to
 * n_rays.  Coordinate search: density (kind in kind_mask: bit 0 uniform
 * lambda in {1,2,3,5,8}, bit 1 two-level lambda in {3,5,8} x leaf lambda in
 * {1.2,2,3}), then radix bits {4,6,8}, then block size {64,128,256}.  The
 * reference fixes lambda by hand (GridBuilder.h:48) and its SAHGridBuilder
 * searches a cost model (SAHGridBuilder.h:117-202); this measures instead. */
typedef struct {
    sg_build_params best;
    float build_ms, trace_ms, cost_ms;   /* of the chosen parameters          */
    uint32_t candidates;                 /* parameter sets measured           */
} sg_tune_result;
sg_status sg_autotune(sg_ctx *ctx, sg_scene *scene, const sg_rays *sample, uint32_t n_sample, uint64_t n_rays,
                      int kind_mask, void *stream, sg_tune_result *out);

/* -------------------------------------------------------------- serving --- */
/* A request loop for ray queries against grids kept resident on the device.
 * Requests (host SoA rays in, host prim/t out) are admitted while free slots
 * and HBM budget (device bytes for request rays and results) allow.
 * SG_SERVE_BATCHED: continuous batching -- every request that is admitted
 * while earlier waves run joins the next wave; a wave is one multi-batch
 * launch (as sg_batch_*) with its copies, up to max_waves waves are
 * in flight on separate streams, so copies and traversal of different waves
 * overlap.  SG_SERVE_PLAIN: the per-request path of the reference -- one
 * request at a time: copy in, one trace launch, copy out, synchronise.
 * Results are identical in both modes and to sg_trace_closest. */
enum { SG_SERVE_BATCHED = 0, SG_SERVE_PLAIN = 1 };
typedef struct {
    int mode;                /* SG_SERVE_BATCHED or SG_SERVE_PLAIN              */
    size_t hbm_budget;       /* device bytes for in-flight request data        */
    uint32_t max_requests;   /* requests admitted at once (slots)              */
    uint32_t max_waves;      /* waves in flight (streams), batched mode        */
    uint32_t max_wave_rays;  /* rays per wave (0: no cap)                       */
} sg_server_config;
typedef struct {
    uint64_t requests, rays, waves;
    double p50_us, p99_us, mean_us, max_us;   /* submit -> results in host memory */
    double wall_s;                            /* first submit -> last completion  */
    double mrays_per_s;                       /* rays / wall_s                    */
    uint64_t regrown;                         /* dynamic rebuilds that outgrew their
                                                 grid and were built again (not reset) */
} sg_server_stats;
typedef struct sg_server sg_server;
sg_status sg_server_create(sg_ctx *ctx, const sg_server_config *cfg, sg_server **out);
/* finishes every admitted request, then stops */
sg_status sg_server_destroy(sg_server *server);
/* a resident grid (built by the caller; it must outlive the server) */
sg_status sg_server_add_grid(sg_server *server, sg_grid *grid, uint32_t *scene_id);
/* a dynamic scene: the server keeps two copies of the mesh and two grids and
 * rebuilds the back copy on its own stream (sg_server_update_dynamic) while
 * waves keep running on the front one (double buffering,
 * DynamicRTEngine::renderFrame, src/Engine/DynamicRTEngine.h:73-158, with
 * FastTwoLevelGridBuilder::rebuild at fixed resolution).  update_dynamic
 * returns once the new grid is in front: requests submitted after it see the
 * new mesh, requests admitted before it finish on the old one.  A rebuild
 * that outgrows the back grid's buffers (a mesh deformed past the capacity
 * slack) is built again with new buffers before the swap; a mesh that
 * cannot be gridded at all returns its status and the front grid stays. */
sg_status sg_server_add_dynamic(sg_server *server, const float *const soa9[9], uint32_t n,
                                const sg_build_params *params, uint32_t *scene_id);
sg_status sg_server_update_dynamic(sg_server *server, uint32_t scene_id, const float *const soa9[9]);
/* host rays7 (7 arrays of n floats; one contiguous 7 x n block is copied
 * with one transfer) in; prim/t written to host memory when the request
 * completes.  The server copies straight from and to these buffers, which
 * must stay valid until then; pinned buffers (sg_host_alloc) make the copies
 * asynchronous.  SG_ERR_BUDGET if the request can never fit the budget. */
sg_status sg_server_submit(sg_server *server, uint32_t scene_id, const float *const rays7[7], uint32_t n,
                           uint32_t *prim_out, float *t_out, uint64_t *request_id);
/* wait for one request; latency_us (optional) = submit -> completion.
 * Returns the request's status: SG_ERR_CAPACITY (or the grid's other
 * failure status) when its grid was traversed as all misses because its
 * last build failed or outgrew the buffers (for example a caller's rebuild
 * of a grid added with sg_server_add_grid). */
sg_status sg_server_wait(sg_server *server, uint64_t request_id, double *latency_us);
/* sg_server_wait that also reports which mesh of the request's scene it was
 * traced against: 0 for a static scene and for the mesh a dynamic scene was
 * added with, k for the mesh of the k-th successful sg_server_update_dynamic
 * of that scene.  No counterpart in the reference (one mesh per frame). */
sg_status sg_server_wait_version(sg_server *server, uint64_t request_id, double *latency_us, uint64_t *version);
/* wait for every submitted request */
sg_status sg_server_drain(sg_server *server);
sg_status sg_server_get_stats(sg_server *server, sg_server_stats *out);
sg_status sg_server_reset_stats(sg_server *server);

/* ---------------------------------------------------------- generators --- */
/* Seeded synthetic inputs (host memory), see shinobi_amd/csrc/sg_gen.cpp.
 * scene kinds: 0 = uniform-in-cube soup (n triangles, edge scale e),
 * 1 = height field (grid x grid vertices), 2 = dense cluster + large tris. */
sg_status sg_gen_soup(uint64_t seed, uint32_t n, float edge, float *const soa9[9]);
sg_status sg_gen_heightfield(uint64_t seed, uint32_t grid, float amplitude, float *const soa9[9]);
float sg_heightfield_height(uint64_t seed, float x, float y, float amplitude);
sg_status sg_gen_cluster(uint64_t seed, uint32_t n_small, uint32_t n_large, float *const soa9[9]);
/* pinhole primary rays with the reference Camera (src/Primitive/Camera.h:41-70)
 * and RegularPixelSampler (src/Primitive/PixelSampler.h:34-54), one per pixel
 * centre; rays7 = 7 host arrays of res_x*res_y floats */
sg_status sg_gen_camera_rays(const float eye[3], const float fwd[3], const float up[3], float fov_deg,
                             uint32_t res_x, uint32_t res_y, float *const rays7[7]);
/* rays with origins uniform in [lo,hi]^3 and directions uniform on the sphere */
sg_status sg_gen_random_rays(uint64_t seed, uint32_t n, float lo, float hi, float *const rays7[7]);

/* ---------------------------------------------------------- scene files --- */
/* Wavefront OBJ geometry, replacing FWObject::read (src/Loaders/FWObjectReader.cpp:181-502)
 * as far as it decides triangles: accepted / rejected lines, (float)strtod
 * coordinates, 1-based strtol indices, fan triangulation (0,k,k+1) of polygons.
 * Rejected lines are skipped and counted, as the reference skips them.
 * SG_ERR_IO if the file cannot be read; SG_ERR_INVALID_SCENE if a face index
 * lies outside the vertex list. Host-only; the triangles feed sg_scene_create
 * (FaceSoup::upload, src/Control/Structure/FaceSoup.h:171-247). */
typedef struct sg_obj sg_obj;
typedef struct sg_obj_info {
    uint32_t n_vertices;   /* accepted 'v' lines */
    uint32_t n_faces;      /* accepted face lines */
    uint32_t n_triangles;  /* after fan triangulation */
    uint32_t n_bad_lines;  /* lines the reader rejects */
} sg_obj_info;
sg_status sg_obj_load(const char *path, sg_obj **out, sg_obj_info *info);
/* triangle i = (v[a], v[b], v[c]) of its face into 9 host arrays of n_triangles floats */
sg_status sg_obj_triangles(const sg_obj *obj, float *const soa9[9]);
void sg_obj_destroy(sg_obj *obj);

/* camera.cfg, replacing CameraManager::read (src/Loaders/CameraManager.cpp:93-165):
 * position / orientation / up / rotation / resX / resY / FOV keyword lines;
 * lines holding '#' are skipped. Feed sg_gen_camera_rays(position, orientation, up, fov_deg, res_x, res_y). */
typedef struct sg_camera {
    float position[3], orientation[3], up[3], rotation[3];
    uint32_t res_x, res_y;
    float fov_deg;
} sg_camera;
sg_status sg_camera_load(const char *path, sg_camera *out);

#ifdef __cplusplus
}
#endif
#endif
