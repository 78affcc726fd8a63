/*
 * gpurt.h -- C ABI of the MI355X (gfx950) geometric-query library.
 *
 * Drop-in for the BVH build and traversal path of TheNumbat/GPU-RT:
 *   - BVH build            replaces VK::Accel(const Mesh&) / Accel::recreate(const Mesh&)
 *                          (src/vk/vulkan.h:256-284, src/vk/vulkan.cpp:881-936), which
 *                          calls vkCmdBuildAccelerationStructuresKHR (vulkan.cpp:877).
 *   - closest-hit rays     replaces trace_ray() -> traceRayEXT + rt.rchit / rt.rmiss
 *                          (src/shaders/rt/rt.rgen:257-270, rt.rchit:11-16, rt.rmiss:9-11).
 *   - any-hit rays         replaces visibility() (rt.rgen:272-291).
 *   - closest-point query  the FCPW-style query the reference's README (README.md:6-8)
 *                          points to; same BVH handle.
 *   - backward passes      gradients of unsigned distance and hit distance t.
 *
 * Conventions
 *   - Every array pointer is a caller-owned DEVICE pointer (hipMalloc or torch tensor
 *     storage on the same device as the context).  BVH memory is library-owned.
 *   - Calls are asynchronous on `stream` (a hipStream_t; NULL = the legacy default
 *     stream) except gpurt_bvh_build / gpurt_bvh_refit, which wait for the build like
 *     the reference's fence wait (vulkan.cpp:1205-1206) and return a complete handle.
 *   - Status: 0 = OK, negative = error (see GPURT_E_*); nothing calls exit().  The
 *     message of the last error on the calling thread is gpurt_last_error().
 *   - A built BVH is immutable: concurrent queries from several threads/streams are
 *     safe; a refit or destroy must not overlap queries on the same handle.
 *   - Triangles are indexed uint32 triples; vertex positions are 3 floats at the start
 *     of each vertex record of `vstride_bytes` bytes (12, 16, or 48 for the reference's
 *     Mesh::Vertex {Vec4 pos, norm, tang}, src/vk/mesh.h:16-22).
 *   - Primitive ids are the triangle index within its mesh (gl_PrimitiveID,
 *     rt.rchit:15).  Ties (equal distance or equal t) go to the LOWEST id.
 *   - Barycentrics of a ray hit are (u,v) with b = (1-u-v, u, v) (rt.rchit:13); the
 *     closest-point query writes all three (b0,b1,b2) of the closest point.
 *   - Bad input (the reference calls die() on errors, log.h:46-49; here a status or a
 *     defined answer):  a build or refit over a non-finite position of a vertex that a
 *     triangle uses (refit: any vertex) returns GPURT_E_INVALID (a refit then leaves the
 *     tree unchanged).  A ray with a non-finite origin or direction component, a zero
 *     direction or a NaN tmin / tmax is a miss (prim -1, t +inf, uv 0; not occluded).  A
 *     closest-point query with a non-finite coordinate finds no triangle: prim -1, dist
 *     +inf, pt and bary 0.
 */
#ifndef GPURT_H
#define GPURT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define GPURT_API __attribute__((visibility("default")))
#else
#define GPURT_API
#endif

#define GPURT_OK 0
#define GPURT_E_INVALID (-1)        /* bad argument: null pointer, bad size, bad index   */
#define GPURT_E_OOM (-2)            /* device allocation failed                          */
#define GPURT_E_STACK_OVERFLOW (-3) /* traversal stack exhausted (results incomplete)    */
#define GPURT_E_DEVICE (-4)         /* HIP runtime error                                 */
#define GPURT_E_UNSUPPORTED (-5)    /* option value not supported                        */

typedef struct gpurt_ctx_s* gpurt_ctx;
typedef struct gpurt_bvh_s* gpurt_bvh;
typedef struct gpurt_batch_s* gpurt_batch;

/* One ray, 32 bytes: origin, tmin, direction, tmax.  A hit must have t in [tmin, tmax].
 * The reference's rays use tmin = EPS = 1e-5 and tmax = LARGE_DIST = 1e7 for closest
 * hits (rtcommon.glsl:6-7, rt.rgen:257-270) and tmax = |b-a| - EPS for visibility
 * (rt.rgen:272-291). */
typedef struct gpurt_ray {
    float o[3];
    float tmin;
    float d[3];
    float tmax;
} gpurt_ray;

typedef struct gpurt_build_opts {
    int32_t node_width;     /* children per node: 2, 4 or 8 (0 = 4)                  */
    int32_t leaf_size;      /* max triangles per leaf, 1..8 (0 = 4)                   */
    int32_t treelet_size;   /* leaves per treelet for SAH restructuring, 0 = off, 5..7 */
    int32_t treelet_rounds; /* bottom-up restructuring rounds: 0 or 1 (0 = 1 when on) */
    int32_t node_format;    /* traversed node encoding (answers are identical for all):
                               1 = FP32 child boxes (32*W bytes per node),
                               2 = 8-bit child boxes relative to the node's box
                                   (48/64/96 bytes for W = 2/4/8),
                               3 = both: 8-bit for rays, FP32 for closest points,
                               0 = library default (3)                                */
} gpurt_build_opts;

typedef struct gpurt_query_opts {
    int32_t stack_depth;  /* per-lane LDS stack entries: 8, 16, 32 or 64 (0 = library
                             default); deeper entries spill to a device buffer         */
    int32_t waves_per_cu; /* resident waves per CU for the launch, 0 = library choice  */
    uint64_t* stats;      /* optional DEVICE uint64[4], accumulated: wide nodes visited,
                             triangles tested, stack entries spilled past the LDS
                             part, reserved (slower counting kernel variant)          */
    int32_t leaf_trigger; /* lanes (of a 64-lane wave) holding a postponed leaf at which
                             the wave switches to leaf tests, 1..64 (0 = library
                             default); changes speed only                             */
    int32_t refill_trigger; /* idle lanes (of 64) at which a wave of the ray kernels
                             writes its finished results and fetches new rays, 1..64
                             (0 = library default); changes speed only              */
} gpurt_query_opts;

typedef struct gpurt_bvh_info {
    int64_t n_tris;
    int64_t n_verts;
    int64_t n_nodes;    /* wide nodes                                                  */
    int64_t node_bytes; /* bytes per wide node                                         */
    int64_t tri_bytes;  /* bytes per packed triangle                                   */
    int32_t depth;      /* wide-node levels                                            */
    int32_t node_width;
    int32_t leaf_size;
    float sah_cost;     /* SAH cost of the wide tree, relative to the root area        */
    int32_t node_format; /* resolved gpurt_build_opts::node_format (1, 2 or 3)          */
    int32_t cpq_node_bytes; /* bytes per node traversed by closest-point queries
                               (node_bytes: by rays)                                    */
} gpurt_bvh_info;

/* ---- context --------------------------------------------------------------------- */

/* Create a context on HIP device `device`.  Replaces the ray-tracing part of
 * VK::Manager::init / init_rt (src/vk/vulkan.cpp:1136-1160, 1326-1385).
 * Query and backward calls need device scratch (work-queue counters, the traversal
 * stacks' spill area, the fixed-order sums' term buffers).  The context keeps one scratch
 * buffer per stream that calls use in turn, grown to the largest call on that stream
 * (growing waits for the device once) and released by gpurt_ctx_destroy, so a call does
 * not allocate and returns without waiting for the device.  A call on a stream being
 * captured into a HIP graph, or on a stream whose buffer another host thread's call
 * holds, takes a stream-ordered allocation of its own instead. */
GPURT_API int gpurt_ctx_create(int device, gpurt_ctx* out);
/* Waits for the device, then frees the context (which must outlive its handles). */
GPURT_API void gpurt_ctx_destroy(gpurt_ctx ctx);

/* Conditions the context's kernels report as they run (sticky bits, set on the device):
 *   GPURT_STATUS_STACK_OVERFLOW        a traversal stack ran out (that launch's answers
 *                                      are incomplete); the library sizes every stack for
 *                                      the deepest tree, so this does not happen
 *   GPURT_STATUS_TLAS_OUTSIDE_DOMAIN   a ray traced through a TLAS started more than 2^8
 *                                      scene extents from the origin, outside the domain
 *                                      where two-level answers are exact (see
 *                                      gpurt_tlas_build); it was answered anyway
 * gpurt_ctx_status waits for the device, returns the bits in *flags and clears them when
 * `clear` is non-zero. */
#define GPURT_STATUS_STACK_OVERFLOW 1u
#define GPURT_STATUS_TLAS_OUTSIDE_DOMAIN 2u
GPURT_API int gpurt_ctx_status(gpurt_ctx ctx, uint32_t* flags, int clear);

/* How the backward passes (gpurt_closest_point_bwd, gpurt_trace_bwd, gpurt_batch_set_grads)
 * add the per-query vertex terms (three per query or ray) into the vertex gradients:
 *   GPURT_SUMS_FIXED_ORDER  the terms are sorted by vertex (a stable radix sort) and each
 *                           vertex's terms added in a fixed order (a tree over query
 *                           order): the same bits on every run; 144 bytes of
 *                           scratch per query (the context's per-stream buffer)
 *   GPURT_SUMS_ATOMIC       FP32 atomic adds, one pass; the summation order, and so the
 *                           last bits, can vary from run to run
 *   GPURT_SUMS_AUTO         (the default) whichever is faster for the call: the fixed
 *                           order from 2^21 terms (~700 K queries; the atomics' device-wide
 *                           rate then costs more than the sort), atomics below
 * Applies to calls made after it; a batch reads it in gpurt_batch_set_grads.  An unknown
 * mode: GPURT_E_INVALID.  No counterpart in the reference (it has no gradients). */
#define GPURT_SUMS_AUTO 0
#define GPURT_SUMS_FIXED_ORDER 1
#define GPURT_SUMS_ATOMIC 2
GPURT_API int gpurt_ctx_set_vertex_sums(gpurt_ctx ctx, int mode);

/* ---- build ----------------------------------------------------------------------- */

/* Build a BVH over one triangle mesh: 30-bit Morton LBVH (Karras 2012) -> optional SAH
 * treelet refinement -> collapse to `node_width`-wide nodes -> packed node and triangle
 * arrays.  Replaces VK::Accel::Accel(const Mesh&) / Accel::recreate(const Mesh&)
 * (src/vk/vulkan.h:258,266; src/vk/vulkan.cpp:881-936: R32G32B32_SFLOAT at
 * vertexStride, UINT32 indices, primitiveCount = n_idx/3, opaque geometry). */
GPURT_API int gpurt_bvh_build(gpurt_ctx ctx, const float* verts, size_t vstride_bytes, size_t n_verts,
                    const uint32_t* tri_idx, size_t n_tris, const gpurt_build_opts* opts,
                    void* stream, gpurt_bvh* out);

/* Update the boxes of a built BVH for moved vertices (same count, same topology).  The
 * reference only rebuilds (mode BUILD, vulkan.cpp:832,915); this is the cheaper update
 * for optimisation loops that move vertices every step. */
GPURT_API int gpurt_bvh_refit(gpurt_bvh bvh, const float* verts, size_t vstride_bytes, void* stream);

/* Free a BVH.  Replaces Accel::destroy (vulkan.cpp:768-776) and its deferred Drop
 * (vulkan.h:305-336): the memory is released in stream order after the work of every
 * call that used the handle (build, refit, queries, backward, TLAS builds and queries
 * over it), without a host wait. */
GPURT_API void gpurt_bvh_destroy(gpurt_bvh bvh);

GPURT_API int gpurt_bvh_get_info(gpurt_bvh bvh, gpurt_bvh_info* out);

/* Caller-owned DEVICE buffers for gpurt_bvh_export (each nullable). */
typedef struct gpurt_bvh_arrays {
    void* nodes;          /* n_nodes * 32 * W bytes: the FP32 wide nodes                   */
    void* qnodes;         /* n_nodes * (48, 64, 96 for W = 2, 4, 8) bytes: the 8-bit nodes
                             (only for node formats 2 and 3)                              */
    float* tris;          /* n_tris * 12 floats: the packed triangles                      */
    uint32_t* leaf_order; /* n_tris: packed position -> triangle index in the mesh         */
} gpurt_bvh_arrays;

/* Copy the built tree into the caller's device buffers (asynchronous on `stream`), for
 * inspection, invariant checks and serialization.  The reference's acceleration
 * structure is opaque (vkGetAccelerationStructureBuildSizesKHR, vulkan.cpp:921-928);
 * this one is laid out as follows.  Node 0 is the root; nodes are stored level by level.
 *   FP32 node (W children):  float lo_x[W], lo_y[W], lo_z[W], hi_x[W], hi_y[W], hi_z[W];
 *                            uint32 child[W]; uint32 count[W] (triangles under the child)
 *   child word: internal node index (< 2^31); leaf: 0x80000000 | (count-1) << 28 | first
 *               packed triangle (count 1..8); empty slot: 0xFFFFFFFF
 *   8-bit node: float origin[3]; W = 2, 4: float scale[3]; W = 8: uint32 exponents (byte
 *               a = e_a, scale 2^(e_a - 127)); uint32 child[W]; uint8 q[6][W] (lo x, y, z,
 *               hi x, y, z), zero-padded; child j's box is fmaf(q, scale, origin) per
 *               plane (scales are powers of two); an empty slot has lo = 255, hi = 0
 *   packed triangle: float4 (a.xyz, id bits), float4 (b.xyz, 0), float4 (c.xyz, 0) -- a, b,
 *               c the mesh's vertices in index order, id the triangle index
 * Child boxes are inflated by 2^-18 of the largest coordinate magnitude, and every 8-bit
 * box contains its FP32 box. */
GPURT_API int gpurt_bvh_export(gpurt_bvh bvh, const gpurt_bvh_arrays* out, void* stream);

/* ---- queries ---------------------------------------------------------------------- */

/* Closest point on the mesh for n query points q (n x 3 floats).  Outputs: prim (int32,
 * -1 only for an empty mesh), pt (n x 3 closest point), dist (unsigned distance), bary
 * (n x 3 barycentrics of pt, nullable).  prim is the FP32 argmin (lowest id on ties);
 * pt, dist and bary are that triangle's closest point computed in FP64 and rounded once
 * to FP32 (a second, gather-only pass).  No counterpart in the mounted reference (its
 * closest-point work is on a branch not included, README.md:6-8); it uses the same
 * handle as the ray queries. */
GPURT_API int gpurt_closest_point(gpurt_bvh bvh, const float* q, size_t n, int32_t* prim, float* pt,
                        float* dist, float* bary, const gpurt_query_opts* opts, void* stream);

/* Closest hit for n rays.  prim = hit triangle (-1 on miss, rt.rmiss:9-11), t = hit
 * distance (+inf on miss), uv = (u,v) (n x 2, nullable; 0 on miss).  Replaces
 * trace_ray()'s traceRayEXT(TLAS, gl_RayFlagsOpaqueEXT, 0xFF, ...) with the rt.rchit
 * payload (rt.rgen:257-270, rt.rchit:11-16, rtcommon.glsl:55-60). */
GPURT_API int gpurt_trace_closest(gpurt_bvh bvh, const gpurt_ray* rays, size_t n, int32_t* prim,
                        float* t, float* uv, const gpurt_query_opts* opts, void* stream);

/* gpurt_trace_closest without any device allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory -- code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc.  Same arguments, results
 * identical bit for bit to gpurt_trace_closest.  The traversal's scratch (work-queue
 * counters and the stack spill area) is a fixed 64 MiB array per device, held from the
 * first gpurt_ctx_create on that device until the process ends; the launch is fitted into
 * it (a tree whose worst-case stack spills more than it holds runs with fewer resident
 * waves), and GPURT_E_INVALID is returned when not even one block's spill fits (a tree
 * thousands of levels deep) or n exceeds 2^32 - 1025.  The counters are cleared by a
 * kernel on `stream` (no memset node).  Every call on one device shares that array:
 * calls that may overlap on one device (different streams, or host threads) race and
 * must be serialized by the caller.  opts->stats must be null. */
GPURT_API int gpurt_trace_closest_noalloc(gpurt_bvh bvh, const gpurt_ray* rays, size_t n, int32_t* prim,
                                          float* t, float* uv, const gpurt_query_opts* opts, void* stream);

/* Any hit (occlusion) for n rays: occluded[i] = 1 iff some triangle is hit with t in
 * [tmin, tmax].  Replaces visibility()'s traceRayEXT with TerminateOnFirstHit |
 * Opaque | SkipClosestHitShader and the payload.hit preset/clear
 * (rt.rgen:272-291, rt.rmiss:9-11). */
GPURT_API int gpurt_trace_any(gpurt_bvh bvh, const gpurt_ray* rays, size_t n, uint8_t* occluded,
                    const gpurt_query_opts* opts, void* stream);

/* ---- backward --------------------------------------------------------------------- */

/* Gradient of sum_i g_dist[i] * dist_i.  Uses the forward pass's saved prim (no
 * re-traversal): the closest point p_i of that triangle, its barycentrics b and q_i - p_i
 * are recomputed in FP64 (the computation the forward pass reports from), then
 * g_q[i] = g_dist[i] (q_i - p_i)/|q_i - p_i|; for the three vertices v_j of prim_i,
 * g_verts[v_j] -= b_j g_q[i]  (g_verts is ACCUMULATED, n_verts x 3: zero it first).
 * Vertex sums use FP32 atomics or a fixed order (gpurt_ctx_set_vertex_sums).  g_q and
 * g_verts are each
 * nullable; so are the saved bary, pt and dist (accepted for the forward's output
 * layout, not read).  A query on the surface (distance 0) contributes zero. */
GPURT_API int gpurt_closest_point_bwd(gpurt_bvh bvh, const float* q, size_t n, const int32_t* prim,
                            const float* bary, const float* pt, const float* dist,
                            const float* g_dist, float* g_q, float* g_verts, void* stream);

/* Gradient of sum_i g_t[i] * t_i for closest hits.  With n = (v1-v0) x (v2-v0):
 * dt/do = -n/(n.d), dt/dd = -t n/(n.d), dt/dv_j = b_j n/(n.d), b = (1-u-v,u,v).
 * g_origin / g_dir are n x 3 (written, nullable); g_verts is ACCUMULATED (nullable), with
 * atomics or in a fixed order (gpurt_ctx_set_vertex_sums).  Misses contribute
 * zero. */
GPURT_API int gpurt_trace_bwd(gpurt_bvh bvh, const gpurt_ray* rays, size_t n, const int32_t* prim,
                    const float* t, const float* uv, const float* g_t, float* g_origin,
                    float* g_dir, float* g_verts, void* stream);

/* ---- batched ---------------------------------------------------------------------- */

/* One mesh of a batch, with its own query points and rays (all device pointers). */
typedef struct gpurt_batch_mesh {
    const float* verts;
    size_t vstride_bytes;
    size_t n_verts;
    const uint32_t* tri_idx;
    size_t n_tris;
    const float* q;       /* n_q x 3 closest-point queries (nullable if n_q == 0) */
    size_t n_q;
    int32_t* q_prim;      /* outputs of the closest-point query                   */
    float* q_pt;
    float* q_dist;
    float* q_bary;        /* nullable                                             */
    const gpurt_ray* rays; /* n_r rays, closest hit                               */
    size_t n_r;
    int32_t* r_prim;
    float* r_t;
    float* r_uv;          /* nullable                                             */
} gpurt_batch_mesh;

/* Prepare a batch over n_meshes meshes: validates, allocates every buffer the build and
 * query stages need, and records them (BVH builds of all meshes in shared launches, then
 * one persistent closest-point + closest-hit launch fed by a device work queue) into a
 * HIP graph.  The mesh and query arrays must stay valid while the batch is used. */
GPURT_API int gpurt_batch_create(gpurt_ctx ctx, const gpurt_batch_mesh* meshes, size_t n_meshes,
                       const gpurt_build_opts* bopts, const gpurt_query_opts* qopts,
                       gpurt_batch* out);
/* Run the batch (graph launch on `stream`): rebuild every BVH and answer every query. */
GPURT_API int gpurt_batch_run(gpurt_batch batch, void* stream);

/* Gradient buffers of one batch mesh (all device pointers, each nullable). */
typedef struct gpurt_batch_grad {
    const float* g_dist;  /* n_q upstream gradients of q_dist (null: no distance term)   */
    float* g_q;           /* n_q x 3 gradient w.r.t. the query points                    */
    const float* g_t;     /* n_r upstream gradients of r_t (null: no hit-distance term)   */
    float* g_origin;      /* n_r x 3 gradient w.r.t. the ray origins                      */
    float* g_dir;         /* n_r x 3 gradient w.r.t. the ray directions                   */
    float* g_verts;       /* n_verts x 3 gradient w.r.t. the vertex positions, overwritten
                             with the sum of the distance and hit-distance terms          */
} gpurt_batch_grad;

/* Record the batched backward pass (one gpurt_batch_grad per mesh, in the order of
 * gpurt_batch_create) into a second HIP graph: it clears the vertex gradients, then
 * forms every mesh's gradients from the outputs of the last gpurt_batch_run (saved ids,
 * closest-point triangles recomputed in FP64, hit t, (u,v)) without re-traversal; the
 * vertex terms of all meshes are added with FP32 atomics or, by the context's
 * gpurt_ctx_set_vertex_sums mode when this is called (the default picks it for batches
 * of 2^21 terms and more), sorted by (mesh, vertex) and each vertex's terms added in a
 * fixed order (the same bits on every run).  Vertex gradients of hit t need the
 * forward's r_uv.  The fixed-order sum's term buffers (allocated here) take 144 bytes
 * per query or ray slot (3 terms of 48 bytes; slots are allotted in 64-item chunks per
 * mesh): about 2.7 GB for 19 M queries and rays.
 * No counterpart in the reference (it has no gradients); it is the batched form of
 * gpurt_closest_point_bwd + gpurt_trace_bwd.  Calling it again replaces the buffers. */
GPURT_API int gpurt_batch_set_grads(gpurt_batch batch, const gpurt_batch_grad* grads);
/* Run the recorded backward graph on `stream` (after gpurt_batch_run on the same stream). */
GPURT_API int gpurt_batch_run_backward(gpurt_batch batch, void* stream);
/* Waits for the device (every stream), then frees the batch and its recorded graphs. */
GPURT_API void gpurt_batch_destroy(gpurt_batch batch);

/* ---- autotuner ------------------------------------------------------------------------ */

/* Candidate values per parameter (each list nullable: its default is used). */
typedef struct gpurt_tune_space {
    const int32_t* node_width;   int32_t n_node_width;   /* default {2, 4, 8}            */
    const int32_t* stack_depth;  int32_t n_stack_depth;  /* default {8, 16, 32, 64}      */
    const int32_t* waves_per_cu; int32_t n_waves_per_cu; /* default {0, 16, 24}; 0 = the
                                                            library's choice             */
    const int32_t* treelet_size; int32_t n_treelet_size; /* default {0, 6}              */
    const int32_t* node_format;  int32_t n_node_format;  /* default {1, 2}              */
    int32_t reps;              /* timed runs per candidate, median kept (0 = 3)          */
    float builds_per_run;      /* gpurt_tune_queries: objective = query time + this many
                                  builds (0: built once, queried many times)             */
    const char* cache_path;    /* nullable: a text file of tuned results; a stored result
                                  for the same key is returned without timing anything,
                                  a new one is appended                                  */
} gpurt_tune_space;

typedef struct gpurt_tune_result {
    gpurt_build_opts build;    /* the chosen build options                               */
    gpurt_query_opts query;    /* the chosen query options (stats null)                  */
    float best_ms;             /* median time of the chosen candidate (objective)        */
    float default_ms;          /* the same with every option at its library default      */
    int32_t candidates;        /* candidates timed (0 when taken from the cache)         */
    int32_t from_cache;
    char key[128];             /* cache key: device, workload kind, log2 size buckets    */
} gpurt_tune_result;

/* Autotuner of the batched form: every candidate of `space` (node width x stack depth x
 * waves per CU x treelet size x node format) is a full gpurt_batch_create + timed
 * gpurt_batch_run replays of the caller's own meshes and queries on `stream` (HIP
 * events), and the fastest is returned, with the time of the library defaults beside
 * it.  Answers do not depend on any of these parameters (the traversal is exact for any
 * tree), only speed.  The meshes' output arrays are written by the trial runs.  The
 * reference's knobs of this kind are host fields (src/vk/rt.h:38-53). */
GPURT_API int gpurt_tune_batch(gpurt_ctx ctx, const gpurt_batch_mesh* meshes, size_t n_meshes,
                               const gpurt_tune_space* space, void* stream, gpurt_tune_result* out);

/* Autotuner of the plain path for one mesh: builds the BVH for every build candidate and
 * times gpurt_closest_point over q (n_q, nullable) plus gpurt_trace_closest over rays
 * (n_r, nullable) for every query candidate; objective = query time + builds_per_run x
 * build time (host wall time of gpurt_bvh_build). */
GPURT_API int gpurt_tune_queries(gpurt_ctx ctx, const float* verts, size_t vstride_bytes, size_t n_verts,
                                 const uint32_t* tri_idx, size_t n_tris, const float* q, size_t n_q,
                                 const gpurt_ray* rays, size_t n_r, const gpurt_tune_space* space,
                                 void* stream, gpurt_tune_result* out);

/* The tuned-result cache: one line per result, "key<TAB>node_width leaf_size treelet_size
 * treelet_rounds node_format stack_depth waves_per_cu leaf_trigger refill_trigger best_ms
 * default_ms"; the last line of a key wins.  Load returns GPURT_E_INVALID when the key is
 * absent. */
GPURT_API int gpurt_tune_cache_load(const char* path, const char* key, gpurt_tune_result* out);
GPURT_API int gpurt_tune_cache_store(const char* path, const gpurt_tune_result* result);

/* ---- two-level instancing (TLAS) ----------------------------------------------------- */

typedef struct gpurt_tlas_s* gpurt_tlas;

/* One instance: a BLAS placed by an object-to-world transform.  Mirrors
 * VkAccelerationStructureInstanceKHR as filled by Accel::recreate(blas, T)
 * (src/vk/vulkan.cpp:777-856): transform = the row-major 3x4 VkTransformMatrixKHR
 * (T[i].T(), vulkan.cpp:795-796), custom_index = instanceCustomIndex (i, line 798; reported
 * as a hit's object id, gl_InstanceCustomIndexEXT in rt.rchit:14), mask = 8-bit visibility
 * mask (0xFF there; rays use cull mask 0xFF as traceRayEXT does in rt.rgen:258,278, so an
 * instance with mask 0 is never hit).  Face culling is off for every instance
 * (TRIANGLE_FACING_CULL_DISABLE, line 802), as for BLAS queries. */
typedef struct gpurt_instance {
    float transform[12];
    uint32_t custom_index;
    uint32_t mask;
    gpurt_bvh blas;
} gpurt_instance;

/* Build a TLAS over n instances (replaces Accel(const std::vector<Drop<Accel>>&,
 * const std::vector<Mat4>&), vulkan.h:259, and its recreate, vulkan.cpp:777-856).  Every
 * BLAS must have the same node width and ray node format; the BLASes must outlive the
 * TLAS, and a BLAS refit requires rebuilding the TLAS (as a Vulkan TLAS build does).
 * Rays are traversed in world space through the TLAS and in object space inside an
 * instance (o' = M^-1 (o - c), d' = M^-1 d with the FP32 rounding of the FP64 inverse of
 * the 3x3 part M; c the translation), so t is the same parameter in both spaces.
 * Results are exact (equal to FP32 brute force over all instances, ties to the lowest
 * instance index, then the lowest primitive id) for ray origins whose coordinates are at
 * most 2^8 times the largest |coordinate| of the instances' world boxes (a ray starting
 * farther out is answered and sets GPURT_STATUS_TLAS_OUTSIDE_DOMAIN, gpurt_ctx_status);
 * transforms must be finite and invertible (else GPURT_E_INVALID).  Returns when the
 * TLAS is complete. */
GPURT_API int gpurt_tlas_build(gpurt_ctx ctx, const gpurt_instance* instances, size_t n,
                               const gpurt_build_opts* opts, void* stream, gpurt_tlas* out);
/* Free a TLAS: like gpurt_bvh_destroy, its memory is released in stream order after the
 * work of every call that used it, without a host wait. */
GPURT_API void gpurt_tlas_destroy(gpurt_tlas tlas);

/* Closest hit through the TLAS: per ray obj (custom_index of the hit instance, -1 on a
 * miss), prim (triangle index in that instance's mesh, -1 on a miss), t, uv (nullable).
 * Replaces traceRayEXT(TLAS, ...) in trace_ray (rt.rgen:257-270) with the payload of
 * rt.rchit:11-16 / rt.rmiss:9-11. */
GPURT_API int gpurt_tlas_trace_closest(gpurt_tlas tlas, const gpurt_ray* rays, size_t n, int32_t* obj,
                                       int32_t* prim, float* t, float* uv, const gpurt_query_opts* opts,
                                       void* stream);
/* Any hit through the TLAS (visibility, rt.rgen:272-291): occluded[i] = 1 if some
 * instance has a hit with t in [tmin, tmax]. */
GPURT_API int gpurt_tlas_trace_any(gpurt_tlas tlas, const gpurt_ray* rays, size_t n, uint8_t* occluded,
                                   const gpurt_query_opts* opts, void* stream);

/* ---- wavefront path tracer ----------------------------------------------------------- */

typedef struct gpurt_scene_s* gpurt_scene;

/* One object of a rendered scene: the per-object record of RTPipe::build_desc
 * (Scene_Desc, src/vk/rt.h:67-78, filled at src/vk/rt.cpp:26-45) with its mesh
 * (Mesh::Vertex records, src/scene/mesh.h:16-22).  Textures are not supported: the
 * albedo / emissive / metal_rough factors are used everywhere (rt.rgen:97-133 with every
 * *_tex index -1).  Every object whose emissive factor is non-zero is a light, in object
 * order (rt.cpp:47-60). */
typedef struct gpurt_object {
    const float* verts;        /* device, n_verts x 12 floats: pos.xyz, u, normal.xyz, v, tangent.xyzw */
    const uint32_t* indices;   /* device, n_tris x 3 */
    uint32_t n_verts;
    uint32_t n_tris;
    float model[16];           /* row-major object-to-world (scale * pose), affine */
    float albedo[3];
    float emissive[3];
    float metal_rough[2];      /* metalness, roughness */
} gpurt_object;

/* Create a render scene over a TLAS whose instance custom indices address `objects`
 * (custom index i = objects[i], as GPURT::build_accel and build_desc number them,
 * gpurt.cpp:228-240).  Copies the meshes, derives the light list and builds a BVH over
 * the world-space light triangles that answers light_pdf (rt.rgen:222-255) in
 * O(log n) instead of looping over every light triangle.  The TLAS must outlive the
 * scene.  Returns when the scene is complete. */
GPURT_API int gpurt_scene_create(gpurt_tlas tlas, const gpurt_object* objects, size_t n_objects,
                                 void* stream, gpurt_scene* out);
/* Waits for the device (every stream), then frees the scene. */
GPURT_API void gpurt_scene_destroy(gpurt_scene scene);

/* Per-frame constants: RTPipe_Constants (src/vk/rt.h:85-102) and the camera's inverse
 * view / projection (CameraConstants, rt.cpp:121-127).  Matrices are row-major (M v). */
typedef struct gpurt_render_params {
    float inv_view[16];
    float inv_proj[16];
    float clear_col[3];   /* camera-ray miss colour */
    float env_light[3];   /* radiance of a miss after the first bounce (env * env_scale) */
    int32_t width;
    int32_t height;
    int32_t frame;        /* 0: overwrite the image; k > 0: image = mix(image, new, 1/(k+1)) */
    int32_t samples;      /* paths per pixel this frame (>= 1) */
    int32_t max_frame;    /* Hammersley sequence length per sample, for qmc */
    int32_t qmc;          /* 1: Hammersley pixel jitter, 0: random (centre on frame 0) */
    int32_t max_depth;    /* bounces per path (1..64) */
    int32_t use_metalness;
    int32_t integrator;   /* 0 direct light sampling, 1 material sampling, 2 MIS */
    int32_t brdf;         /* 0 Blinn-Phong, 1 GGX */
    int32_t use_rr;       /* Russian roulette after every bounce */
    uint32_t seed;        /* per-frame RNG seed: pixel p starts from tea(p, seed) */
    uint64_t* ray_counts; /* host, nullable: [0] += camera / bounce rays traced (closest hit),
                             [1] += shadow (any hit) or light (closest hit) rays traced */
    int32_t concurrency;  /* pixel groups rendered concurrently on their own streams, 1..8
                             (0 = library default: 2 / 4 / 1 for integrators 0 / 1 / 2);
                             results do not depend on it */
    int32_t reserved;
} gpurt_render_params;

/* Render one frame of `samples` paths per pixel into image (height x width x 4 floats,
 * device, row y = launch row y; .w = 1), the wavefront form of rt.rgen main()
 * (rt.rgen:611-677) with integrate_direct / integrate_mats / integrate_mis
 * (rt.rgen:303-414), Blinn-Phong and GGX (rtcommon.glsl:267-369) and the reference's RNG
 * (tea / lcg, rtcommon.glsl:99-124) consumed in the reference's order per pixel.  Each
 * bounce is a closest-hit launch over the live paths, a shading launch that writes the
 * next rays and the shadow (direct: any hit) or light (MIS: closest hit) rays in
 * compacted order, and the shadow launch.  The G-buffer images (nullable, same layout)
 * receive the first sample's first hit position, shading normal and albedo (0 on a
 * miss).  The host waits on `stream` once per bounce to size the next launch.  The
 * reference's clockARB() seeding is replaced by `seed`.  Not supported: textures,
 * normal maps, ReSTIR (integrators 3, 4) and debug views. */
GPURT_API int gpurt_render(gpurt_scene scene, const gpurt_render_params* params, float* image,
                           float* gbuf_pos, float* gbuf_norm, float* gbuf_albedo, void* stream);

/* light_pdf(p, d) of rt.rgen:222-255 for n (point, unit direction) pairs (device n x 3
 * each): the sum over light triangles crossed by the ray at t >= 0 (triangle_hit,
 * rtcommon.glsl:181-202) of triangle_pdf (2 / |cross| * dist^2 / |cos|, lines 204-214),
 * each divided by its light's triangle count, over the light count; found through the
 * scene's light BVH. */
GPURT_API int gpurt_scene_light_pdf(gpurt_scene scene, const float* p, const float* d, size_t n,
                                    float* pdf, void* stream);

/* ---- errors ----------------------------------------------------------------------- */

/* Message of the last failing call on this thread ("" if none). */
GPURT_API const char* gpurt_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* GPURT_H */
