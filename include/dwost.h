/* dwost.h -- C ABI of libdwost, the MI355X Walk-on-Stars engine.
 *
 * Plain C: pointers, sizes and status codes; no torch or C++ types cross the
 * boundary.  Every array argument of a query / solve entry point is a DEVICE
 * pointer (HIP global memory on the current device); `stream` is a hipStream_t
 * (NULL = default stream).  Calls are asynchronous on `stream`; outputs are
 * valid once the stream is synchronised.  Functions return DW_OK (0) or a
 * negative DW_ERR_* code; dw_last_error() gives a thread-local message.
 *
 * The entry points replace the reference's nanobind module and the Dr.Jit
 * traversals that the Python layer runs per query (zihay/dwost):
 *   scene construction  -> src/python/bind_bvh.cpp:101-109, bind_snch.cpp:110-134
 *                          + diff_wost/shapes/polyline.py:623-692, mesh.py:592-635
 *   queries             -> diff_wost/shapes/bvh.py:111-257, snch.py:312-378,
 *                          snch3d.py:334-476,594-663, polyline.py:304-314
 *   solvers             -> diff_wost/solvers/solver.py:36-71, wost.py:21-221,
 *                          wost_3d.py:43-175
 */
#ifndef DWOST_H
#define DWOST_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DW_ABI_VERSION 1

#define DW_OK 0
#define DW_ERR_ARG (-1)
#define DW_ERR_HIP (-2)
#define DW_ERR_INTERNAL (-3)

/* sub-scene selector (Scene2D/Scene3D and their dirichlet_scene / neumann_scene) */
#define DW_SUB_FULL 0
#define DW_SUB_DIRICHLET 1
#define DW_SUB_NEUMANN 2

/* boundary types (diff_wost/shapes/primitive.py:18-19) */
#define DW_NEUMANN 0
#define DW_DIRICHLET 1

/* boundary-value codes: g = u(p) on Dirichlet, h = n . grad u(p) on Neumann */
#define DW_BC_EXPCOS 0     /* u = e^x cos y                                  */
#define DW_BC_QUAD 1       /* u = x^2 - y^2                                  */
#define DW_BC_LINEAR 2     /* u = c0 + c1 x + c2 y + c3 z                    */
#define DW_BC_PER_VERTEX 3 /* g, h interpolated from per-vertex buffers      */
/* source codes (the reference's scene.f, wost.py:128-143) */
#define DW_SRC_ZERO 0
#define DW_SRC_CONST 1     /* f = src_value                                  */
#define DW_SRC_LINEAR 2    /* f = s0 + s1 x + s2 y + s3 z  (src_coeffs)      */

typedef struct dw_scene dw_scene;

typedef struct dw_scene_desc {
    int32_t dim;            /* 2: polyline segments, 3: triangle mesh               */
    int64_t n_vertices;
    int64_t n_prims;
    const float* vertices;  /* HOST, n_vertices x dim, row-major                    */
    const int32_t* indices; /* HOST, n_prims x dim                                  */
    const int32_t* types;   /* HOST, n_prims, DW_NEUMANN / DW_DIRICHLET             */
    int32_t flags;          /* DW_SCENE_HOST_ONLY: build on the host, no upload    */
} dw_scene_desc;

/* Build the trees on the host only (no HIP call; query / solve calls on such
 * a scene fail with DW_ERR_ARG).  Used to inspect and check the builder. */
#define DW_SCENE_HOST_ONLY 1

typedef struct dw_boundary {
    int32_t bc_code;        /* DW_BC_*                                              */
    float coeffs[4];        /* DW_BC_LINEAR coefficients                            */
    const float* g_vertex;  /* DEVICE, n_vertices (DW_BC_PER_VERTEX)                */
    const float* h_vertex;  /* DEVICE, n_vertices (DW_BC_PER_VERTEX)                */
    int32_t src_code;       /* DW_SRC_*                                             */
    float src_value;        /* DW_SRC_CONST                                         */
    float src_coeffs[4];    /* DW_SRC_LINEAR                                        */
} dw_boundary;

typedef struct dw_solver_params { /* Solver / WoSt fields (solver.py:30-34, wost.py:17-19) */
    int32_t nwalks;
    int32_t nsteps;
    int32_t nsubwalks;
    int32_t control_variates;
    int32_t antithetic_variates;
    uint32_t seed;
    float eps;
    float min_R;
    int32_t flags;          /* DW_SOLVE_* bits                                      */
    int64_t point_base;     /* global index of pts[0]: walk streams are keyed by the
                               global (point, walk) index, so a launch over a slice
                               of the points draws what one launch over all does  */
    /* launch shape of the walk kernels; never changes results (0: defaults) */
    int32_t block_size;     /* threads per block: 64, 128 or 256 (0: 256)           */
    int32_t blocks_per_cu;  /* cap on resident blocks per CU (0: occupancy limit)  */
    int32_t spatial_order;  /* DW_ORDER_AUTO / _ON / _OFF: sort points by Morton
                               cell and give each XCD a contiguous range         */
    int32_t schedule;       /* DW_SCHED_*: lane scheduling of the walk kernels     */
    int32_t gate;           /* interleaved schedule: waiting lanes of a wave that
                               trigger a walk-logic pass, 1..64 (0: 24)           */
    int32_t node_width;     /* tree records of the walks' closest-point queries:
                               2 (one binary level per fetch) or 4 (two levels per
                               fetch); 0: the default, 2.  Same visit order, same
                               results                                            */
    int32_t walks_per_lane; /* fused solve (dw_solve_fused, batches): walk pairs of
                               one point a lane runs back to back per job it takes
                               from the queue, 1, 2, 4 or 8 (0: 1).  Scheduling
                               only: each pair keeps its own stream            */
} dw_solver_params;

#define DW_ORDER_AUTO 0 /* on from 8192 points */
#define DW_ORDER_ON 1
#define DW_ORDER_OFF 2

/* Lane scheduling of the persistent walk kernels (results never depend on it).
 * STEP: every lane advances its walk by one whole step per loop iteration, the
 * step's traversals back to back.  INTERLEAVED: lanes run traversal iterations
 * of whatever query each has in flight side by side and run the walk logic
 * between queries in gated passes, so no lane waits for a wave's longest
 * traversal.  AUTO: interleaved in 3D, step-synchronous in 2D. */
#define DW_SCHED_AUTO 0
#define DW_SCHED_INTERLEAVED 1
#define DW_SCHED_STEP 2

/* Evaluate the SNCH cone test literally (acos/asin/atan2 per child, snch.py:46-97)
 * instead of the transcendental-free form |cos| <= sin(half angle + view angle)
 * with a 4e-6 conservative margin; the two differ only where the reference's
 * own rounding would prune a node holding the nearest silhouette. */
#define DW_SOLVE_EXACT_CONES 1
/* Plain Walk on Spheres (WoS.u / WoS.dude, wos.py:16-105; WoS3D, wos_3d.py:15-102):
 * every boundary primitive is treated as Dirichlet, each step jumps to a uniform
 * point on the largest empty ball.  dw_solve_u and dw_solve_dude only. */
#define DW_SOLVE_WOS 2

const char* dw_last_error(void);
int dw_version(void);

/* Scene2D(vertices, indices, types) / Scene3D(...): builds the full,
 * Dirichlet and Neumann sub-scenes with their BVH + SNCH on the host and
 * uploads them once (polyline.py:623-692, mesh.py:592-635; builder
 * include/bvh.h:359-652, include/snch.h:387-489). */
int dw_scene_create(const dw_scene_desc* desc, dw_scene** out);
int dw_scene_destroy(dw_scene* scene);

/* out8 = {present, n_vertices, n_prims, n_nodes, n_silhouette_refs, max_depth,
 *         has_silhouettes, n_topology_silhouettes} of one sub-scene (HOST). */
int dw_scene_info(const dw_scene* scene, int32_t sub, int64_t* out8);

/* Sub-scene primitive id -> scene primitive id (HOST out[n_prims of the
 * sub-scene]): Scene2D._dirichlet_2_global_indices / _neumann_2_global_indices
 * (polyline.py:636-645); the full scene maps to itself. */
int dw_scene_prim_map(const dw_scene* scene, int32_t sub, int32_t* out);

/* HOST copies of one sub-scene's tree, sized from dw_scene_info:
 * box[n_nodes][2*dim] (min, max), ints[n_nodes][4] (offset, n_refs,
 * silhouette offset, n_silhouettes), cone[n_nodes][dim+2] (axis, half angle,
 * radius), prim_index[n_prims], refs[n_silhouette_refs].  Any may be NULL.
 * Mirrors convert_snch_nodes / silhouette_refs_soa (bind_snch.cpp:48-134). */
int dw_scene_export(const dw_scene* scene, int32_t sub, float* box, int32_t* ints, float* cone, int32_t* prim_index,
                    int32_t* refs);

/* BVH.closest_point (bvh.py:111-188) / SNCH3D.closest_point (snch3d.py:334-409).
 * pts[n][dim]; outputs (each may be NULL): prim[n] (sub-scene primitive id),
 * dist[n], point[n][dim], normal[n][dim], tuv[n][2] (2D: t, 0; 3D: u, v). */
int dw_closest_point(const dw_scene* scene, int32_t sub, const float* pts, int64_t n, int32_t* prim, float* dist,
                     float* point, float* normal, float* tuv, void* stream);

/* ray query modes of dw_intersect */
#define DW_RAY_REFERENCE 0 /* the reference's BVH query: 2D first valid hit in traversal order (bvh.py:215-217,
                              what the walks use), 3D nearest hit                                              */
#define DW_RAY_NEAREST 1   /* nearest hit in 2D too: the answer of Polyline.intersect_baseline
                              (polyline.py:468-496), through the BVH; 3D is unchanged                          */

/* BVH.intersect (bvh.py:190-257, first hit in traversal order) /
 * SNCH3D.intersect (snch3d.py:411-476, nearest hit; prim = BVH-sorted index).
 * origin[n][dim], dir[n][dim], rmax[n]; mode: DW_RAY_*; prim[n] = -1 on a miss. */
int dw_intersect(const dw_scene* scene, int32_t sub, const float* origin, const float* dir, const float* rmax,
                 int64_t n, int32_t mode, int32_t* prim, float* dist, float* point, float* normal, float* tuv,
                 void* stream);

/* SNCH.closest_silhouette (snch.py:312-378, snch3d.py:594-663).  sil[n] is the
 * silhouette reference index (-1: none within rmax). */
int dw_closest_silhouette(const dw_scene* scene, int32_t sub, const float* pts, const float* rmax, int64_t n,
                          int32_t* sil, float* dist, float* point, void* stream);

/* Polyline.sdf / Mesh.sdf (polyline.py:188-194, mesh.py:290-293) on the full
 * scene: sdf[n] = closest-point distance signed by the side of the interpolated
 * vertex normal (2D) or face normal (3D), negative inside; Polyline.inside /
 * Mesh.inside (polyline.py:232-234, mesh.py:295-297) is sdf < 0. */
int dw_sdf(const dw_scene* scene, const float* pts, int64_t n, float* sdf, void* stream);

/* Polyline.sample_boundary (polyline.py:498-512) / Mesh.sample_boundary
 * (mesh.py:474-490) on one sub-scene.  The reference draws from the caller's
 * PCG32 sampler; here the caller passes those uniform draws in [0, 1): u[n]
 * (2D: the edge CDF draw, whose reused remainder is t) or u[n][3] (3D: face CDF
 * draw, r1, r2).  Outputs (each may be NULL): point[n][dim], normal[n][dim],
 * tuv[n][2] (2D: t, 0; 3D: u, v), pdf[n], prim[n] (sub-scene primitive id).
 * A draw outside [0, 1) gives prim -1 and NaN values. */
int dw_sample_boundary(const dw_scene* scene, int32_t sub, const float* u, int64_t n, float* point, float* normal,
                       float* tuv, float* pdf, int32_t* prim, void* stream);

/* kinds of dw_query_baseline */
#define DW_BASE_CLOSEST 0    /* closest_point_baseline (polyline.py:246-279, mesh.py:312-335)         */
#define DW_BASE_RAY 1        /* intersect_baseline (polyline.py:468-496, mesh.py:440-472): nearest hit  */
#define DW_BASE_SILHOUETTE 2 /* closest_silhouette_baseline (polyline.py:289-302, mesh.py:558-570)     */

/* The reference's brute-force queries: each query loops over the whole
 * sub-scene in the reference's order (first of equally near candidates wins).
 * pts[n][dim]; dir[n][dim] (DW_BASE_RAY only); rmax[n] (NULL: +inf; closest
 * point ignores it).  prim[n] = sub-scene primitive id (silhouettes: topology
 * id, 2D the local vertex, 3D the edge), -1 if none; dist[n] (may be NULL). */
int dw_query_baseline(const dw_scene* scene, int32_t sub, int32_t kind, const float* pts, const float* dir,
                      const float* rmax, int64_t n, int32_t* prim, float* dist, void* stream);

/* Polyline.star_radius (polyline.py:304-314) / Mesh.star_radius (mesh.py:337-346). */
int dw_star_radius(const dw_scene* scene, int32_t sub, const float* pts, const float* rmax, int64_t n,
                   float* radius, void* stream);

/* Polyline.star_radius_2 (polyline.py:317-330) / Mesh.star_radius_2
 * (mesh.py:349-361): the star radius along direction dirs[n*dim] bounded by
 * the Poisson-kernel ratio rho_max (WoStGrad's ball radius on the Neumann
 * part).  radius[n]. */
int dw_star_radius_rho(const dw_scene* scene, int32_t sub, const float* pts, const float* dirs, float rho_max,
                       int64_t n, float* radius, void* stream);

/* Device math self-check; no counterpart in the reference.  The cone tests
 * decide from fp32 acosf / asinf / atanf / cosf values bracketed by a bound of
 * out[4] float steps to the fp64-evaluated, once-rounded values, running the
 * fp64 form only when a bracket straddles a threshold.  out (DEVICE, 5 int32)
 * receives the largest distance found over every float of each domain:
 * out[0] acos and out[1] asin on [-1, 1], out[2] atan on [0, inf], out[3] cos
 * on [-4, 8]; out[4] = the bound assumed.  A few seconds of GPU time. */
int dw_math_check(int32_t* out, void* stream);

/* Solver.solve with WoSt.u / WoSt3D.u (solver.py:36-71, wost.py:145-221,
 * wost_3d.py:43-106).  u[n] = per-point mean; optional per-walk outputs
 * walk_u / walk_steps / walk_term / walk_flags [n*nwalks]; total_steps (DEVICE uint64) =
 * sum of loop iterations.  Without walk_u the library uses an internal
 * device workspace. */
int dw_solve_u(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
               const dw_boundary* boundary, float* u, float* walk_u, int32_t* walk_steps, int32_t* walk_term,
               uint8_t* walk_flags, uint64_t* total_steps, void* stream);

/* Adjoint of dw_solve_u for per-vertex boundary data (CustomSolver backward,
 * solver.py:247-252): replays each walk from its seed with weight
 * dl_du[i] / (nwalks * nsubwalks) and scatter-adds dL/dg, dL/dh (accumulated).
 * walk_flags[n*nwalks] come from the forward dw_solve_u (bit s: subwalk s not
 * zeroed by the outlier guard); nsubwalks <= 8. */
int dw_adjoint_u(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                 const dw_boundary* boundary, const float* dl_du, const uint8_t* walk_flags, float* dl_dg,
                 float* dl_dh, void* stream);

/* WoSt.solve_dude / WoSt3D.solve_dude (wost.py:21-102, wost_3d.py:108-175):
 * directional derivative du/de at pts, one PCG32 stream per point. */
int dw_solve_dude(const dw_scene* scene, const float* pts, const float* e, int64_t n, const dw_solver_params* params,
                  const dw_boundary* boundary, float* out, uint64_t* total_steps, void* stream);

/* WoStGrad fields beyond dw_solver_params (wost_grad.py:17-26): the bound on
 * the reflected walk's throughput growth used by the star radius, and the
 * switches that drop the boundary, silhouette and source terms. */
typedef struct dw_grad_params {
    float rho_max;             /* 1.2 in the reference                          */
    int32_t ignore_boundary;
    int32_t ignore_silhouette;
    int32_t ignore_source;
} dw_grad_params;

/* WoStGrad.solve_dude (wost_grad.py:91-254) and WoStGrad3D.solve_dude
 * (wost_grad_3d.py:93-206) through Solver.solve_dude / walk_dude
 * (solver.py:113-166): du/de at pts along e[n][dim], nwalks
 * walks per point (one PCG32 stream each, keyed as Solver.solve), nsubwalks
 * dudes per walk.  params: nwalks, nsteps, nsubwalks, eps, min_R (1e-2 in the
 * reference), control_variates (of the nested normal-derivative walk), seed,
 * point_base.  Boundary data must be an analytic field (its tangential
 * derivatives enter the estimator).  3D scenes must be all-Neumann (the
 * reference's 3D walk defines no normal-derivative estimate at Dirichlet
 * points).  out[n]; per-walk walk_out[n*nwalks] and walk_steps (outer steps +
 * nested walk steps) optional. */
int dw_grad_solve_dude(const dw_scene* scene, const float* pts, const float* e, int64_t n,
                       const dw_solver_params* params, const dw_grad_params* grad, const dw_boundary* boundary,
                       float* out, float* walk_out, int32_t* walk_steps, uint64_t* total_steps, void* stream);

/* StreamlineTracer.next_position (examples/example_2d/trace.py:19-68) with a
 * WoStGrad solver, as the reference example drives it (examples/example_2d/
 * run.py:85-127): sample j of point i runs on the stream of lane
 * i * samples_per_point + j seeded with params->seed; walk_dude = the mean of
 * nsubwalks WoStGrad dudes along each axis; out[n][2] = the mean of the samples'
 * new positions (the 2D step rule).  2D scenes, analytic boundary data. */
int dw_grad_trace_step(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                       const dw_grad_params* grad, const dw_boundary* boundary, int32_t samples_per_point,
                       float step_size, float* out, uint64_t* total_steps, void* stream);

/* StreamlineTracer.next_position (examples/example_2d/trace.py:19-68,
 * examples/example_3d/trace.py:21-60) with a WoSt / WoSt3D solver
 * (Solver.walk_dude, solver.py:148-166, along each axis): out[n*dim] = the mean
 * of samples_per_point streamline steps of length step_size from pts; sample j
 * of point i runs on the stream of lane i * samples_per_point + j, seeded with
 * params->seed (the iteration index in the reference example). */
int dw_trace_step(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                  const dw_boundary* boundary, int32_t samples_per_point, float step_size, float* out,
                  uint64_t* total_steps, void* stream);

/* Fused u + grad u (the throughput form; no single reference counterpart):
 * per antithetic pair, the WoSt.dude first-ball step (wost.py:51-99) yields
 * both u(p) (mean value over the ball) and grad u.  u[n], grad[n][dim];
 * optional per-pair outputs pair_vals[n*npairs][dim+1], pair_steps[n*npairs]
 * with npairs = (nwalks+1)/2; pair_flags[n*npairs] (optional) records which
 * walks survived the outlier guards, for dw_adjoint_fused. */
int dw_solve_fused(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                   const dw_boundary* boundary, float* u, float* grad, float* pair_vals, int32_t* pair_steps,
                   uint8_t* pair_flags, uint64_t* total_steps, void* stream);

/* dw_solve_fused for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * No counterpart in the reference.  Same kernels and the same bits as
 * dw_solve_fused without per-pair outputs: u[n], grad[n][dim], total_steps
 * (DEVICE uint64, optional).  Its scratch (queue counters, per-point first-ball
 * radius and inside flag, per-pair partial results, spatial order) is one
 * static device array of DW_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; it calls no allocating routine.  A call whose scratch would
 * exceed that size fails with DW_ERR_ARG before any launch: 3D needs about
 * 20 B and 2D 16 B per (point, walk pair), plus 16 B per point (e.g. 3D
 * n * ceil(nwalks / 2) <= ~3.3 million).  All calls on one GPU share its array
 * and each leaves its contents behind: calls that overlap on one GPU (other
 * streams, other threads) race and must be serialised by the caller. */
#define DW_NOALLOC_SCRATCH_BYTES (64u << 20)
int dw_solve_fused_noalloc(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                           const dw_boundary* boundary, float* u, float* grad, uint64_t* total_steps, void* stream);

/* Adjoint of dw_solve_fused for per-vertex boundary data (the reference's
 * CustomSolver backward pass, solver.py:202-252, which replays the solver with
 * the same seed under Dr.Jit AD).  Replays every walk pair from its seed and
 * scatter-adds dL/dg and dL/dh into dl_dg / dl_dh (n_vertices each, accumulated:
 * zero them first).  dl_du[n], dl_dgrad[n][dim] are the upstream gradients;
 * pair_flags[n*npairs] come from the forward dw_solve_fused with the same
 * params and seed (bit 0/1: walk +/- not zeroed by an outlier guard). */
int dw_adjoint_fused(const dw_scene* scene, const float* pts, int64_t n, const dw_solver_params* params,
                     const dw_boundary* boundary, const float* dl_du, const float* dl_dgrad, const uint8_t* pair_flags,
                     float* dl_dg, float* dl_dh, void* stream);

/* Many scenes and query sets in one persistent launch (fused u + grad u).
 * No reference counterpart (the reference solves one scene per call,
 * solver.py:36-71); each item's results are bit-identical to dw_solve_fused
 * with the same arguments, since walk streams are keyed by (point_base + i,
 * pair) per item.  All items share the dimension and the DW_SOLVE_EXACT_CONES
 * bit; the launch shape and schedule (block_size, blocks_per_cu, schedule,
 * gate) are the first item's; pts / u / grad are device arrays that must outlive the batch (their
 * contents may change between solves).  total_steps (device uint64, optional)
 * receives the walk steps of each solve. */
typedef struct dw_batch_item {
    const dw_scene* scene;
    const float* pts; /* [n][dim] */
    int64_t n;
    dw_solver_params params;
    dw_boundary boundary;
    float* u;    /* [n] */
    float* grad; /* [n][dim] */
} dw_batch_item;
typedef struct dw_batch dw_batch;

/* Capture the launch sequence into a hipGraph on the first solve and replay it
 * afterwards (the sequence holds no memset and no host-memory reads). */
#define DW_BATCH_GRAPH 1

int dw_batch_create(const dw_batch_item* items, int32_t n_items, uint64_t* total_steps, int32_t flags,
                    dw_batch** out);
/* Enqueue one solve of every item on `stream`.  Solves of one batch must not
 * overlap (they share its buffers and queue counters). */
int dw_batch_solve(dw_batch* batch, void* stream);
void dw_batch_destroy(dw_batch* batch);

#ifdef __cplusplus
}
#endif

#endif /* DWOST_H */
