/*
 * wob_c.h -- C ABI of the MI355X walk-on-boundary engine (libwob_hip.so).
 *
 * The reference (rsugimoto/WoBToolbox) is a header-only C++ template library whose device
 * work is a thrust::for_each over evaluation points.  Its user-facing surface on the hot
 * path is:
 *   SceneHost<S,Dim,IsVec>(std::vector<ElementXs>)          wob/wob/scene.cuh:264-299
 *   SceneHost::set_volume_cache / get_device_repr             wob/wob/scene.cuh:301-318
 *   Estimator<...>::compute_sample_path_contribution_domain    wob/wob/estimator.cuh:154-191
 *   Estimator<...>::compute_sample_path_contribution_boundary  wob/wob/estimator.cuh:193-222
 *   utils::run<DeviceFunctor>(args, n, bie, domain, est, prob) wob/wob/runtime_utils.cuh:105-153
 *   utils::random_states_init                                  wob/wob/runtime_utils.cuh:89-103
 *   string_to_*_type / *_type_to_string                        bie_models.cuh:24-51, estimator.cuh:25-54
 * Arbitrary user __device__ functors cannot cross a C ABI; the entry points below run the
 * library's own estimators over caller-supplied evaluation points and accumulate per-point
 * sums the way the reference's app functors do (apps/general_2d.cu:40-71).
 *
 * Conventions
 *   - All pointers named d_* are device pointers (HBM); everything else is host memory.
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream).
 *   - No exceptions cross the ABI; every entry point returns a wob_status and records a
 *     message retrievable with wob_last_error_message() (thread-local).
 *   - A scene owns the scratch memory of its walks (per-item partial sums, a work counter):
 *     estimate calls on one scene run in stream order, so issue them on one stream (or
 *     synchronise between streams).  Handles are not thread-safe.
 *   - Element records follow ElementXs (scene.cuh:378-382, 646-650): per element its Dim
 *     vertices, one normal, Dim boundary values, a boundary type, Dim Robin coefficients.
 *   - Boundary point records follow BoundaryPoint (points.hpp:32-39) as 2*Dim+4 floats:
 *     [p (Dim), n (Dim), interior_angle, boundary_value, boundary_type, robin_alpha].
 */
#ifndef WOB_C_H
#define WOB_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum wob_status {
    WOB_OK = 0,
    WOB_INVALID_ARG = 1,        /* null pointer, zero size, size mismatch */
    WOB_INVALID_ENUM = 2,       /* enum value out of range (the reference silently launches nothing:
                                   runtime_utils.cuh:117-120) */
    WOB_UNSUPPORTED_COMBO = 3,  /* e.g. double layer with non-Dirichlet data (bie_models.cuh:383, 394) */
    WOB_OOM = 4,
    WOB_DEVICE = 5,             /* a HIP runtime error */
    WOB_CANCELLED = 6
} wob_status;

/* Enum values match the reference's enums. */
enum { WOB_DIRICHLET = 0, WOB_NEUMANN = 1, WOB_ROBIN = 2 };                               /* points.hpp:22 */
enum { WOB_POISSON_KERNEL = 0 };                                                           /* equation_kernel.cuh:20 */
enum { WOB_SINGLE_LAYER = 0, WOB_DOUBLE_LAYER = 1, WOB_DIRECT = 2 };                      /* bie_models.cuh:21 */
enum { WOB_INTERIOR = 0, WOB_EXTERIOR = 1 };                                               /* bie_models.cuh:22 */
enum { WOB_BACKWARD = 0, WOB_FORWARD = 1 };                                                /* estimator.cuh:24 */
enum { WOB_DIRICHLET_PROBLEM = 0, WOB_NEUMANN_PROBLEM = 1, WOB_ROBIN_PROBLEM = 2, WOB_MIXED_PROBLEM = 3 }; /* estimator.cuh:38 */

typedef struct wob_scene wob_scene;

/* Replaces SceneHost(const std::vector<ElementXs>&) (scene.cuh:264-299).  Host arrays are
 * copied; nothing refers to caller memory after the call.  Builds the wide BVH, the area and
 * boundary-value CDFs (FP32 Kahan sums, scene.cuh:271-298) and their guide tables.
 *   vertices        n_elements * dim * dim floats
 *   normals         n_elements * dim
 *   boundary_values n_elements * dim
 *   boundary_types  n_elements  (WOB_DIRICHLET / WOB_NEUMANN / WOB_ROBIN)
 *   robin_alphas    n_elements * dim (may be NULL: zeros)                                   */
typedef struct wob_scene_desc {
    uint32_t dim;
    uint32_t n_elements;
    const float *vertices;
    const float *normals;
    const float *boundary_values;
    const int32_t *boundary_types;
    const float *robin_alphas;
} wob_scene_desc;

wob_status wob_scene_create(const wob_scene_desc *desc, wob_scene **out_scene);

/* wob_scene_create with build options.  WOB_SCENE_DEVICE_BVH builds the BVH on the current
 * device, as the reference's SceneHost does with lbvh (scene.cuh:264-299; deps/lbvh): Morton
 * codes of the element boxes, a radix sort, the binary radix tree and its boxes, two passes
 * of SAH treelet restructuring, then a collapse level by level into the same 4-wide node
 * layouts (subtrees of at most two primitives become leaves).  Attributes and CDFs are built
 * on the host as in wob_scene_create.  Walk results, intersections and
 * closest points are the same as with the host-built scene: they do not depend on the tree.
 * A device tree whose traversal stack bound exceeds 48 entries is replaced by the host SAH
 * build (which falls back to median splits).  WOB_DEVICE if the device build fails;
 * WOB_INVALID_ARG for unknown flags or a stack bound over 48 entries after that.            */
enum { WOB_SCENE_DEVICE_BVH = 1 };
/* WOB_SCENE_BVH_WIDTH(w), w = 2, 3 or 4 (default 4): every BVH node takes at most w children
 * (the node layout stays 4-wide; a narrower tree is deeper).  wob_tune_bvh_width picks w.   */
#define WOB_SCENE_BVH_WIDTH_SHIFT 4
#define WOB_SCENE_BVH_WIDTH_MASK (15u << WOB_SCENE_BVH_WIDTH_SHIFT)
#define WOB_SCENE_BVH_WIDTH(w) ((uint32_t)(w) << WOB_SCENE_BVH_WIDTH_SHIFT)
wob_status wob_scene_create_ex(const wob_scene_desc *desc, uint32_t flags, wob_scene **out_scene);
wob_status wob_scene_destroy(wob_scene *scene);

/* Scene facts: dim, element count, BVH node count, BVH depth, totals (area, boundary value),
 * device memory held, the traversal stack entries a line query of this BVH needs and the
 * most children a BVH node takes.                                                           */
typedef struct wob_scene_info {
    uint32_t dim, n_elements, n_nodes, depth;
    float total_area, total_boundary_value;
    uint64_t device_bytes;
    uint32_t stack_need, bvh_width;
} wob_scene_info;
wob_status wob_scene_get_info(const wob_scene *scene, wob_scene_info *info);
/* Replaces SceneHost::set_volume_cache(grid_res, domain_size, cache) (scene.cuh:301-306,
 * 206-240): a grid source over (-domain_size/2, domain_size/2)^dim; estimators with
 * num_volume_samples > 0 sample a cell with probability |cache| / sum |cache| and a uniform
 * point in it (scene.cuh:100-120).  As in the reference there is no inside test: cells
 * outside the domain must hold 0.  cache: grid_res^dim host floats, x fastest
 * (utils.cuh:92-99).  grid_res == 0 or cache == NULL removes the source.  Replaces any
 * earlier source of the scene; waits for the device to be idle.                             */
wob_status wob_scene_set_volume_grid(wob_scene *scene, uint32_t grid_res, float domain_size, const float *cache);

/* Gaussian-mixture source f(x) = sum_j A_j exp(-|x - c_j|^2 / (2 w_j^2)) (no counterpart in
 * the reference, whose only source is the grid cache).  Volume samples come from the mixture
 * with weights |A_j| w_j^dim (inv_pdf = 1 / mixture density) and, unless
 * WOB_VOLUME_NO_MASK is given, a sample on the far side of the boundary (outside an interior
 * domain, inside the boundary for an exterior one; parity of the crossings of a fixed ray)
 * contributes 0, so the source is restricted to the domain.  The masked source builds an
 * occupancy grid over the vertex box (64^3 cells in 3D, 512^2 in 2D, one byte each, on the
 * scene's device) that decides the samples away from the boundary without a ray.  centers:
 * n*dim floats, widths and amplitudes: n floats (host).  n == 0 removes the source.          */
enum { WOB_VOLUME_NO_MASK = 1 };
wob_status wob_scene_set_volume_gaussians(wob_scene *scene, uint32_t n, const float *centers, const float *widths,
                                          const float *amplitudes, uint32_t flags);

/* Estimator configuration: the reference's Estimator template and constructor arguments
 * (estimator.cuh:118-152) plus the JSON keys of the apps (general_2d.cu:174-195).           */
typedef struct wob_estimator_desc {
    int32_t equation_kernel;      /* WOB_POISSON_KERNEL */
    int32_t bie_model;            /* WOB_SINGLE_LAYER / WOB_DOUBLE_LAYER / WOB_DIRECT */
    int32_t domain;               /* WOB_INTERIOR / WOB_EXTERIOR */
    int32_t estimator;            /* WOB_BACKWARD / WOB_FORWARD */
    int32_t problem;              /* WOB_*_PROBLEM */
    uint32_t path_length;
    uint32_t num_resampling_candidates;   /* 1 = plain uniform sampling */
    uint32_t num_volume_samples;          /* > 0: add the scene's volume source term (Poisson) */
    float eps;                            /* kernel clamp, default FLT_EPSILON */
    float first_kind_equation_scaling_constant;  /* default 1 */
    uint32_t flags;                       /* WOB_EST_* options, 0 = the reference's behaviour */
    uint32_t cv_samples;                  /* WOB_EST_CONTROL_VARIATE: samples of the domain-level
                                             control-variate integral (0 = 64) */
} wob_estimator_desc;

/* Estimator options (no counterpart in the reference).
 * WOB_EST_EXCLUDE_ORIGIN_ELEMENT: a line sampled from a point on the boundary ignores the
 *   element that point lies on.  The reference only rejects hits with |t| <= 10 FLT_EPSILON
 *   (scene.cuh:431, 726); an FP32 point sits ~1e-7 off its own element, so a line within a
 *   few degrees of that element re-hits it at a larger |t| and adds a spurious hit (a few
 *   percent of bounces), which inflates the variance and makes paths sensitive to rounding. */
enum { WOB_EST_EXCLUDE_ORIGIN_ELEMENT = 1 };
/* WOB_EST_CONTROL_VARIATE (single-layer model, backward estimator): use the value table bound
 *   with wob_scene_set_control_variate as a control variate.  Every added term has zero mean
 *   for any table, so the estimator's expectation is exactly that of plain WoB; a table close
 *   to the walk's value functions removes most of the path-choice variance.
 * WOB_EST_BOUNDARY_DENSITY (backward estimator, boundary entry points): return the walk's
 *   boundary-density estimate at each record (compute_sample_unknown_boundary,
 *   estimator.cuh:232-357) instead of the solution; this generates control-variate training
 *   data.                                                                                    */
enum { WOB_EST_CONTROL_VARIATE = 2, WOB_EST_BOUNDARY_DENSITY = 4 };
/* WOB_EST_INDEPENDENT_VOLUME_SAMPLES: each of the num_volume_samples addends of a volume term
 *   uses its own sample.  The reference draws one sample per term and adds it
 *   num_volume_samples times (bie_models.cuh:85-101), which leaves the variance unchanged.   */
enum { WOB_EST_INDEPENDENT_VOLUME_SAMPLES = 8 };
/* WOB_EST_HARMONIC_CV (single-layer or double-layer model, or the direct model in an
 *   interior domain; backward or forward estimator; domain estimates or boundary densities;
 *   the fit must serve the domain's side, see wob_scene_fit_harmonic_cv): use the scene's
 *   harmonic function u_h.  Every boundary datum
 *   the walk reads is replaced by the residual's: g - u_h at Dirichlet points, g - du_h/dn at
 *   Neumann points, g - (du_h/dn + alpha u_h) at Robin points; each domain estimate adds
 *   u_h(x) and grad u_h(x).  u - u_h solves the same problem with the residual data, so the
 *   expectation is that of plain WoB for any u_h harmonic in the domain (of the layer
 *   potential's class at infinity in an exterior domain; the path-length truncation then
 *   applies to the residual's series).  The forward estimator then draws its start point,
 *   and the direct model its known terms' sample, uniformly rather than by the data's density
 *   (the residual data need not vanish where the data do).  Combines with
 *   WOB_EST_CONTROL_VARIATE (a value table of the residual walk).                          */
enum { WOB_EST_HARMONIC_CV = 16 };

/* Launch tuning and timing (in/out; may be NULL).  paths_per_lane: paths one lane
 * accumulates before it writes its partial sums; block_size: threads per workgroup
 * (64..256); 0 = automatic.  With WOB_LAUNCH_TIME_WALK in flags the call brackets the walk
 * kernel with HIP events on the launch stream, waits for the end event and writes the
 * kernel's duration to walk_ms (and the work-item geometry it chose to the other outputs).
 * schedule: WOB_SCHEDULE_AUTO picks the per-lane walk (the faster one on every line walk
 * measured); WOB_SCHEDULE_LANE runs the wave's queries to their end, lanes whose query ended
 * taking traversal work from the others; WOB_SCHEDULE_WAVEFRONT lets a lane whose query ended run
 * its walk to the next query while the others traverse (an error where unsupported).  The
 * results are bit-identical either way.  regen_threshold: lanes that must have finished their
 * query before a wave advances them (wavefront walk; 0 = default 32).                       */
enum { WOB_LAUNCH_TIME_WALK = 1 };
/* WOB_LAUNCH_ONE_LANE_PER_ITEM: launch one lane per work item instead of the default resident
 * grid whose waves take work items from a counter (same results).                          */
enum { WOB_LAUNCH_ONE_LANE_PER_ITEM = 2 };
enum { WOB_SCHEDULE_AUTO = 0, WOB_SCHEDULE_LANE = 1, WOB_SCHEDULE_WAVEFRONT = 2 };
typedef struct wob_launch_config {
    uint32_t paths_per_lane;
    uint32_t block_size;
    uint32_t flags;
    float walk_ms;              /* out */
    uint32_t used_paths_per_lane; /* out */
    uint32_t used_work_items;     /* out */
    uint32_t schedule;
    uint32_t regen_threshold;
    uint32_t used_schedule;       /* out */
} wob_launch_config;

/* Domain estimates: replaces utils::run<DeviceFunctor> with an app functor that calls
 * compute_sample_path_contribution_domain num_sample_paths times per point and adds the
 * Kahan sum into per-point buffers (apps/general_2d.cu:44-70).
 *   d_points   n_points * dim floats (device)
 *   point_base global index of d_points[0] (RNG counter word; lets callers split point sets)
 *   path_base  global index of the first path of every point in this call (progressive passes
 *              use path_base = pass * paths_per_point so every path has its own stream).
 *              The stream's path word is 32 bits: path_base + paths_per_point must not exceed
 *              2^32 (WOB_INVALID_ARG otherwise, before any work); later paths of a point take
 *              another seed or stream_id, whose streams are independent of these.  Batches
 *              check path_base + (passes so far + n_passes) * paths_per_point the same way.
 *   stream_id  RNG stream word (0 for a single scene; batch entries use their scene index)
 *   d_sum, d_sumsq  n_points * (dim + 1) doubles [u, grad u] each, ACCUMULATED into (+=)
 * Per path the result is [u, grad u]; per lane paths are Kahan-summed in FP32, chunks are
 * reduced per point in FP64 in a fixed order (deterministic for fixed inputs).             */
wob_status wob_estimate_domain(const wob_scene *scene, const wob_estimator_desc *est, const float *d_points,
                               uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                               uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                               wob_launch_config *launch, void *stream);

/* No counterpart in the reference.  wob_estimate_domain for callers that cannot allocate
 * device memory: code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc.  Same arguments, kernels and results (bit for bit) as wob_estimate_domain; the
 * caller still passes the points and the output sums.  The launch's work counter and
 * per-work-item partial sums live in a static device array of WOB_NOALLOC_SCRATCH_BYTES
 * per GPU, taken when the library first runs this entry point on that GPU and held until the
 * process ends; the call itself allocates nothing.  A launch needing more than the array
 * holds (256 bytes + 8 * (dim + 1) bytes per work item; about 1M items at the automatic
 * paths_per_lane) returns WOB_INVALID_ARG.  launch->flags may not ask for
 * WOB_LAUNCH_TIME_WALK (timing creates events).  Every call on one GPU shares that GPU's
 * array: calls that overlap on one GPU (several streams or threads) race and must be
 * serialised by the caller.                                                                 */
#define WOB_NOALLOC_SCRATCH_BYTES (64u << 20)
wob_status wob_estimate_domain_noalloc(const wob_scene *scene, const wob_estimator_desc *est, const float *d_points,
                                       uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                                       uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                                       wob_launch_config *launch, void *stream);

/* On-boundary estimates: compute_sample_path_contribution_boundary (estimator.cuh:193-222,
 * forward :456-496), as used by apps/boundary_estimate.cu:65-98.
 *   d_records  n_points * (2*dim + 4) floats (BoundaryPoint records, device)
 *   d_sum, d_sumsq  n_points doubles each, accumulated into.                                 */
wob_status wob_estimate_boundary(const wob_scene *scene, const wob_estimator_desc *est, const float *d_records,
                                 uint32_t n_points, uint32_t point_base, uint32_t paths_per_point, uint64_t seed,
                                 uint64_t path_base, uint32_t stream_id, double *d_sum, double *d_sumsq,
                                 wob_launch_config *launch, void *stream);

/* Virtual point lights (apps/potential_flow_with_vpl.cu): the path phase and the dense
 * aggregation of one iteration.
 *
 * wob_vpl_paths replaces PathComputeFunc (potential_flow_with_vpl.cu:73-88): for i < n_paths,
 * y_i = sample_boundary_uniform (scene.cuh:441-462 / 734-758) and c_i =
 * compute_sample_unknown_boundary(y_i) (estimator.cuh:232-357), one backward walk per start
 * point.  y_i uses the Philox stream (seed, path_base, i, stream_id | 2^30) and the walk from
 * it (seed, path_base, i, stream_id); an iteration loop advances path_base by 1.
 *   d_records  n_paths * (2*dim + 4) floats (BoundaryPoint records, written)
 *   d_contrib  n_paths floats (c_i, written)
 *   stream_id  below 2^30; path_base below 2^32
 * The estimator must be backward without the control variate; launch (may be NULL) tunes and
 * times the walk as for wob_estimate_boundary.                                               */
wob_status wob_vpl_paths(const wob_scene *scene, const wob_estimator_desc *est, uint32_t n_paths, uint64_t seed,
                         uint64_t path_base, uint32_t stream_id, float *d_records, float *d_contrib,
                         wob_launch_config *launch, void *stream);
/* wob_vpl_aggregate replaces AggregateFunc (potential_flow_with_vpl.cu:44-71) over all points:
 *   d_sum[i*(dim+1) + 0]     += sum_j c_j * solution_domain_unknown(x_i, y_j, |boundary|)
 *   d_sum[i*(dim+1) + 1..dim] += sum_j c_j * gradient_domain_unknown(x_i, y_j, |boundary|)
 * for the estimator's bie_model, domain and eps (bie_models.cuh:228-243, 341-356, 428-462;
 * the other fields are ignored).  d_points: n_points * dim floats; d_records / d_contrib as
 * written by wob_vpl_paths (or any BoundaryPoint records and weights).  FP32 terms,
 * compensated FP32 sums over VPL ranges, FP64 accumulation into d_sum (deterministic).  The
 * caller divides by the number of paths, as the reference app does (:250-257), and skips
 * masked points by leaving them out of d_points.
 * Concurrency: an aggregation and a path phase of the same scene work in disjoint scene
 * buffers, so iteration k's aggregation may run on one stream while iteration k + 1's path
 * phase runs on another (into other record / contribution buffers); two path phases, or two
 * aggregations, of one scene must not overlap.  The first call of each may allocate.        */
wob_status wob_vpl_aggregate(const wob_scene *scene, const wob_estimator_desc *est, const float *d_records,
                             const float *d_contrib, uint32_t n_vpl, const float *d_points, uint32_t n_points,
                             double *d_sum, void *stream);

/* Closest points (the lbvh nearest query with the point-segment distance of
 * apps/simple_wob_and_wos.cu:48-79; triangles use the closest point on the triangle).  For
 * each point i: d_elem[i] = the nearest element (caller order; ties go to the smaller id),
 * d_dist[i] = its distance, d_value[i] = the boundary value interpolated at the closest point
 * (simple_wob_and_wos.cu:70-74; barycentric on triangles).  d_points: n * dim floats.      */
wob_status wob_closest_points(const wob_scene *scene, const float *d_points, uint32_t n, uint32_t *d_elem, float *d_dist,
                              float *d_value, void *stream);
/* Walk on spheres for interior Dirichlet problems (WoSFunc, simple_wob_and_wos.cu:116-141):
 * from x, R = distance to the boundary; while R >= eps, x += R * (uniform direction); then
 * the path scores the boundary value at the closest point.  Path j of point i uses the Philox
 * stream (seed, path_base + j, point_base + i, stream_id), one uniform per 2D step and two
 * per 3D step.  A path still walking after max_steps closest-point queries (0 = 65536)
 * scores its current closest value (the reference loops without a bound).  d_sum / d_sumsq:
 * n_points doubles each, accumulated into, as wob_estimate_domain's u column.            */
wob_status wob_wos_estimate(const wob_scene *scene, const float *d_points, uint32_t n_points, uint32_t point_base,
                            uint32_t paths_per_point, uint64_t seed, uint64_t path_base, uint32_t stream_id, float eps,
                            uint32_t max_steps, double *d_sum, double *d_sumsq, void *stream);

/* Batched launches (no counterpart in the reference, whose apps run one scene per
 * thrust::for_each): many (scene, estimator, point set) entries of uneven sizes run in ONE
 * walk launch and ONE reduction launch per pass.  All entries must share the dimension and
 * the kernel flavour (model, estimator and sampling strategy; problem types may differ when
 * they select the same strategy, e.g. single-layer Neumann / Robin / mixed scenes with the
 * backward estimator).  Pass p of entry e walks paths [path_base + p * paths_per_point, +
 * paths_per_point) of every point and accumulates into its d_sum / d_sumsq.  With use_graph
 * the pass (walk, reduction, counter update) is captured once in a hipGraph and replayed.  */
typedef struct wob_batch wob_batch;
typedef struct wob_batch_entry {
    const wob_scene *scene;
    wob_estimator_desc est;
    const float *d_points;      /* domain points, or BoundaryPoint records when boundary != 0 */
    uint32_t n_points;
    uint32_t paths_per_point;
    uint32_t point_base;
    uint32_t stream_id;
    uint64_t path_base;
    double *d_sum, *d_sumsq;    /* n_points * (dim + 1) doubles each (1 for boundary batches) */
} wob_batch_entry;

typedef struct wob_batch_config {
    uint32_t boundary;          /* 0: domain estimates, 1: on-boundary estimates */
    uint32_t paths_per_lane;    /* 0 = automatic */
    uint32_t block_size;        /* 0 = automatic (128) */
    uint32_t persistent;        /* 1: resident grid, waves pull 64-item blocks from a counter */
    uint32_t use_graph;         /* 1: capture a pass in a hipGraph */
    uint32_t schedule;          /* WOB_SCHEDULE_* (the wavefront walk is always resident); the
                                   autotuner tries both where the wavefront walk exists */
    uint32_t regen_threshold;   /* wavefront walk; 0 = default */
} wob_batch_config;

wob_status wob_batch_create(const wob_batch_entry *entries, uint32_t n_entries, const wob_batch_config *cfg,
                            wob_batch **out_batch);
wob_status wob_batch_run(wob_batch *batch, uint64_t seed, uint32_t n_passes, void *stream);
/* BVH width autotuner (no counterpart in the reference; the reference's lbvh is binary):
 * builds the scene (desc, flags as in wob_scene_create_ex without a width) with nodes of at
 * most 2, 3 and 4 children, runs one untimed and one timed estimate of paths_per_point paths
 * at the device points with each (the walk kernel's time from HIP events on its stream),
 * and returns the fastest width in *best_width.  ms_by_width (host, 3 floats, may be NULL):
 * the times for widths 2, 3, 4, NaN where the scene cannot be built (a stack bound over 48).
 * Walk results do not depend on the width.  Plain walks without a volume term only.        */
wob_status wob_tune_bvh_width(const wob_scene_desc *desc, uint32_t flags, const wob_estimator_desc *est,
                              const float *d_points, uint32_t n_points, uint32_t paths_per_point, uint64_t seed,
                              uint32_t *best_width, float *ms_by_width, void *stream);

/* Autotuner: times one pass for each candidate (paths per lane x block size x persistent)
 * into scratch outputs (the entries' outputs are not touched), keeps the fastest and
 * reports it in *best (may be NULL) with its pass time in *best_ms.                          */
wob_status wob_batch_tune(wob_batch *batch, uint64_t seed, wob_batch_config *best, float *best_ms, void *stream);
wob_status wob_batch_destroy(wob_batch *batch);

/* Learned control variate (no counterpart in the reference): a Fourier-feature MLP
 * (64 Gaussian frequencies of scale `sigma`, 4 hidden layers of width 128, ReLU,
 * n_outputs <= 16 outputs) trained and evaluated by BF16 MFMA kernels with FP32
 * accumulation and FP32 Adam master weights.  Output r - 1 models V_r, the expected value of
 * a single-layer walk of r bounces from a boundary point.                                   */
typedef struct wob_cv wob_cv;
wob_status wob_cv_create(uint32_t dim, uint32_t n_outputs, float sigma, uint64_t seed, wob_cv **out_cv);
wob_status wob_cv_destroy(wob_cv *cv);
/* n_params: FP32 parameter count (layer l: W_l (out x in) then b_l).                        */
wob_status wob_cv_get_info(const wob_cv *cv, uint32_t *n_params, uint32_t *n_outputs);
/* host copies of the FP32 master parameters and the 64 x 3 Fourier frequencies             */
wob_status wob_cv_get_params(const wob_cv *cv, float *params, float *freq);
wob_status wob_cv_set_params(wob_cv *cv, const float *params, const float *freq);
/* Training: `steps` Adam steps over consecutive mini-batches of `batch` samples (wrapping
 * around n); sample i has position d_pos[i] (dim floats), output index d_out_index[i] and
 * target d_target[i] (masked mean squared error).  losses (host, may be NULL) receives the
 * mean loss of every step.  lr decays by a cosine schedule to lr / 20.                      */
wob_status wob_cv_train(wob_cv *cv, const float *d_pos, const int32_t *d_out_index, const float *d_target, uint32_t n,
                        uint32_t batch, uint32_t steps, float lr, float *losses, void *stream);
/* Gradient of the batch loss (no update): grad (host, n_params floats), loss (host).        */
wob_status wob_cv_gradient(wob_cv *cv, const float *d_pos, const int32_t *d_out_index, const float *d_target,
                           uint32_t n, float *grad, float *loss, void *stream);
wob_status wob_cv_infer(wob_cv *cv, const float *d_pos, uint32_t n, float *d_out, void *stream);
/* Evaluate the network at every element vertex and bind the table (values r = 1..n_values)
 * to the scene; cv == NULL with d_table != NULL binds a caller table [n_values][n][dim]
 * (device) instead; both NULL unbinds.                                                      */
wob_status wob_scene_set_control_variate(wob_scene *scene, wob_cv *cv, const float *d_table, uint32_t n_values,
                                         void *stream);

/* Learned harmonic control variate (no counterpart in the reference) for single-layer and
 * double-layer walks of interior and exterior domains (WOB_EST_HARMONIC_CV), e.g. the Poisson
 * problem's double-layer estimator (poisson_equation.cu:40-61), the mixed single-layer
 * workload and the batched exterior scenes:
 *   u_h(x) = c0 + sum_j w_j phi(|x - s_j|^2),  phi(z) = z^(-1/2) (3D) or log2 z (2D),
 * a combination of fundamental solutions with sources s_j on a sphere (circle) outside the
 * scene (interior fits), so u_h is harmonic in the interior domain, or inside the body
 * (WOB_HCV_EXTERIOR), so u_h is harmonic in the exterior domain; an exterior fit has c0 = 0,
 * since a layer potential has no constant term at infinity.  The variance of the walk
 * falls with the size of the residual data.
 * Fit: the targets are the walk's own known-term data at n_train uniform boundary samples
 * (sample i on the Philox stream (seed, path i, point 0, stream_id)): the boundary datum plus
 * the volume term as the estimator samples it (the mean of volume_samples independent
 * volume samples; 0 without a volume source); each sample contributes the row of its type
 * (u_h, du_h/dn or du_h/dn + alpha u_h at the sample), (c0, w) minimise the ridge-regularised
 * squared error in FP64 (columns scaled to unit RMS, ridge relative to the mean diagonal),
 * and the result is bound to the scene.  An interior fit serves interior walks only, an
 * exterior fit exterior walks only (double-layer ones with WOB_HCV_ZERO_CHARGE).
 * Exterior source placement: a sphere (circle) of radius radius_scale * R_in about the centre
 * c of the scene's vertex box, R_in the smallest distance from c to an element; the sources
 * lie inside the body when c does and the body holds the ball of radius R_in about c (a body
 * star-shaped about c does).                                                                */
#define WOB_HCV_EXTERIOR 1u
/* With WOB_HCV_EXTERIOR: the fit keeps sum_j w_j = 0, so u_h decays at infinity like a
 * double-layer potential (no log term in 2D, no 1/r term in 3D); exterior double-layer walks
 * need such a fit, exterior single-layer walks take either.                                 */
#define WOB_HCV_ZERO_CHARGE 2u
typedef struct wob_hcv_desc {
    uint32_t n_sources;       /* 1 .. 1024 */
    float radius_scale;       /* interior: source radius / the scene's bounding radius about its box
                                 centre, > 1; exterior: source radius / R_in, in (0, 1) */
    uint32_t n_train;         /* boundary samples, >= 2 (n_sources + 1) */
    uint32_t volume_samples;  /* volume-term samples per target (the scene's source; 0: data only) */
    float eps;                /* kernel clamp of the volume term (the walk's eps) */
    float ridge;              /* relative Tikhonov weight, >= 0 */
    uint64_t seed;
    uint32_t stream_id;
    uint32_t flags;           /* 0, WOB_HCV_EXTERIOR, or WOB_HCV_EXTERIOR | WOB_HCV_ZERO_CHARGE */
} wob_hcv_desc;
/* rms_residual (host, may be NULL): RMS of target - (the sample's row) over the samples.  */
wob_status wob_scene_fit_harmonic_cv(wob_scene *scene, const wob_hcv_desc *desc, float *rms_residual, void *stream);
/* Training data of the fit as BoundaryPoint records (n_train * (2 dim + 4) floats, device):
 * [p, n, 0.5, target, type, robin_alpha].                                                   */
wob_status wob_harmonic_cv_targets(const wob_scene *scene, const wob_hcv_desc *desc, float *d_records, void *stream);
/* Bind explicit sources (host, n_sources * (dim + 1) floats: s_j then w_j) and c0;
 * n_sources == 0 unbinds.                                                                   */
wob_status wob_scene_set_harmonic_cv(wob_scene *scene, uint32_t n_sources, const float *sources, float c0);
/* The bound function: *n_sources, and (if not NULL) the sources in the layout above and c0. */
wob_status wob_scene_get_harmonic_cv(const wob_scene *scene, uint32_t *n_sources, float *sources, float *c0);
/* u_h and grad u_h at n device points (d_points: n * dim, d_out: n * (dim + 1) floats), with
 * the walk's own arithmetic.                                                                */
wob_status wob_eval_harmonic_cv(const wob_scene *scene, const float *d_points, uint32_t n, float *d_out, void *stream);

/* Uniform boundary samples as BoundaryPoint records (Scene::sample_boundary_uniform,
 * scene.cuh:441-462 / 734-758): record i uses the Philox stream (seed, path i, point 0,
 * stream_id).  d_records: n * (2*dim + 4) floats (device).                                  */
wob_status wob_sample_boundary(const wob_scene *scene, uint32_t n, uint64_t seed, uint32_t stream_id, float *d_records,
                               void *stream);

/* One line-intersection sample per domain point with the walk's sampler
 * (Scene::sample_boundary_line_intersection, scene.cuh:464-535 / 760-840): a uniformly
 * oriented line through x_i, all its hits, one hit taken uniformly (the order-independent
 * pick of the walk) and the BoundaryPoint there with its inverse pdf.  Point i uses the
 * Philox stream (seed, path, point i, stream_id).  d_points: n*dim floats; d_records:
 * n * (2*dim + 4) floats; d_inv_pdf: n floats (device).  No hit: a zero record, inv_pdf 0. */
wob_status wob_sample_lines(const wob_scene *scene, const float *d_points, uint32_t n, uint64_t seed, uint32_t path,
                            uint32_t stream_id, float *d_records, float *d_inv_pdf, void *stream);

/* All-hit line query (the lbvh query_line_intersect call of scene.cuh:481-486 / 787-790,
 * without its 64-hit cap).  For each line i (origin, direction; device arrays of n*dim
 * floats): d_counts[i] = number of hits; the first max_hits hits are written as element ids
 * (caller's element order) and line parameters t, in traversal order.                      */
wob_status wob_intersect_lines(const wob_scene *scene, const float *d_origins, const float *d_dirs, uint32_t n,
                               uint32_t max_hits, uint32_t *d_counts, uint32_t *d_ids, float *d_t, void *stream);

/* Deterministic Laplace kernels (equation_kernel.cuh:67-145).  Per sample, out holds
 * [G, dG_dx (dim), dG_dny, d2G_dxdny (dim), d2G_dnxdny] = 3 + 2*dim floats.                 */
wob_status wob_eval_kernels(uint32_t dim, uint32_t n, const float *d_y, const float *d_x, const float *d_ny,
                            const float *d_nx, float eps, float *d_out, void *stream);

/* Deterministic BIE weights (bie_models.cuh) for given pairs of boundary records x, y, the
 * inverse pdf, a domain point and the uniform that drives the first-kind splitting.
 * Per pair, out holds [fredholm_known, fredholm_unknown, split_taken,
 * solution_domain_unknown, gradient_domain_unknown (dim), solution_boundary_unknown].       */
wob_status wob_eval_weights(uint32_t dim, int32_t bie_model, int32_t domain, float eps, float fkesc, uint32_t n,
                            const float *d_x, const float *d_y, const float *d_inv_pdf, const float *d_xdomain,
                            const float *d_u_split, float *d_out, void *stream);

/* Counter-based RNG used by every estimator: Philox4x32-10 (Salmon et al., SC'11).
 * out[4*i..] = philox(ctr[4*i..], key[2*i..]) for i < n (device arrays).                    */
wob_status wob_philox4x32_10(uint32_t n, const uint32_t *d_ctr, const uint32_t *d_key, uint32_t *d_out, void *stream);

/* String <-> enum maps of the reference (bie_models.cuh:24-51, estimator.cuh:25-54,
 * equation_kernel.cuh:25-34).  Unknown strings return -1.                                   */
int32_t wob_string_to_enum(const char *s);
const char *wob_enum_to_string(const char *kind, int32_t value);

const char *wob_last_error_message(void);
const char *wob_version(void);

#ifdef __cplusplus
}
#endif

#endif /* WOB_C_H */
