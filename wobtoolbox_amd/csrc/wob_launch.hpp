// Host-visible launch interface between the C ABI (wob_capi.cpp) and the kernels
// (wob_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "wob_device.hpp"

namespace wobl {

// Per-lane LDS traversal stack (uint32 entries).  The scene builder rejects BVHs whose
// depth-first stack bound exceeds this.
constexpr int kStackEntries = 48;

// One (scene, estimator, point set) of a launch.  A single estimate call is a one-entry
// launch; a batch packs many entries (scenes of different sizes, uneven point sets) into one
// launch whose work items are numbered entry by entry.
struct WalkEntry {
    wobdev::DevScene scene;
    const float *points;      // domain points (n * D) or boundary records (n * (2D + 4))
    double *sum, *sumsq;      // n_points * V each, accumulated into
    uint32_t n_points;
    uint32_t chunks;          // work items per point
    uint32_t paths_per_point;
    uint32_t paths_per_chunk;
    uint32_t point_base;      // RNG counter word of points[0]
    uint32_t stream;          // RNG stream word
    uint32_t path_length, nrc;
    float dsign, eps, fkesc;
    uint32_t item_begin;      // first global work item of this entry
    uint32_t point_begin;     // first global point of this entry (reduction)
    uint32_t partial_base;    // offset of this entry's partials (floats)
    uint32_t flags;           // WOB_EST_* options
    uint32_t cv_samples;      // domain-level control-variate samples
    uint32_t nvs;             // num_volume_samples (volume kernels only)
    uint32_t pad_;
    uint64_t path_base;       // first path index of pass 0
};

struct WalkLaunch {
    WalkEntry e0;                  // used when table == nullptr
    const WalkEntry *table;        // device table of n_entries (batches)
    uint32_t n_entries;
    uint32_t total_items;
    uint32_t total_points;
    uint32_t persistent;           // 1: waves pull 64-item blocks from *work_counter
    uint32_t *work_counter;        // device counter (persistent mode)
    const uint32_t *pass_counter;  // device pass index (graph replays), may be null
    uint64_t seed;
    float *partials;
    uint32_t wavefront;            // 1: wavefront walk (walk_kernel_wf); needs work_counter at 0
    uint32_t regen_threshold;      // wavefront walk: lanes that must wait before a wave advances them
};

hipError_t launch_walk(int dim, bool boundary, bool volume, const WalkLaunch &a, int bie, int est, int flavour,
                       uint32_t block, uint32_t grid, hipStream_t st);
hipError_t launch_reduce(const WalkLaunch &a, uint32_t V, hipStream_t st);
hipError_t launch_finalize(uint32_t *work_counter, uint32_t *pass_counter, hipStream_t st);
hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st);
hipError_t launch_eval_kernels(int dim, uint32_t n, const float *y, const float *x, const float *ny, const float *nx,
                               float eps, float *out, hipStream_t st);
hipError_t launch_eval_weights(int dim, int bie, float dsign, float eps, float fkesc, uint32_t n, const float *xr,
                               const float *yr, const float *inv, const float *xd, const float *us, float *out,
                               hipStream_t st);
hipError_t launch_intersect(int dim, const wobdev::DevScene &S, uint32_t n, const float *o, const float *d,
                            uint32_t max_hits, uint32_t *counts, uint32_t *ids, float *ts, hipStream_t st);
hipError_t launch_sample_lines(int dim, const wobdev::DevScene &S, uint32_t n, const float *points, uint64_t seed,
                               uint32_t path, uint32_t stream, float *rec, float *inv_pdf, hipStream_t st);
// path < 0: record i uses the stream (seed, path i, point 0, stream); path >= 0: (seed, path,
// point i, stream)
hipError_t launch_sample_boundary(int dim, const wobdev::DevScene &S, uint32_t n, uint64_t seed, uint32_t stream,
                                  float *out, hipStream_t st, int64_t path = -1);
hipError_t launch_philox(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t st);
// the m fused area-guide records of a 3D scene (4 float4 rows each): row 0 = (i = guide[k],
// cdf[i - 1], cdf[i], cdf[i + 1]) with 0 outside the CDF, rows 1-3 = geo rows 4 i .. 4 i + 2
hipError_t launch_fused_records(const uint32_t *guide, uint32_t m, const float *cdf, uint32_t n, const float4 *geo,
                                float4 *out, hipStream_t st);

// Virtual-point-light aggregation (wob_vpl.hip).  vpl_plan picks the VPL split count and
// chunk; the stage kernel writes n * vpl_record_floats records; the aggregate kernel writes
// splits * n_pts * (dim + 1) doubles of partials that the reduce kernel adds into out.
uint32_t vpl_record_floats(int dim, int bie);
void vpl_plan(uint32_t n_pts, uint32_t n_vpl, uint32_t &splits, uint32_t &chunk);
hipError_t launch_vpl_stage(int dim, int bie, const float *rec_in, const float *contrib, uint32_t n, float inv_pdf,
                            float dsign, float *out, hipStream_t st);
hipError_t launch_vpl_aggregate(int dim, int bie, const float *pts, uint32_t n_pts, const float *rec, uint32_t n_vpl,
                                uint32_t splits, uint32_t chunk, float eps, double *partial, hipStream_t st);
hipError_t launch_vpl_reduce(const double *partial, uint32_t n_pts, uint32_t V, uint32_t splits, double *out,
                             hipStream_t st);
hipError_t launch_f64_to_f32(const double *in, uint32_t n, float *out, hipStream_t st);

// Closest-point queries and walk on spheres (wob_wos.hip).  The WoS launch uses e0 of the
// launch (points, n_points, chunks, paths_per_chunk, paths_per_point, point_base, stream,
// path_base) and writes 2 * total_items partial floats for reduce_kernel with V = 1.
size_t closest_lds_bytes(uint32_t block);
hipError_t launch_closest(int dim, const wobdev::DevScene &S, uint32_t n, const float *pts, uint32_t *elem, float *dist,
                          float *value, hipStream_t st);
hipError_t launch_wos(int dim, const WalkLaunch &L, float eps, uint32_t max_steps, hipStream_t st);

// Per (dimension, volume term) entry points of the walk translation units (wob_walk.hip).
#define WOB_DECLARE_WALK(d, v)                                                                                  \
    hipError_t launch_walk_##d##v(bool boundary, const WalkLaunch &a, int bie, int est, int flavour, uint32_t block, \
                                  uint32_t grid, hipStream_t st);                                               \
    uint32_t persistent_grid_##d##v(bool boundary, int bie, int est, int flavour, uint32_t block, bool wavefront,  \
                                    bool multi);                                                               \
    bool supports_wavefront_##d##v(bool boundary, int bie, int est, int flavour);
WOB_DECLARE_WALK(2, 0)
WOB_DECLARE_WALK(2, 1)
WOB_DECLARE_WALK(3, 0)
WOB_DECLARE_WALK(3, 1)
#undef WOB_DECLARE_WALK

// Training targets of the harmonic control variate (wob_walk.hip, volume translation units):
// n uniform boundary samples as BoundaryPoint records (n * (2 dim + 4) floats) whose boundary
// value is the target; lds_bytes = walk_lds_bytes(FL_UNIFORM, true, 256).
hipError_t launch_hcv_targets_2(const wobdev::DevScene &S, uint32_t n, uint64_t seed, uint32_t stream, uint32_t nvs,
                                float eps, bool interior, float *rec, size_t lds_bytes, hipStream_t st);
hipError_t launch_hcv_targets_3(const wobdev::DevScene &S, uint32_t n, uint64_t seed, uint32_t stream, uint32_t nvs,
                                float eps, bool interior, float *rec, size_t lds_bytes, hipStream_t st);
// The inside test's occupancy grid (wob_walk.hip, volume translation units): classifies every
// cell whose flag is 255 by ray parity; S.vol carries the grid layout; lds_bytes =
// walk_lds_bytes(FL_UNIFORM, true, 256).
hipError_t launch_inside_grid_2(const wobdev::DevScene &S, uint8_t *flag, uint32_t n_cells, size_t lds_bytes,
                                hipStream_t st);
hipError_t launch_inside_grid_3(const wobdev::DevScene &S, uint8_t *flag, uint32_t n_cells, size_t lds_bytes,
                                hipStream_t st);
// The fit's augmented FP64 Gram matrix (G: (m + 2)^2 doubles) of n target records and m
// source positions (src: m * dim doubles); the walk's residual datum at each record (out: n).
size_t hcv_gram_scratch(uint32_t n, uint32_t m);   // bytes of `partial`
hipError_t launch_hcv_gram(int dim, const float *rec, uint32_t n, const double *src, uint32_t m, double *partial,
                           double *G, hipStream_t st);
hipError_t launch_hcv_residual(int dim, const wobdev::DevScene &S, const float *rec, uint32_t n, float *out,
                               hipStream_t st);
// u_h and grad u_h of the scene's harmonic control variate at n points (out: n * (dim + 1))
hipError_t launch_hcv_eval(int dim, const wobdev::DevScene &S, uint32_t n, const float *pts, float *out, hipStream_t st);

// Occupancy-derived grid for persistent launches (waves that fit at once on the device).
uint32_t persistent_grid(int dim, bool boundary, bool volume, int bie, int est, int flavour, uint32_t block,
                         bool wavefront = false, bool multi = false);
bool supports_wavefront(int dim, bool boundary, bool volume, int bie, int est, int flavour);
// LDS bytes of a walk launch (per-lane traversal stacks)
size_t walk_lds_bytes(int flavour, bool volume, uint32_t block);

// Device BVH build (wob_bvh.hip): the float and quantized 4-wide nodes (at most `width` = 2..4
// children each), the primitive records and the leaf-order element ids of a scene, built on
// the current device from host vertex data (n * dim * dim floats).  The buffers are device allocations owned by the
// caller afterwards.  Returns an empty string on success.
struct DeviceBvh {
    float *fnodes = nullptr;
    uint32_t *qnodes = nullptr;
    float *prims = nullptr;
    uint32_t *prim_id = nullptr;
    uint32_t n_nodes = 0, depth = 0, stack_need = 0;
    size_t bytes = 0;
};
std::string build_bvh_device(int dim, uint32_t n, const float *h_vertices, DeviceBvh &out, int width = 4);

}  // namespace wobl
