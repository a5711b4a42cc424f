// Host-visible launch interface between the C ABI (wob_capi.cpp) and the kernels
// (wob_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wob_device.hpp"

namespace wobl {

// Per-lane LDS traversal stack (uint32 entries).  The scene builder rejects BVHs whose
// depth-first stack bound exceeds this.
constexpr int kStackEntries = 32;

struct WalkArgs {
    wobdev::DevScene scene;
    const float *points;      // domain points (n * D) or boundary records (n * (2D + 4))
    float *partials;          // 2 * V * n_points * chunks floats (structure of arrays)
    uint32_t n_points;
    uint32_t chunks;          // work items per point
    uint32_t paths_per_point;
    uint32_t paths_per_chunk;
    uint32_t point_base;
    uint32_t stream;
    uint64_t seed;
    uint64_t path_base;
    float dsign, eps, fkesc;
    uint32_t path_length, nrc;
};

hipError_t launch_walk(int dim, bool boundary, const WalkArgs &a, int bie, int est, int flavour, uint32_t block,
                       hipStream_t st);
hipError_t launch_reduce(const float *partials, uint32_t n_points, uint32_t chunks, uint32_t V, double *sum,
                         double *sumsq, hipStream_t st);
hipError_t launch_eval_kernels(int dim, uint32_t n, const float *y, const float *x, const float *ny, const float *nx,
                               float eps, float *out, hipStream_t st);
hipError_t launch_eval_weights(int dim, int bie, float dsign, float eps, float fkesc, uint32_t n, const float *xr,
                               const float *yr, const float *inv, const float *xd, const float *us, float *out,
                               hipStream_t st);
hipError_t launch_intersect(int dim, const wobdev::DevScene &S, uint32_t n, const float *o, const float *d,
                            uint32_t max_hits, uint32_t *counts, uint32_t *ids, float *ts, hipStream_t st);
hipError_t launch_philox(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t st);

}  // namespace wobl
