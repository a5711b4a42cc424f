// Kernels and launchers of the MI355X walk-on-boundary engine.
//
// walk_kernel   one lane = one (point, chunk of paths) work item; paths of the chunk run
//               back to back (a path that stops early is followed at once by the next), the
//               lane Kahan-sums [u, grad u] and their squares in FP32 and writes per-item
//               partials (structure of arrays, coalesced).  Items are chunk-major so that
//               neighbouring lanes walk from neighbouring points.
// reduce_kernel per point, sums its chunk partials in FP64 in a fixed order and adds them
//               into the caller's FP64 accumulators (deterministic for fixed inputs).
// The deterministic parity kernels (kernels, weights, all-hit lists, Philox) share the
// device functions of the walk.
#include <hip/hip_runtime.h>

#include "wob_device.hpp"
#include "wob_launch.hpp"

using namespace wobdev;

namespace wobk {

using wobl::WalkArgs;

struct Kahan {
    float s, c;
    __device__ __forceinline__ void init() { s = 0.f; c = 0.f; }
    __device__ __forceinline__ void add(float v) {   // utils.cuh:158-164
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
};

template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY>
__global__ __launch_bounds__(256) void walk_kernel(WalkArgs a) {
    extern __shared__ uint32_t lds_stack[];
    constexpr int V = BOUNDARY ? 1 : D + 1;
    const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t n_items = a.n_points * a.chunks;
    if (item >= n_items) return;
    const uint32_t chunk = item / a.n_points;
    const uint32_t pt = item - chunk * a.n_points;
    const uint32_t j0 = chunk * a.paths_per_chunk;
    const uint32_t j1 = min(a.paths_per_point, j0 + a.paths_per_chunk);

    Model<D, BIE> m;
    m.k.eps = a.eps;
    m.dsign = a.dsign;
    m.c = a.fkesc;
    const WalkParams wp{a.path_length, a.nrc};
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;

    Kahan ks[V], kq[V];
#pragma unroll
    for (int k = 0; k < V; ++k) { ks[k].init(); kq[k].init(); }

    if constexpr (BOUNDARY) {
        const float *r = a.points + (size_t)pt * (2 * D + 4);
        BP<D> x;
#pragma unroll
        for (int k = 0; k < D; ++k) { x.p[k] = r[k]; x.n[k] = r[D + k]; }
        x.bv = r[2 * D + 1];
        x.type = (int)r[2 * D + 2];
        x.alpha = r[2 * D + 3];
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(a.seed, (uint32_t)(a.path_base + j), a.point_base + pt, a.stream);
            const float v = boundary_path<D, BIE, EST, FL, STACK>(a.scene, m, wp, x, rng, stack, stride);
            ks[0].add(v);
            kq[0].add(v * v);
        }
    } else {
        float x[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = a.points[(size_t)pt * D + k];
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(a.seed, (uint32_t)(a.path_base + j), a.point_base + pt, a.stream);
            float out[D + 1];
            domain_path<D, BIE, EST, FL, STACK>(a.scene, m, wp, x, rng, stack, stride, out);
#pragma unroll
            for (int k = 0; k < V; ++k) { ks[k].add(out[k]); kq[k].add(out[k] * out[k]); }
        }
    }
#pragma unroll
    for (int k = 0; k < V; ++k) {
        a.partials[(size_t)k * n_items + item] = ks[k].s;
        a.partials[(size_t)(V + k) * n_items + item] = kq[k].s;
    }
}

__global__ void reduce_kernel(const float *__restrict__ partials, uint32_t n_points, uint32_t chunks, uint32_t V,
                              double *__restrict__ sum, double *__restrict__ sumsq) {
    const uint32_t pt = blockIdx.x * blockDim.x + threadIdx.x;
    if (pt >= n_points) return;
    const size_t n_items = (size_t)n_points * chunks;
    for (uint32_t k = 0; k < V; ++k) {
        double s = 0.0, q = 0.0;
        for (uint32_t c = 0; c < chunks; ++c) {
            s += (double)partials[k * n_items + (size_t)c * n_points + pt];
            q += (double)partials[(V + k) * n_items + (size_t)c * n_points + pt];
        }
        sum[(size_t)pt * V + k] += s;
        sumsq[(size_t)pt * V + k] += q;
    }
}

// ---------------------------------------------------------------------------------------
// deterministic parity kernels
// ---------------------------------------------------------------------------------------
template <int D>
__global__ void eval_kernels_kernel(uint32_t n, const float *y, const float *x, const float *ny, const float *nx,
                                    float eps, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Kern<D> k{eps};
    float Y[D], X[D], NY[D], NX[D];
#pragma unroll
    for (int j = 0; j < D; ++j) { Y[j] = y[D * i + j]; X[j] = x[D * i + j]; NY[j] = ny[D * i + j]; NX[j] = nx[D * i + j]; }
    float *o = out + (size_t)i * (3 + 2 * D);
    o[0] = k.G(Y, X);
    float g[D];
    k.dG_dx(Y, X, g);
#pragma unroll
    for (int j = 0; j < D; ++j) o[1 + j] = g[j];
    o[1 + D] = k.dG_dny(Y, X, NY);
    k.d2G_dxdny(Y, X, NY, g);
#pragma unroll
    for (int j = 0; j < D; ++j) o[2 + D + j] = g[j];
    o[2 + 2 * D] = k.d2G_dnxdny(Y, X, NY, NX);
}

template <int D> __device__ __forceinline__ BP<D> load_record(const float *r) {
    BP<D> b;
#pragma unroll
    for (int k = 0; k < D; ++k) { b.p[k] = r[k]; b.n[k] = r[D + k]; }
    b.bv = r[2 * D + 1];
    b.type = (int)r[2 * D + 2];
    b.alpha = r[2 * D + 3];
    return b;
}

template <int D, int BIE>
__global__ void eval_weights_kernel(uint32_t n, float dsign, float eps, float fkesc, const float *xr, const float *yr,
                                    const float *inv_pdf, const float *xd, const float *u_split, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Model<D, BIE> m;
    m.k.eps = eps; m.dsign = dsign; m.c = fkesc;
    const BP<D> x = load_record<D>(xr + (size_t)i * (2 * D + 4));
    BP<D> y = load_record<D>(yr + (size_t)i * (2 * D + 4));
    float X[D];
#pragma unroll
    for (int k = 0; k < D; ++k) X[k] = xd[(size_t)i * D + k];
    const float inv = inv_pdf[i];
    float *o = out + (size_t)i * (5 + D);
    o[0] = m.fredholm_known(x, y, inv);
    // the first-kind split with the supplied uniform (bie_models.cuh:272-279, 517-524)
    float fu;
    float split = 0.f;
    if (BIE != kDL && x.type == kDirichlet && u_split[i] * (1.0f + fkesc) < 1.0f) {
        fu = 1.0f + fkesc;
        split = 1.f;
    } else {
        float iv = inv;
        if (BIE != kDL && x.type == kDirichlet) iv *= (1.0f + fkesc) / fkesc;
        fu = m.fredholm_unknown_kernel(x, y, iv);
    }
    o[1] = fu;
    o[2] = split;
    o[3] = m.solution_domain_unknown(X, y, inv);
    float g[D];
    m.gradient_domain_unknown(X, y, inv, g);
#pragma unroll
    for (int k = 0; k < D; ++k) o[4 + k] = g[k];
    o[4 + D] = m.solution_boundary_unknown(x, y, inv);
}

template <int D, int STACK>
__global__ __launch_bounds__(256) void intersect_kernel(DevScene S, uint32_t n, const float *origins, const float *dirs,
                                                        uint32_t max_hits, uint32_t *counts, uint32_t *ids, float *ts) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float o[D], d[D];
#pragma unroll
    for (int k = 0; k < D; ++k) { o[k] = origins[(size_t)i * D + k]; d[k] = dirs[(size_t)i * D + k]; }
    uint32_t c = 0;
    traverse<D, STACK>(S, o, d, lds_stack + threadIdx.x, blockDim.x, [&](uint32_t prim, float t, float, float) {
        if (c < max_hits) {
            ids[(size_t)i * max_hits + c] = S.prim_id[prim];
            ts[(size_t)i * max_hits + c] = t;
        }
        ++c;
    });
    counts[i] = c;
}

__global__ void philox_kernel(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t c0 = ctr[4 * i], c1 = ctr[4 * i + 1], c2 = ctr[4 * i + 2], c3 = ctr[4 * i + 3];
    philox4x32_10(c0, c1, c2, c3, key[2 * i], key[2 * i + 1]);
    out[4 * i] = c0; out[4 * i + 1] = c1; out[4 * i + 2] = c2; out[4 * i + 3] = c3;
}

}  // namespace wobk

// ---------------------------------------------------------------------------------------
// launchers
// ---------------------------------------------------------------------------------------
namespace wobl {

using namespace wobk;

template <int D, int BIE, int EST, int FL, bool BOUNDARY>
static hipError_t launch_walk_t(const WalkArgs &a, uint32_t block, hipStream_t st) {
    const uint32_t n_items = a.n_points * a.chunks;
    const uint32_t grid = (n_items + block - 1) / block;
    constexpr bool lines = FL == FL_LINE_LINE || FL == FL_LINE_BV || FL == FL_FWD_LINE;
    constexpr int STACK = lines ? kStackEntries : 0;
    const size_t lds = lines ? (size_t)STACK * block * sizeof(uint32_t) : 0;
    hipLaunchKernelGGL((walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY>), dim3(grid), dim3(block), lds, st, a);
    return hipGetLastError();
}

template <int D, bool BOUNDARY>
static hipError_t launch_walk_d(const WalkArgs &a, int bie, int est, int fl, uint32_t block, hipStream_t st) {
    if (est == kBackward) {
        if (bie == kDL && fl == FL_LINE_LINE) return launch_walk_t<D, kDL, kBackward, FL_LINE_LINE, BOUNDARY>(a, block, st);
        if (bie == kDL) return launch_walk_t<D, kDL, kBackward, FL_UNIFORM, BOUNDARY>(a, block, st);
        if (bie == kDirect && fl == FL_LINE_BV) return launch_walk_t<D, kDirect, kBackward, FL_LINE_BV, BOUNDARY>(a, block, st);
        if (bie == kDirect) return launch_walk_t<D, kDirect, kBackward, FL_UNIFORM, BOUNDARY>(a, block, st);
        return launch_walk_t<D, kSL, kBackward, FL_UNIFORM, BOUNDARY>(a, block, st);
    }
    if (bie == kSL && fl == FL_FWD_LINE) return launch_walk_t<D, kSL, kForward, FL_FWD_LINE, BOUNDARY>(a, block, st);
    if (bie == kSL) return launch_walk_t<D, kSL, kForward, FL_FWD_UNIFORM, BOUNDARY>(a, block, st);
    if (bie == kDL) return launch_walk_t<D, kDL, kForward, FL_FWD_UNIFORM, BOUNDARY>(a, block, st);
    return launch_walk_t<D, kDirect, kForward, FL_FWD_UNIFORM, BOUNDARY>(a, block, st);
}

hipError_t launch_walk(int dim, bool boundary, const WalkArgs &a, int bie, int est, int fl, uint32_t block,
                       hipStream_t st) {
    if (dim == 2) return boundary ? launch_walk_d<2, true>(a, bie, est, fl, block, st) : launch_walk_d<2, false>(a, bie, est, fl, block, st);
    return boundary ? launch_walk_d<3, true>(a, bie, est, fl, block, st) : launch_walk_d<3, false>(a, bie, est, fl, block, st);
}

hipError_t launch_reduce(const float *partials, uint32_t n_points, uint32_t chunks, uint32_t V, double *sum,
                         double *sumsq, hipStream_t st) {
    const uint32_t block = 256;
    hipLaunchKernelGGL(reduce_kernel, dim3((n_points + block - 1) / block), dim3(block), 0, st, partials, n_points,
                       chunks, V, sum, sumsq);
    return hipGetLastError();
}

hipError_t launch_eval_kernels(int dim, uint32_t n, const float *y, const float *x, const float *ny, const float *nx,
                               float eps, float *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    if (dim == 2) hipLaunchKernelGGL(eval_kernels_kernel<2>, dim3(grid), dim3(block), 0, st, n, y, x, ny, nx, eps, out);
    else hipLaunchKernelGGL(eval_kernels_kernel<3>, dim3(grid), dim3(block), 0, st, n, y, x, ny, nx, eps, out);
    return hipGetLastError();
}

hipError_t launch_eval_weights(int dim, int bie, float dsign, float eps, float fkesc, uint32_t n, const float *xr,
                               const float *yr, const float *inv, const float *xd, const float *us, float *out,
                               hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
#define WOB_EW(DD, BB) hipLaunchKernelGGL((eval_weights_kernel<DD, BB>), dim3(grid), dim3(block), 0, st, n, dsign, eps, fkesc, xr, yr, inv, xd, us, out)
    if (dim == 2) {
        if (bie == kSL) WOB_EW(2, kSL); else if (bie == kDL) WOB_EW(2, kDL); else WOB_EW(2, kDirect);
    } else {
        if (bie == kSL) WOB_EW(3, kSL); else if (bie == kDL) WOB_EW(3, kDL); else WOB_EW(3, kDirect);
    }
#undef WOB_EW
    return hipGetLastError();
}

hipError_t launch_intersect(int dim, const DevScene &S, uint32_t n, const float *o, const float *d, uint32_t max_hits,
                            uint32_t *counts, uint32_t *ids, float *ts, hipStream_t st) {
    const uint32_t block = 128, grid = (n + block - 1) / block;
    const size_t lds = (size_t)kStackEntries * block * sizeof(uint32_t);
    if (dim == 2)
        hipLaunchKernelGGL((intersect_kernel<2, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, o, d, max_hits, counts, ids, ts);
    else
        hipLaunchKernelGGL((intersect_kernel<3, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, o, d, max_hits, counts, ids, ts);
    return hipGetLastError();
}

hipError_t launch_philox(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(philox_kernel, dim3(grid), dim3(block), 0, st, n, ctr, key, out);
    return hipGetLastError();
}

}  // namespace wobl
