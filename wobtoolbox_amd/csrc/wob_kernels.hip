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

#include <algorithm>

#include "wob_device.hpp"
#include "wob_launch.hpp"

using namespace wobdev;

namespace wobk {

using wobl::WalkEntry;
using wobl::WalkLaunch;

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

// entry of a global work item / point (binary search over the entry table)
__device__ __forceinline__ const WalkEntry &entry_of_item(const WalkLaunch &L, uint32_t item) {
    if (L.table == nullptr) return L.e0;
    uint32_t lo = 0, hi = L.n_entries - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (L.table[mid].item_begin <= item) lo = mid; else hi = mid - 1;
    }
    return L.table[lo];
}

__device__ __forceinline__ const WalkEntry &entry_of_point(const WalkLaunch &L, uint32_t gp) {
    if (L.table == nullptr) return L.e0;
    uint32_t lo = 0, hi = L.n_entries - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (L.table[mid].point_begin <= gp) lo = mid; else hi = mid - 1;
    }
    return L.table[lo];
}

// One work item: `paths_per_chunk` consecutive paths of one point, back to back.
template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY>
__device__ __forceinline__ void walk_item(const WalkLaunch &L, uint32_t item, uint32_t pass, uint32_t *stack,
                                          uint32_t stride) {
    constexpr int V = BOUNDARY ? 1 : D + 1;
    const WalkEntry &e = entry_of_item(L, item);
    const uint32_t local = item - e.item_begin;
    const uint32_t chunk = local / e.n_points;
    const uint32_t pt = local - chunk * e.n_points;
    const uint32_t j0 = chunk * e.paths_per_chunk;
    const uint32_t j1 = min(e.paths_per_point, j0 + e.paths_per_chunk);
    const uint64_t path_base = e.path_base + (uint64_t)pass * e.paths_per_point;
    const DevScene S = e.scene;

    Model<D, BIE> m;
    m.k.eps = e.eps;
    m.dsign = e.dsign;
    m.c = e.fkesc;
    const WalkParams wp{e.path_length, e.nrc, e.flags & 1u, (e.flags >> 1) & 1u, e.cv_samples, (e.flags >> 2) & 1u};

    Kahan ks[V], kq[V];
#pragma unroll
    for (int k = 0; k < V; ++k) { ks[k].init(); kq[k].init(); }

    if constexpr (BOUNDARY) {
        const float *r = e.points + (size_t)pt * (2 * D + 4);
        BP<D> x;
#pragma unroll
        for (int k = 0; k < D; ++k) { x.p[k] = r[k]; x.n[k] = r[D + k]; }
        x.bv = r[2 * D + 1];
        x.type = (int)r[2 * D + 2];
        x.alpha = r[2 * D + 3];
        x.elem = 0xFFFFFFFFu;
#pragma unroll
        for (int k = 0; k < D; ++k) x.bw[k] = 0.f;
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(L.seed, (uint32_t)(path_base + j), e.point_base + pt, e.stream);
            const float v = boundary_path<D, BIE, EST, FL, STACK>(S, m, wp, x, rng, stack, stride);
            ks[0].add(v);
            kq[0].add(v * v);
        }
    } else {
        float x[D];
#pragma unroll
        for (int k = 0; k < D; ++k) x[k] = e.points[(size_t)pt * D + k];
        for (uint32_t j = j0; j < j1; ++j) {
            Rng rng;
            rng.init(L.seed, (uint32_t)(path_base + j), e.point_base + pt, e.stream);
            float out[D + 1];
            domain_path<D, BIE, EST, FL, STACK>(S, m, wp, x, rng, stack, stride, out);
#pragma unroll
            for (int k = 0; k < V; ++k) { ks[k].add(out[k]); kq[k].add(out[k] * out[k]); }
        }
    }
    const size_t n_items = (size_t)e.n_points * e.chunks;
    float *part = L.partials + e.partial_base;
#pragma unroll
    for (int k = 0; k < V; ++k) {
        part[(size_t)k * n_items + local] = ks[k].s;
        part[(size_t)(V + k) * n_items + local] = kq[k].s;
    }
}

// Work items are numbered entry by entry, chunk-major inside an entry so that neighbouring
// lanes walk from neighbouring points.  Persistent mode: each wave takes 64 items at a time
// from a device counter until the items run out, so lanes of a finished wave move on to new
// (point, chunk) work at once.
template <int D, int BIE, int EST, int FL, int STACK, bool BOUNDARY>
__global__ __launch_bounds__(256) void walk_kernel(WalkLaunch L) {
    extern __shared__ uint32_t lds_stack[];
    uint32_t *stack = lds_stack + threadIdx.x;
    const uint32_t stride = blockDim.x;
    const uint32_t pass = L.pass_counter ? *L.pass_counter : 0u;
    if (!L.persistent) {
        const uint32_t item = blockIdx.x * blockDim.x + threadIdx.x;
        if (item < L.total_items) walk_item<D, BIE, EST, FL, STACK, BOUNDARY>(L, item, pass, stack, stride);
        return;
    }
    const uint32_t lane = threadIdx.x & 63u;
    while (true) {
        uint32_t base = 0;
        if (lane == 0) base = atomicAdd(L.work_counter, 64u);
        base = __shfl(base, 0, 64);
        if (base >= L.total_items) break;
        const uint32_t item = base + lane;
        if (item < L.total_items) walk_item<D, BIE, EST, FL, STACK, BOUNDARY>(L, item, pass, stack, stride);
    }
}

// Per point: FP64 sum of its chunk partials in chunk order, added into the entry's outputs.
__global__ void reduce_kernel(WalkLaunch L, uint32_t V) {
    const uint32_t gp = blockIdx.x * blockDim.x + threadIdx.x;
    if (gp >= L.total_points) return;
    const WalkEntry &e = entry_of_point(L, gp);
    const uint32_t pt = gp - e.point_begin;
    const size_t n_items = (size_t)e.n_points * e.chunks;
    const float *part = L.partials + e.partial_base;
    for (uint32_t k = 0; k < V; ++k) {
        double s = 0.0, q = 0.0;
        for (uint32_t c = 0; c < e.chunks; ++c) {
            s += (double)part[k * n_items + (size_t)c * e.n_points + pt];
            q += (double)part[(V + k) * n_items + (size_t)c * e.n_points + pt];
        }
        e.sum[(size_t)pt * V + k] += s;
        e.sumsq[(size_t)pt * V + k] += q;
    }
}

// end of a pass: clear the persistent work counter, advance the pass counter
__global__ void finalize_kernel(uint32_t *work_counter, uint32_t *pass_counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (work_counter) *work_counter = 0u;
        if (pass_counter) *pass_counter += 1u;
    }
}

__global__ void fill_u32_kernel(uint32_t *p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
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
    b.elem = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < D; ++k) b.bw[k] = 0.f;
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

template <int D>
__global__ void sample_boundary_kernel(DevScene S, uint32_t n, uint64_t seed, uint32_t stream, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng rng;
    rng.init(seed, i, 0u, stream);
    const Sample<D> z = sample_uniform<D>(S, rng);
    float *r = out + (size_t)i * (2 * D + 4);
#pragma unroll
    for (int k = 0; k < D; ++k) { r[k] = z.y.p[k]; r[D + k] = z.y.n[k]; }
    r[2 * D] = 0.5f;
    r[2 * D + 1] = z.y.bv;
    r[2 * D + 2] = (float)z.y.type;
    r[2 * D + 3] = z.y.alpha;
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

template <int D, int BIE, int EST, int FL, bool BOUNDARY> struct WalkKernel {
    static constexpr bool lines = FL == FL_LINE_LINE || FL == FL_LINE_BV || FL == FL_FWD_LINE;
    static constexpr int STACK = lines ? kStackEntries : 0;
    static size_t lds(uint32_t block) { return lines ? (size_t)STACK * block * sizeof(uint32_t) : 0; }
    static const void *fn() { return (const void *)&walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY>; }
    static hipError_t launch(const WalkLaunch &a, uint32_t block, uint32_t grid, hipStream_t st) {
        if (grid == 0) grid = (a.total_items + block - 1) / block;
        hipLaunchKernelGGL((walk_kernel<D, BIE, EST, FL, STACK, BOUNDARY>), dim3(grid), dim3(block), lds(block), st, a);
        return hipGetLastError();
    }
    static uint32_t resident_grid(uint32_t block) {
        int per_cu = 0, dev = 0, cus = 0;
        (void)hipGetDevice(&dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)fn(), (int)block, lds(block)) != hipSuccess)
            per_cu = 1;
        return (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    }
};

// dispatch (model, estimator, flavour) -> instantiation; F(kernel-type) is called with the chosen one
template <int D, bool BOUNDARY, typename F> static auto dispatch(int bie, int est, int fl, F &&f) {
    if (est == kBackward) {
        if (bie == kDL && fl == FL_LINE_LINE) return f(WalkKernel<D, kDL, kBackward, FL_LINE_LINE, BOUNDARY>{});
        if (bie == kDL) return f(WalkKernel<D, kDL, kBackward, FL_UNIFORM, BOUNDARY>{});
        if (bie == kDirect && fl == FL_LINE_BV) return f(WalkKernel<D, kDirect, kBackward, FL_LINE_BV, BOUNDARY>{});
        if (bie == kDirect) return f(WalkKernel<D, kDirect, kBackward, FL_UNIFORM, BOUNDARY>{});
        return f(WalkKernel<D, kSL, kBackward, FL_UNIFORM, BOUNDARY>{});
    }
    if (bie == kSL && fl == FL_FWD_LINE) return f(WalkKernel<D, kSL, kForward, FL_FWD_LINE, BOUNDARY>{});
    if (bie == kSL) return f(WalkKernel<D, kSL, kForward, FL_FWD_UNIFORM, BOUNDARY>{});
    if (bie == kDL) return f(WalkKernel<D, kDL, kForward, FL_FWD_UNIFORM, BOUNDARY>{});
    return f(WalkKernel<D, kDirect, kForward, FL_FWD_UNIFORM, BOUNDARY>{});
}

template <typename F> static auto dispatch_all(int dim, bool boundary, int bie, int est, int fl, F &&f) {
    if (dim == 2) return boundary ? dispatch<2, true>(bie, est, fl, f) : dispatch<2, false>(bie, est, fl, f);
    return boundary ? dispatch<3, true>(bie, est, fl, f) : dispatch<3, false>(bie, est, fl, f);
}

hipError_t launch_walk(int dim, bool boundary, const WalkLaunch &a, int bie, int est, int fl, uint32_t block,
                       uint32_t grid, hipStream_t st) {
    return dispatch_all(dim, boundary, bie, est, fl, [&](auto k) { return decltype(k)::launch(a, block, grid, st); });
}

uint32_t persistent_grid(int dim, bool boundary, int bie, int est, int fl, uint32_t block) {
    return dispatch_all(dim, boundary, bie, est, fl, [&](auto k) { return decltype(k)::resident_grid(block); });
}

hipError_t launch_reduce(const WalkLaunch &a, uint32_t V, hipStream_t st) {
    const uint32_t block = 256;
    hipLaunchKernelGGL(reduce_kernel, dim3((a.total_points + block - 1) / block), dim3(block), 0, st, a, V);
    return hipGetLastError();
}

hipError_t launch_finalize(uint32_t *work_counter, uint32_t *pass_counter, hipStream_t st) {
    hipLaunchKernelGGL(finalize_kernel, dim3(1), dim3(64), 0, st, work_counter, pass_counter);
    return hipGetLastError();
}

hipError_t launch_fill_u32(uint32_t *p, uint32_t n, uint32_t v, hipStream_t st) {
    const uint32_t block = 256;
    hipLaunchKernelGGL(fill_u32_kernel, dim3((n + block - 1) / block), dim3(block), 0, st, p, n, v);
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

hipError_t launch_sample_boundary(int dim, const DevScene &S, uint32_t n, uint64_t seed, uint32_t stream, float *out,
                                  hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    if (dim == 2) hipLaunchKernelGGL(sample_boundary_kernel<2>, dim3(grid), dim3(block), 0, st, S, n, seed, stream, out);
    else hipLaunchKernelGGL(sample_boundary_kernel<3>, dim3(grid), dim3(block), 0, st, S, n, seed, stream, out);
    return hipGetLastError();
}

hipError_t launch_philox(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(philox_kernel, dim3(grid), dim3(block), 0, st, n, ctr, key, out);
    return hipGetLastError();
}

}  // namespace wobl
