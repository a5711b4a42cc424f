// Kernels and launchers of the MI355X walk-on-boundary engine (the walk kernels themselves
// are in wob_walk.hip).
//
// reduce_kernel per point, sums its chunk partials in FP64 in a fixed order and adds them
//               into the caller's FP64 accumulators (deterministic for fixed inputs).
// The deterministic parity kernels (kernels, weights, all-hit lists, Philox) share the
// device functions of the walk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <type_traits>

#include "wob_device.hpp"
#include "wob_launch.hpp"

using namespace wobdev;

namespace wobk {

using wobl::WalkEntry;
using wobl::WalkLaunch;

__device__ __forceinline__ const WalkEntry &entry_of_point(const WalkLaunch &L, uint32_t gp) {
    if (L.table == nullptr) return L.e0;
    uint32_t lo = 0, hi = L.n_entries - 1;
    while (lo < hi) {
        const uint32_t mid = (lo + hi + 1) >> 1;
        if (L.table[mid].point_begin <= gp) lo = mid; else hi = mid - 1;
    }
    return L.table[lo];
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
    traverse<D, STACK>(S, o, d, lds_stack + threadIdx.x, blockDim.x, [&](bool hit, uint32_t prim, float t, float, float) {
        if (!hit) return;
        if (c < max_hits) {
            ids[(size_t)i * max_hits + c] = S.prim_id[prim];
            ts[(size_t)i * max_hits + c] = t;
        }
        ++c;
    });
    counts[i] = c;
}

// One line-intersection sample per domain point, with the walk's own sampler (line_sample:
// the work-sharing all-hit query, the uniform pick and the hit's attributes) on the Philox
// stream (seed, path, point i, stream).  Lanes past n leave before the query; line_query_ws
// shares work only among the lanes active at its call.
template <int D, int STACK>
__global__ __launch_bounds__(256) void line_sample_kernel(DevScene S, uint32_t n, const float *points, uint64_t seed,
                                                          uint32_t path, uint32_t stream, float *rec, float *inv_pdf) {
    extern __shared__ uint32_t lds_stack[];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = points[(size_t)i * D + k];
    Rng rng;
    rng.init(seed, path, i, stream);
    const Sample<D> z = line_sample<D, STACK>(S, x, 0xFFFFFFFFu, rng, lds_stack + threadIdx.x, blockDim.x);
    float *r = rec + (size_t)i * (2 * D + 4);
#pragma unroll
    for (int k = 0; k < D; ++k) { r[k] = z.y.p[k]; r[D + k] = z.y.n[k]; }
    r[2 * D] = 0.5f;
    r[2 * D + 1] = z.y.bv;
    r[2 * D + 2] = (float)z.y.type;
    r[2 * D + 3] = z.y.alpha;
    inv_pdf[i] = z.inv_pdf;
}

// per_point == 0: record i uses the Philox stream (seed, path i, point 0, stream);
// per_point == 1: (seed, path `path`, point i, stream) -- one stream per VPL of a pass.
template <int D>
__global__ void sample_boundary_kernel(DevScene S, uint32_t n, uint64_t seed, uint32_t stream, uint32_t path,
                                       uint32_t per_point, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Rng rng;
    if (per_point) rng.init(seed, path, i, stream);
    else rng.init(seed, i, 0u, stream);
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

size_t walk_lds_bytes(int fl, bool volume, uint32_t block) {
    const bool lines = fl == FL_LINE_LINE || fl == FL_LINE_BV || fl == FL_FWD_LINE || volume;
    return lines ? (size_t)line_query_lds_words<kStackEntries>() * block * sizeof(uint32_t) : 0;
}

hipError_t launch_walk(int dim, bool boundary, bool volume, const WalkLaunch &a, int bie, int est, int fl,
                       uint32_t block, uint32_t grid, hipStream_t st) {
    if (dim == 2)
        return volume ? launch_walk_21(boundary, a, bie, est, fl, block, grid, st)
                      : launch_walk_20(boundary, a, bie, est, fl, block, grid, st);
    return volume ? launch_walk_31(boundary, a, bie, est, fl, block, grid, st)
                  : launch_walk_30(boundary, a, bie, est, fl, block, grid, st);
}

uint32_t persistent_grid(int dim, bool boundary, bool volume, int bie, int est, int fl, uint32_t block, bool wavefront,
                         bool multi) {
    if (dim == 2)
        return volume ? persistent_grid_21(boundary, bie, est, fl, block, wavefront, multi)
                      : persistent_grid_20(boundary, bie, est, fl, block, wavefront, multi);
    return volume ? persistent_grid_31(boundary, bie, est, fl, block, wavefront, multi)
                  : persistent_grid_30(boundary, bie, est, fl, block, wavefront, multi);
}

bool supports_wavefront(int dim, bool boundary, bool volume, int bie, int est, int fl) {
    if (dim == 2)
        return volume ? supports_wavefront_21(boundary, bie, est, fl) : supports_wavefront_20(boundary, bie, est, fl);
    return volume ? supports_wavefront_31(boundary, bie, est, fl) : supports_wavefront_30(boundary, bie, est, fl);
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
    const size_t lds = (size_t)(kStackEntries + 1) * block * sizeof(uint32_t);
    if (dim == 2)
        hipLaunchKernelGGL((intersect_kernel<2, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, o, d, max_hits, counts, ids, ts);
    else
        hipLaunchKernelGGL((intersect_kernel<3, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, o, d, max_hits, counts, ids, ts);
    return hipGetLastError();
}

hipError_t launch_sample_lines(int dim, const DevScene &S, uint32_t n, const float *points, uint64_t seed,
                               uint32_t path, uint32_t stream, float *rec, float *inv_pdf, hipStream_t st) {
    const uint32_t block = 128, grid = (n + block - 1) / block;
    const size_t lds = (size_t)line_query_lds_words<kStackEntries>() * block * sizeof(uint32_t);
    if (dim == 2)
        hipLaunchKernelGGL((line_sample_kernel<2, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, points, seed,
                           path, stream, rec, inv_pdf);
    else
        hipLaunchKernelGGL((line_sample_kernel<3, kStackEntries>), dim3(grid), dim3(block), lds, st, S, n, points, seed,
                           path, stream, rec, inv_pdf);
    return hipGetLastError();
}

hipError_t launch_sample_boundary(int dim, const DevScene &S, uint32_t n, uint64_t seed, uint32_t stream, float *out,
                                  hipStream_t st, int64_t path) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    const uint32_t pp = path >= 0 ? 1u : 0u, pw = path >= 0 ? (uint32_t)path : 0u;
    if (dim == 2) hipLaunchKernelGGL(sample_boundary_kernel<2>, dim3(grid), dim3(block), 0, st, S, n, seed, stream, pw, pp, out);
    else hipLaunchKernelGGL(sample_boundary_kernel<3>, dim3(grid), dim3(block), 0, st, S, n, seed, stream, pw, pp, out);
    return hipGetLastError();
}

// The harmonic control variate's fit: the augmented Gram matrix G = [B t]^T [B t] of the rows
// B (basis 1, phi_1 .. phi_m, each as the datum of the record's type: value at Dirichlet,
// normal derivative at Neumann, normal derivative + alpha value at Robin records) and the
// targets t, in FP64, Q = m + 2 columns (column Q - 1 is the target).  Records are taken in
// chunks; thread (a, b) of a chunk sums row a times row b over the chunk's records in order,
// and the chunks' sums are added in order (deterministic).  src: m positions (D doubles).
template <int D>
__device__ __forceinline__ double hcv_row(const float *r, const double *src, uint32_t m, uint32_t c) {
    if (c == m + 1) return (double)r[2 * D + 1];   // the target
    const int type = (int)r[2 * D + 2];
    const double alpha = (double)r[2 * D + 3];
    if (c == 0) return type == kDirichlet ? 1.0 : type == kNeumann ? 0.0 : alpha;
    const double *s = src + (size_t)(c - 1) * D;
    double z = 0.0, dn = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const double d = (double)r[k] - s[k];
        z += d * d;
        dn += d * (double)r[D + k];
    }
    double val, der;
    if constexpr (D == 3) {
        const double ri = 1.0 / sqrt(z);
        val = ri;
        der = -dn * ri * ri * ri;
    } else {
        val = log2(z);
        der = 2.8853900817779268 * dn / z;   // 2 / ln 2
    }
    return type == kDirichlet ? val : type == kNeumann ? der : der + alpha * val;
}

// Block (a, b tile, chunk c) sums the products over the records of chunk c into the partial
// matrix of that chunk; hcv_gram_sum_kernel then adds the chunks in order.
constexpr uint32_t kGramChunk = 256;

template <int D>
__global__ __launch_bounds__(64) void hcv_gram_kernel(const float *rec, uint32_t n, const double *src, uint32_t m,
                                                      double *partial) {
    const uint32_t Q = m + 2;
    const uint32_t a = blockIdx.x, b = blockIdx.y * 64 + threadIdx.x, c = blockIdx.z;
    if (b >= Q || b < a) return;
    const uint32_t i0 = c * kGramChunk, i1 = min(n, i0 + kGramChunk);
    double acc = 0.0;
    for (uint32_t i = i0; i < i1; ++i) {
        const float *r = rec + (size_t)i * (2 * D + 4);
        acc += hcv_row<D>(r, src, m, a) * hcv_row<D>(r, src, m, b);
    }
    partial[((size_t)c * Q + a) * Q + b] = acc;
}

__global__ void hcv_gram_sum_kernel(const double *partial, uint32_t Q, uint32_t chunks, double *G) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Q * Q) return;
    const uint32_t a = t / Q, b = t - a * Q;
    const uint32_t lo = min(a, b), hi = max(a, b);   // partials hold the upper triangle
    double acc = 0.0;
    for (uint32_t c = 0; c < chunks; ++c) acc += partial[((size_t)c * Q + lo) * Q + hi];
    G[t] = acc;
}

size_t hcv_gram_scratch(uint32_t n, uint32_t m) {
    const size_t Q = m + 2, chunks = (n + kGramChunk - 1) / kGramChunk;
    return chunks * Q * Q * sizeof(double);
}

hipError_t launch_hcv_gram(int dim, const float *rec, uint32_t n, const double *src, uint32_t m, double *partial,
                           double *G, hipStream_t st) {
    const uint32_t Q = m + 2, chunks = (n + kGramChunk - 1) / kGramChunk;
    const dim3 grid(Q, (Q + 63) / 64, chunks);
    if (dim == 2) hipLaunchKernelGGL(hcv_gram_kernel<2>, grid, dim3(64), 0, st, rec, n, src, m, partial);
    else hipLaunchKernelGGL(hcv_gram_kernel<3>, grid, dim3(64), 0, st, rec, n, src, m, partial);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(hcv_gram_sum_kernel, dim3((Q * Q + 255) / 256), dim3(256), 0, st, partial, Q, chunks, G);
    return hipGetLastError();
}

// The walk's residual datum (hcv_residual) at each record: the fit's residual as the walk sees it
template <int D> __global__ void hcv_residual_kernel(DevScene S, const float *rec, uint32_t n, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float *r = rec + (size_t)i * (2 * D + 4);
    BP<D> y;
#pragma unroll
    for (int k = 0; k < D; ++k) { y.p[k] = r[k]; y.n[k] = r[D + k]; y.bw[k] = 0.f; }
    y.bv = r[2 * D + 1];
    y.type = (int)r[2 * D + 2];
    y.alpha = r[2 * D + 3];
    y.elem = 0xFFFFFFFFu;
    out[i] = hcv_residual<D>(S, y).bv;
}

hipError_t launch_hcv_residual(int dim, const DevScene &S, const float *rec, uint32_t n, float *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 256, grid = (n + block - 1) / block;
    if (dim == 2) hipLaunchKernelGGL(hcv_residual_kernel<2>, dim3(grid), dim3(block), 0, st, S, rec, n, out);
    else hipLaunchKernelGGL(hcv_residual_kernel<3>, dim3(grid), dim3(block), 0, st, S, rec, n, out);
    return hipGetLastError();
}

// u_h and grad u_h of the scene's harmonic control variate (the walk's own functions)
template <int D> __global__ void hcv_eval_kernel(DevScene S, uint32_t n, const float *pts, float *out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x[D], g[D];
#pragma unroll
    for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
    hcv_gradient<D>(S, x, g);
    out[(size_t)i * (D + 1)] = hcv_value<D>(S, x);
#pragma unroll
    for (int k = 0; k < D; ++k) out[(size_t)i * (D + 1) + 1 + k] = g[k];
}

hipError_t launch_hcv_eval(int dim, const DevScene &S, uint32_t n, const float *pts, float *out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint32_t block = 256, grid = (n + block - 1) / block;
    if (dim == 2) hipLaunchKernelGGL(hcv_eval_kernel<2>, dim3(grid), dim3(block), 0, st, S, n, pts, out);
    else hipLaunchKernelGGL(hcv_eval_kernel<3>, dim3(grid), dim3(block), 0, st, S, n, pts, out);
    return hipGetLastError();
}

// Record k of the fused area-guide table: row 0 is the guide row (i = guide[k], cdf[i - 1],
// cdf[i], cdf[i + 1]); rows 1-3 are the geometry of the element that covers more of entry k's
// interval [k/m, (k+1)/m), i or i + 1 (the search for a uniform in the interval ends there
// more often); bit 31 of the index word says which (element indices stay below 2^27).
__global__ void fused_records_kernel(const uint32_t *guide, uint32_t m, const float *cdf, uint32_t n, const float4 *geo,
                                     float4 *out) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= m) return;
    const uint32_t i = guide[k];
    const float lo = (float)k / (float)m, hi = (float)(k + 1) / (float)m;
    const float ci = cdf[i], cn = i + 1 < n ? cdf[i + 1] : ci;
    const bool next = i + 1 < n && fminf(cn, hi) - fmaxf(ci, lo) > fminf(ci, hi) - lo;
    const uint32_t g = next ? i + 1 : i;
    out[4 * (size_t)k] = make_float4(__uint_as_float(i | (next ? 0x80000000u : 0u)), i > 0 ? cdf[i - 1] : 0.f, ci,
                                     i + 1 < n ? cn : 0.f);
    out[4 * (size_t)k + 1] = geo[4 * (size_t)g];
    out[4 * (size_t)k + 2] = geo[4 * (size_t)g + 1];
    out[4 * (size_t)k + 3] = geo[4 * (size_t)g + 2];
}

hipError_t launch_fused_records(const uint32_t *guide, uint32_t m, const float *cdf, uint32_t n, const float4 *geo,
                                float4 *out, hipStream_t st) {
    if (m == 0) return hipSuccess;
    const uint32_t block = 256, grid = (m + block - 1) / block;
    hipLaunchKernelGGL(fused_records_kernel, dim3(grid), dim3(block), 0, st, guide, m, cdf, n, geo, out);
    return hipGetLastError();
}

hipError_t launch_philox(uint32_t n, const uint32_t *ctr, const uint32_t *key, uint32_t *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(philox_kernel, dim3(grid), dim3(block), 0, st, n, ctr, key, out);
    return hipGetLastError();
}

}  // namespace wobl
