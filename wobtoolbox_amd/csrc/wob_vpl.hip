// Virtual-point-light (VPL) aggregation: for every evaluation point x and every stored path
// start y_i with contribution c_i, sum c_i * solution_domain_unknown(x, y_i, |boundary|) and
// c_i * gradient_domain_unknown(x, y_i, |boundary|) -- the AggregateFunc of
// apps/potential_flow_with_vpl.cu:44-71 -- for all points at once.
//
// The sum is a dense (points x VPLs) evaluation, bound by the FP32 VALU.  Design:
//   * vpl_stage_kernel folds each VPL's contribution, the model's per-type coefficients
//     (bie_models.cuh:228-243, 341-356, 428-462) and the kernels' constants into one record
//     of 4 or 8 floats, so the pair loop reads nothing but position, normal and two scales.
//   * vpl_aggregate_kernel: a thread owns kPts evaluation points; all lanes of a workgroup
//     walk the same VPL sequence, so the VPL index is uniform and the compiler reads the
//     records with scalar loads (SGPR operands, no LDS and no per-lane memory traffic).
//     blockIdx.y splits the VPL range so that small point sets still fill the 256 CUs.
//   * Per point and component, the terms of each group of kGroup consecutive VPLs are added
//     in FP32 and the group sums are Kahan-summed (utils.cuh:148-166; the error stays that of
//     a compensated sum plus kGroup - 1 roundings per group), each split's compensated sum is
//     written in FP64, and vpl_reduce_kernel adds the splits
//     in split order into the caller's FP64 accumulators (deterministic).
//
// Pair terms, r = y - x, z = |r|^2, with the reference's clamps (equation_kernel.cuh:67-145):
//   2D  G         = -log(max(sqrt z, eps)) / 2pi     = -(ln 2 / 4pi) log2(max(z, eps^2))
//       dG_dx     =  r / (2pi sqrt z max(sqrt z, eps)) = r / (2pi max(z, eps sqrt z))
//       dG_dny    = -(r.n) / (2pi max(z, eps sqrt z))
//       d2G_dxdny =  (n - 2 (r.n) r / z) / (2pi max(z, eps))
//   3D  G         =  1 / (4pi max(sqrt z, eps))       =  min(rsqrt z, 1/eps) / 4pi
//       dG_dx     =  r rsqrt(z) / (4pi max(z, eps))
//       dG_dny    = -(r.n) rsqrt(z) / (4pi max(z, eps))
//       d2G_dxdny =  (n - 3 (r.n) r / z) / (4pi max(z^1.5, eps))
// evaluated with the hardware log2 / rcp / rsqrt (about 1 ulp each); r/|r| is taken as r
// when z == 0, as the reference does.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "wob_device.hpp"
#include "wob_launch.hpp"

#pragma clang fp contract(off)

namespace wobv {

using wobdev::kDirect;
using wobdev::kDirichlet;
using wobdev::kDL;
using wobdev::kNeumann;
using wobdev::kPi;
using wobdev::kSL;

constexpr int kBlock = 256;
constexpr int kPts = 2;   // evaluation points per thread
constexpr int kGroup = 8; // VPLs whose terms are summed before one compensated addition

template <int D, int M> constexpr int rec_floats() { return (M == kSL) ? 4 : 8; }

// Record layouts (floats):
//   SL     2D [px py aU aG]                 3D [px py pz a4]
//   DL     2D [px py nx ny bU . . .]        3D [px py pz nx ny nz bU .]
//   Direct 2D [px py nx ny aU aG bU .]      3D [px py pz nx ny nz a4 bU]
// u += aU * log2(max(z, eps^2)) (2D) or a4 * min(rsqrt z, 1/eps) (3D) for the G part and
// bU * (r.n) * s for the dG_dny part; the gradient parts use aG (2D), a4 (3D) and -bU.
template <int D, int M>
__global__ void vpl_stage_kernel(const float *__restrict__ rec_in, const float *__restrict__ contrib, uint32_t n,
                                 float inv_pdf, float dsign, float *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    constexpr int W = 2 * D + 4, K = rec_floats<D, M>();
    const float *r = rec_in + (size_t)i * W;
    const float w = contrib[i] * inv_pdf;
    // model coefficients: u = a G + b dG_dny, grad u = a dG_dx + b d2G_dxdny
    float a = 0.f, b = 0.f;
    if constexpr (M == kSL) {
        a = w;                                      // bie_models.cuh:228-231, 240-243
    } else if constexpr (M == kDL) {
        b = -w;                                     // :341-344, 353-356
    } else {                                        // :428-436, 452-462
        const int type = (int)r[2 * D + 2];
        const float alpha = r[2 * D + 3];
        if (type == kDirichlet) a = dsign * w;
        else if (type == kNeumann) b = -dsign * w;
        else { a = -dsign * w * alpha; b = -dsign * w; }
    }
    float o[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < D; ++k) o[k] = r[k];
    if constexpr (M != kSL) {
#pragma unroll
        for (int k = 0; k < D; ++k) o[D + k] = r[D + k];
    }
    constexpr float kLn2 = 0.693147180559945309f;
    const float g2 = 1.0f / (2.0f * kPi), g3 = 1.0f / (4.0f * kPi);
    if constexpr (D == 2) {
        const float aU = a * (-kLn2 / (4.0f * kPi)), aG = a * g2, bU = -b * g2;
        if constexpr (M == kSL) { o[2] = aU; o[3] = aG; }
        else if constexpr (M == kDL) { o[4] = bU; }
        else { o[4] = aU; o[5] = aG; o[6] = bU; }
    } else {
        const float a4 = a * g3, bU = -b * g3;
        if constexpr (M == kSL) { o[3] = a4; }
        else if constexpr (M == kDL) { o[6] = bU; }
        else { o[6] = a4; o[7] = bU; }
    }
#pragma unroll
    for (int k = 0; k < K; ++k) out[(size_t)i * K + k] = o[k];
}

__device__ __forceinline__ void kahan_add(float &s, float &c, float t) {
    const float y = t - c;
    const float u = s + y;
    c = (u - s) - y;
    s = u;
}

// One (point, VPL) pair: adds the value and gradient terms into t[0..D].
template <int D, int M>
__device__ __forceinline__ void pair_terms(const float *x, const float *q, float eps, float eps2,
                                           float inv_eps, float log2_eps2, float *t) {
    constexpr bool HA = M != kDL, HB = M != kSL;
    float r[D];
#pragma unroll
    for (int k = 0; k < D; ++k) r[k] = q[k] - x[k];
    float z = r[0] * r[0];
#pragma unroll
    for (int k = 1; k < D; ++k) z = __builtin_fmaf(r[k], r[k], z);
    float rn = 0.f;
    if constexpr (HB) {
        rn = r[0] * q[D];
#pragma unroll
        for (int k = 1; k < D; ++k) rn = __builtin_fmaf(r[k], q[D + k], rn);
    }
    if constexpr (D == 2) {
        // common case sqrt z >= eps: 1/z, log2 z and 1 / max(z, eps sqrt z) = 1/z.  A wave takes
        // the clamped forms only when one of its lanes is that close to the VPL (z == 0
        // included: there r/|r| = r = 0 and the reference's terms stay finite).
        float rz = __builtin_amdgcn_rcpf(z);
        float inv1 = rz;
        float lz = 0.f;
        if constexpr (HA) lz = __builtin_amdgcn_logf(z);
        const bool near = z < eps2;
        if (__builtin_amdgcn_ballot_w64(near) != 0) {
            asm volatile("");   // keeps the branch (no speculated rsqrt on the common path)
            if (near) {
                inv1 = z > 0.f ? __builtin_amdgcn_rsqf(z) * inv_eps : 0.f;
                rz = z > 0.f ? rz : 0.f;
                lz = log2_eps2;
            }
        }
        if constexpr (HA) {
            const float aU = M == kSL ? q[2] : q[4], aG = M == kSL ? q[3] : q[5];
            t[0] = __builtin_fmaf(aU, lz, t[0]);
            const float w = aG * inv1;
            t[1] = __builtin_fmaf(w, r[0], t[1]);
            t[2] = __builtin_fmaf(w, r[1], t[2]);
        }
        if constexpr (HB) {
            const float bU = M == kDL ? q[4] : q[6];
            t[0] = __builtin_fmaf(bU * rn, inv1, t[0]);
            // d2G_dxdny part: -bU (n - 2 (r.n) r / z) / max(z, eps)
            const float inv2 = z >= eps ? rz : inv_eps;
            const float w = -bU * inv2, h = 2.0f * rn * rz;
            t[1] = __builtin_fmaf(w, q[2] - h * r[0], t[1]);
            t[2] = __builtin_fmaf(w, q[3] - h * r[1], t[2]);
        }
    } else {
        const float rs = __builtin_amdgcn_rsqf(z);                              // inf at z == 0
        const float rs0 = z > 0.f ? rs : 0.f;
        const float rz = rs0 * rs0;                                             // 1/z, 0 at z == 0
        const float inv2 = z >= eps ? rz : inv_eps;                             // 1 / max(z, eps)
        if constexpr (HA) {
            const float a4 = M == kSL ? q[3] : q[6];
            t[0] = __builtin_fmaf(a4, fminf(rs, inv_eps), t[0]);
            const float w = a4 * rs0 * inv2;
#pragma unroll
            for (int k = 0; k < 3; ++k) t[1 + k] = __builtin_fmaf(w, r[k], t[1 + k]);
        }
        if constexpr (HB) {
            const float bU = M == kDL ? q[6] : q[7];
            t[0] = __builtin_fmaf(bU * rn, rs0 * inv2, t[0]);
            // 1 / max(z^1.5, eps) and the d2G_dxdny part -bU (n - 3 (r.n) r / z) / max(z^1.5, eps)
            const float inv3 = z * (z * rs0) >= eps ? rz * rs0 : inv_eps;
            const float w = -bU * inv3, h = 3.0f * rn * rz;
#pragma unroll
            for (int k = 0; k < 3; ++k) t[1 + k] = __builtin_fmaf(w, q[3 + k] - h * r[k], t[1 + k]);
        }
    }
}

// partial[(split * n_pts + i) * V + v]: compensated sum of split `split` for point i.
template <int D, int M>
__global__ __launch_bounds__(kBlock) void vpl_aggregate_kernel(const float *__restrict__ pts, uint32_t n_pts,
                                                               const float *__restrict__ rec, uint32_t n_vpl,
                                                               uint32_t chunk, float eps, double *__restrict__ partial) {
    constexpr int V = D + 1, K = rec_floats<D, M>();
    const uint32_t split = blockIdx.y;
    const uint32_t j0 = split * chunk;
    const uint32_t j1 = min(j0 + chunk, n_vpl);
    const float eps2 = eps * eps, inv_eps = 1.0f / eps, log2_eps2 = __builtin_amdgcn_logf(eps2);
    float x[kPts][D], s[kPts][V], c[kPts][V];
    uint32_t idx[kPts];
#pragma unroll
    for (int p = 0; p < kPts; ++p) {
        idx[p] = blockIdx.x * (kBlock * kPts) + p * kBlock + threadIdx.x;
        const uint32_t ii = idx[p] < n_pts ? idx[p] : n_pts - 1;   // out-of-range lanes repeat a point
#pragma unroll
        for (int k = 0; k < D; ++k) x[p][k] = pts[(size_t)ii * D + k];
#pragma unroll
        for (int v = 0; v < V; ++v) { s[p][v] = 0.f; c[p][v] = 0.f; }
    }
    // groups of kGroup VPLs: the group's terms are summed plainly, the group sum is Kahan-added
    uint32_t j = j0;
    for (; j + kGroup <= j1; j += kGroup) {
        float t[kPts][V];
#pragma unroll
        for (int p = 0; p < kPts; ++p)
#pragma unroll
            for (int v = 0; v < V; ++v) t[p][v] = 0.f;
        // the group's records, read up front (scalar loads: the address is uniform)
        float q[kGroup][K];
#pragma unroll
        for (int g = 0; g < kGroup; ++g)
#pragma unroll
            for (int k = 0; k < K; ++k) q[g][k] = rec[(size_t)(j + g) * K + k];
#pragma unroll
        for (int g = 0; g < kGroup; ++g)
#pragma unroll
            for (int p = 0; p < kPts; ++p) pair_terms<D, M>(x[p], q[g], eps, eps2, inv_eps, log2_eps2, t[p]);
#pragma unroll
        for (int p = 0; p < kPts; ++p)
#pragma unroll
            for (int v = 0; v < V; ++v) kahan_add(s[p][v], c[p][v], t[p][v]);
    }
    for (; j < j1; ++j) {
        const float *q = rec + (size_t)j * K;
#pragma unroll
        for (int p = 0; p < kPts; ++p) {
            float t[V];
#pragma unroll
            for (int v = 0; v < V; ++v) t[v] = 0.f;
            pair_terms<D, M>(x[p], q, eps, eps2, inv_eps, log2_eps2, t);
#pragma unroll
            for (int v = 0; v < V; ++v) kahan_add(s[p][v], c[p][v], t[v]);
        }
    }
#pragma unroll
    for (int p = 0; p < kPts; ++p) {
        if (idx[p] >= n_pts) continue;
        double *o = partial + ((size_t)split * n_pts + idx[p]) * V;
#pragma unroll
        for (int v = 0; v < V; ++v) o[v] = (double)s[p][v] - (double)c[p][v];
    }
}

// out[i * V + v] += sum over splits (in split order) of partial[split][i][v]
__global__ void vpl_reduce_kernel(const double *__restrict__ partial, uint32_t n_pts, uint32_t V, uint32_t splits,
                                  double *__restrict__ out) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const size_t n = (size_t)n_pts * V;
    if (t >= n) return;
    double acc = 0.0;
    for (uint32_t k = 0; k < splits; ++k) acc += partial[(size_t)k * n + t];
    out[t] += acc;
}

__global__ void f64_to_f32_kernel(const double *__restrict__ in, uint32_t n, float *__restrict__ out) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = (float)in[i];
}

}  // namespace wobv

namespace wobl {

uint32_t vpl_record_floats(int dim, int bie) { return bie == wobdev::kSL ? 4u : 8u; (void)dim; }

void vpl_plan(uint32_t n_pts, uint32_t n_vpl, uint32_t &splits, uint32_t &chunk) {
    const uint32_t tiles = (n_pts + wobv::kBlock * wobv::kPts - 1) / (wobv::kBlock * wobv::kPts);
    // about 2048 workgroups (8 per CU), no split below 256 VPLs, at most 64 splits
    uint32_t s = (2048u + tiles - 1) / tiles;
    const uint32_t max_s = n_vpl / 256u > 0 ? n_vpl / 256u : 1u;
    s = s < max_s ? s : max_s;
    s = s < 64u ? s : 64u;
    s = s > 0 ? s : 1u;
    chunk = (n_vpl + s - 1) / s;
    splits = (n_vpl + chunk - 1) / chunk;
}

hipError_t launch_vpl_stage(int dim, int bie, const float *rec_in, const float *contrib, uint32_t n, float inv_pdf,
                            float dsign, float *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
#define WOB_VS(D, M) hipLaunchKernelGGL((wobv::vpl_stage_kernel<D, M>), dim3(grid), dim3(block), 0, st, rec_in, contrib, n, \
                                        inv_pdf, dsign, out)
    if (dim == 2) {
        if (bie == wobdev::kSL) WOB_VS(2, wobdev::kSL); else if (bie == wobdev::kDL) WOB_VS(2, wobdev::kDL); else WOB_VS(2, wobdev::kDirect);
    } else {
        if (bie == wobdev::kSL) WOB_VS(3, wobdev::kSL); else if (bie == wobdev::kDL) WOB_VS(3, wobdev::kDL); else WOB_VS(3, wobdev::kDirect);
    }
#undef WOB_VS
    return hipGetLastError();
}

hipError_t launch_vpl_aggregate(int dim, int bie, const float *pts, uint32_t n_pts, const float *rec, uint32_t n_vpl,
                                uint32_t splits, uint32_t chunk, float eps, double *partial, hipStream_t st) {
    const uint32_t tiles = (n_pts + wobv::kBlock * wobv::kPts - 1) / (wobv::kBlock * wobv::kPts);
    const dim3 grid(tiles, splits);
#define WOB_VA(D, M) hipLaunchKernelGGL((wobv::vpl_aggregate_kernel<D, M>), grid, dim3(wobv::kBlock), 0, st, pts, n_pts, rec, \
                                        n_vpl, chunk, eps, partial)
    if (dim == 2) {
        if (bie == wobdev::kSL) WOB_VA(2, wobdev::kSL); else if (bie == wobdev::kDL) WOB_VA(2, wobdev::kDL); else WOB_VA(2, wobdev::kDirect);
    } else {
        if (bie == wobdev::kSL) WOB_VA(3, wobdev::kSL); else if (bie == wobdev::kDL) WOB_VA(3, wobdev::kDL); else WOB_VA(3, wobdev::kDirect);
    }
#undef WOB_VA
    return hipGetLastError();
}

hipError_t launch_vpl_reduce(const double *partial, uint32_t n_pts, uint32_t V, uint32_t splits, double *out,
                             hipStream_t st) {
    const size_t n = (size_t)n_pts * V;
    const uint32_t block = 256, grid = (uint32_t)((n + block - 1) / block);
    hipLaunchKernelGGL(wobv::vpl_reduce_kernel, dim3(grid), dim3(block), 0, st, partial, n_pts, V, splits, out);
    return hipGetLastError();
}

hipError_t launch_f64_to_f32(const double *in, uint32_t n, float *out, hipStream_t st) {
    const uint32_t block = 256, grid = (n + block - 1) / block;
    hipLaunchKernelGGL(wobv::f64_to_f32_kernel, dim3(grid), dim3(block), 0, st, in, n, out);
    return hipGetLastError();
}

}  // namespace wobl
