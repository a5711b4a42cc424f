// Device code of the MI355X walk-on-boundary engine (gfx950, wave64).
//
// Mirrors, in FP32, the reference's hot path:
//   equation_kernel.cuh:67-145   Laplace kernels            -> Kern<D>
//   bie_models.cuh:53-588        BIE weights                -> Model<D, BIE>
//   scene.cuh:441-622, 734-918   boundary samplers          -> sample_uniform / sample_bv / line_sample / ris
//   scene.cuh:414-436, 707-730   line-element intersection  -> prim_test
//   lbvh query_line_intersect    all-hit traversal          -> line_query (4-wide BVH, LDS stack)
//   estimator.cuh:113-504        backward / forward walk    -> walk_backward, domain_path, boundary_path
// Randomness is a Philox4x32-10 stream per path (key = seed, counter = (block, path, point,
// stream)); the hit taken by a line sample is the one with the smallest
// triple32(pick ^ id * 0x9E3779B9), which is uniform over the hits and independent of the
// traversal order (the reference takes buf[floor(N u)] in lbvh order, scene.cuh:500-502).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define WOB_HD __host__ __device__
#include "wob_trig.hpp"

// FP32 expressions are evaluated as written (no fused multiply-add contraction), the
// operation order of the reference's formulas; transcendental calls aside, this keeps the
// kernel and weight values bit-identical to a non-contracted CPU evaluation.
#pragma clang fp contract(off)

namespace wobdev {

constexpr float kFltEps = 1.1920928955078125e-07f;
constexpr float kPi = 3.14159274101257324f;   // (float)M_PI

enum { kDirichlet = 0, kNeumann = 1, kRobin = 2 };
enum { kSL = 0, kDL = 1, kDirect = 2 };
enum { kBackward = 0, kForward = 1 };
// sampling flavour selected from (model, problem, estimator) exactly as estimator.cuh does
enum { FL_LINE_LINE = 0,   // backward, double layer, Dirichlet problem   (estimator.cuh:158-161, 255-263)
       FL_LINE_BV = 1,     // backward, direct, Neumann problem           (estimator.cuh:162-164, 264-275)
       FL_UNIFORM = 2,     // backward, everything else (SL uses RIS when nrc > 1: :276-347)
       FL_FWD_LINE = 3,    // forward, single layer, Neumann problem      (estimator.cuh:423-424, 445-448)
       FL_FWD_UNIFORM = 4  // forward, everything else
};

// Volume source of a Poisson problem: the reference's grid cache (scene.cuh:78-120,
// value-proportional cell sampling) or a Gaussian mixture f(x) = sum_j A_j exp(-|x - c_j|^2 /
// (2 w_j^2)) sampled from the mixture with weights proportional to |A_j| w_j^D.
enum { kVolNone = 0, kVolGrid = 1, kVolGauss = 2 };
struct VolSrc {
    uint32_t kind;
    uint32_t n;              // grid: cells; mixture: components
    uint32_t grid_res;
    uint32_t mask;           // mixture: samples on the far side of the boundary contribute 0
    float domain_size;       // grid covers (-domain_size / 2, domain_size / 2)^D
    const float *cache;      // grid values
    const float *cdf;        // grid cell CDF (n) or mixture CDF (n)
    const uint32_t *guide;   // grid guide table
    const float4 *comp;      // mixture: [cx cy cz w] [A pn inv2w2 0] per component
    // mixture with mask: the inside test's occupancy grid, gres^D cells over the vertex box
    // (cell (x, y, z) at x + gres (y + gres z)): 0 outside, 1 inside, 2 near the boundary
    const uint8_t *gflag;
    uint32_t gres;
    float glo[3], gh[3], ginv[3];
};

struct DevScene {
    const float4 *nodes;      // quantized nodes (line traversal)
    const float4 *fnodes;     // float nodes (closest-point queries)
    const float4 *prims;
    const uint32_t *prim_id;
    const float4 *attr;
    const float4 *geo;        // 3D: vertices and normal, 4 rows per element (64-byte records)
    const float4 *area_fused; // 3D: fused area-guide records, 4 rows each (guide row + geometry)
    uint32_t n_fused;
    const float *area_cdf;
    const float4 *area_guide;   // guide records (i, cdf[i-1], cdf[i], cdf[i+1]) per entry
    const float *bv_cdf;
    const float4 *bv_guide;
    uint32_t n;
    float total_area;
    float total_bv;
    uint32_t cv_R;            // number of value functions in cv_table (0: none)
    const float *cv_table;    // control variate values per (r - 1, element, vertex): [cv_R][n][D]
    VolSrc vol;
    // harmonic control variate of double-layer Dirichlet walks (0 sources: none):
    // u_h(p) = hcv_c0 + sum_j w_j phi(|p - s_j|^2), phi(z) = z^(-1/2) (3D) or log2 z (2D);
    // hcv_src[j] = (s_j, w_j) in 3D, (s_j, w_j, 0) in 2D
    uint32_t hcv_m;
    float hcv_c0;
    const float4 *hcv_src;
};

// Scene pointers reach a kernel either in its argument or through an entry table in global
// memory; where both are possible the compiler cannot tell and addresses them with FLAT
// instructions (which also wait on LDS traffic).  They all point to device global memory, and
// the hot loads below say so (address space 1), so they compile to GLOBAL loads.
typedef float gf4_t __attribute__((ext_vector_type(4)));
__device__ __forceinline__ float4 gld4(const float4 *p) {
    const gf4_t v = *(const __attribute__((address_space(1))) gf4_t *)p;
    return make_float4(v.x, v.y, v.z, v.w);
}
__device__ __forceinline__ float gldf(const float *p) { return *(const __attribute__((address_space(1))) float *)p; }
__device__ __forceinline__ uint32_t gldu(const uint32_t *p) {
    return *(const __attribute__((address_space(1))) uint32_t *)p;
}

// ----------------------------------------------------------------------------------------
// RNG
// ----------------------------------------------------------------------------------------
__device__ __forceinline__ void philox4x32_10(uint32_t &c0, uint32_t &c1, uint32_t &c2, uint32_t &c3, uint32_t k0,
                                              uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const uint32_t lo0 = 0xD2511F53u * c0, hi0 = __umulhi(0xD2511F53u, c0);
        const uint32_t lo1 = 0xCD9E8D57u * c2, hi1 = __umulhi(0xCD9E8D57u, c2);
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ __forceinline__ uint32_t triple32(uint32_t x) {
    x ^= x >> 17; x *= 0xED5AD4BBu;
    x ^= x >> 11; x *= 0xAC4C1B51u;
    x ^= x >> 15; x *= 0x31848BABu;
    x ^= x >> 14;
    return x;
}

struct Rng {
    uint32_t k0, k1, block, path, point, stream;
    uint32_t b0, b1, b2, b3;
    int left;
    __device__ __forceinline__ void init(uint64_t seed, uint32_t path_, uint32_t point_, uint32_t stream_) {
        k0 = (uint32_t)seed; k1 = (uint32_t)(seed >> 32);
        block = 0; path = path_; point = point_; stream = stream_;
        left = 0;
    }
    __device__ __forceinline__ uint32_t next_u32() {
        if (left == 0) {
            uint32_t c0 = block, c1 = path, c2 = point, c3 = stream;
            philox4x32_10(c0, c1, c2, c3, k0, k1);
            b0 = c0; b1 = c1; b2 = c2; b3 = c3;
            block += 1;
            left = 4;
        }
        const uint32_t r = b0;
        b0 = b1; b1 = b2; b2 = b3;
        left -= 1;
        return r;
    }
    // (0, 1] with 24 bits, like curand_uniform's range (utils.cuh:35-37)
    __device__ __forceinline__ float uniform() { return (float)((next_u32() >> 8) + 1u) * (1.0f / 16777216.0f); }
};

// ----------------------------------------------------------------------------------------
// small vector helpers (Eigen evaluation order: left-to-right dot, v / norm normalisation)
// ----------------------------------------------------------------------------------------
template <int D> __device__ __forceinline__ float dotv(const float *a, const float *b) {
    float s = a[0] * b[0];
#pragma unroll
    for (int i = 1; i < D; ++i) s = s + a[i] * b[i];
    return s;
}

template <int D> struct BP {
    float p[D], n[D];
    float bv;
    float alpha;
    int type;
    uint32_t elem;   // element the point lies on (0xFFFFFFFF: none / unknown)
    float bw[D];     // weights of the element's vertices at p (control-variate lookups)
};

template <int D> __device__ __forceinline__ BP<D> bp_zero() {
    BP<D> b;
#pragma unroll
    for (int k = 0; k < D; ++k) { b.p[k] = 0.f; b.n[k] = 0.f; }
    b.bv = 0.f; b.alpha = 0.f; b.type = kDirichlet; b.elem = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < D; ++k) b.bw[k] = 0.f;
    return b;
}

// ----------------------------------------------------------------------------------------
// Laplace kernels (equation_kernel.cuh:67-145).  r = y - x; rhat = r / |r| when |r| > 0.
// ----------------------------------------------------------------------------------------
template <int D> struct Kern;

template <> struct Kern<2> {
    float eps;
    // The logarithm is correctly rounded (wobtrig::log_cr), so G stays within 2 FP32 ulps of
    // any FP32 log that is within one ulp of the exact value, as the host C library's is.
    __device__ __forceinline__ float G(const float *y, const float *x) const {
        const float r0 = y[0] - x[0], r1 = y[1] - x[1];
        return -1.0f / (2.0f * kPi) * wobtrig::log_cr(fmaxf(sqrtf(r0 * r0 + r1 * r1), eps));
    }
    __device__ __forceinline__ void dG_dx(const float *y, const float *x, float *out) const {
        const float r[2] = {y[0] - x[0], y[1] - x[1]};
        const float z = dotv<2>(r, r), nr = sqrtf(z);
        const float m = fmaxf(nr, eps);
        const float c = 1.0f / (2.0f * kPi);
#pragma unroll
        for (int k = 0; k < 2; ++k) out[k] = c * (z > 0.f ? r[k] / nr : r[k]) / m;
    }
    __device__ __forceinline__ float dG_dny(const float *y, const float *x, const float *ny) const {
        const float r[2] = {y[0] - x[0], y[1] - x[1]};
        const float z = dotv<2>(r, r), nr = sqrtf(z);
        float rh[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        return -1.0f / (2.0f * kPi) * dotv<2>(rh, ny) / fmaxf(nr, eps);
    }
    __device__ __forceinline__ void d2G_dxdny(const float *y, const float *x, const float *ny, float *out) const {
        const float r[2] = {y[0] - x[0], y[1] - x[1]};
        const float z = dotv<2>(r, r), nr = sqrtf(z);
        float rh[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        const float s = 1.0f / (2.0f * kPi * fmaxf(z, eps));
        const float t = 2.0f * dotv<2>(rh, ny);
#pragma unroll
        for (int k = 0; k < 2; ++k) out[k] = s * (ny[k] - t * rh[k]);
    }
    __device__ __forceinline__ float d2G_dnxdny(const float *y, const float *x, const float *ny, const float *nx) const {
        const float r[2] = {y[0] - x[0], y[1] - x[1]};
        const float z = dotv<2>(r, r), nr = sqrtf(z);
        float rh[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        return 1.0f / (2.0f * kPi * fmaxf(z, eps)) * (dotv<2>(nx, ny) - 2.0f * dotv<2>(rh, nx) * dotv<2>(rh, ny));
    }
};

template <> struct Kern<3> {
    float eps;
    __device__ __forceinline__ float G(const float *y, const float *x) const {
        const float r0 = y[0] - x[0], r1 = y[1] - x[1], r2 = y[2] - x[2];
        return 1.0f / (4.0f * kPi * fmaxf(sqrtf(r0 * r0 + r1 * r1 + r2 * r2), eps));
    }
    __device__ __forceinline__ void dG_dx(const float *y, const float *x, float *out) const {
        const float r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        const float z = dotv<3>(r, r), nr = sqrtf(z);
        const float m = fmaxf(z, eps);
        const float c = 1.0f / (4.0f * kPi);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = c * (z > 0.f ? r[k] / nr : r[k]) / m;
    }
    __device__ __forceinline__ float dG_dny(const float *y, const float *x, const float *ny) const {
        const float r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        const float z = dotv<3>(r, r), nr = sqrtf(z);
        float rh[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        return -1.0f / (4.0f * kPi) * dotv<3>(rh, ny) / fmaxf(z, eps);
    }
    __device__ __forceinline__ void d2G_dxdny(const float *y, const float *x, const float *ny, float *out) const {
        const float r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        const float z = dotv<3>(r, r), nr = sqrtf(z);
        float rh[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        const float s = 1.0f / (4.0f * kPi * fmaxf(nr * z, eps));
        const float t = 3.0f * dotv<3>(rh, ny);
#pragma unroll
        for (int k = 0; k < 3; ++k) out[k] = s * (ny[k] - t * rh[k]);
    }
    __device__ __forceinline__ float d2G_dnxdny(const float *y, const float *x, const float *ny, const float *nx) const {
        const float r[3] = {y[0] - x[0], y[1] - x[1], y[2] - x[2]};
        const float z = dotv<3>(r, r), nr = sqrtf(z);
        float rh[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) rh[k] = z > 0.f ? r[k] / nr : r[k];
        return 1.0f / (4.0f * kPi * fmaxf(nr * z, eps)) * (dotv<3>(nx, ny) - 3.0f * dotv<3>(rh, nx) * dotv<3>(rh, ny));
    }
};

// ----------------------------------------------------------------------------------------
// element attributes and samplers
// ----------------------------------------------------------------------------------------
// Inverse CDF with a guide table: the first i with u <= cdf[i], clamped to n - 1, which is
// what binary_search (scene.cuh:30-43) returns.
__device__ __forceinline__ uint32_t cdf_lookup(const float *__restrict__ cdf, const uint32_t *__restrict__ guide,
                                               uint32_t n, float u) {
    uint32_t k = (uint32_t)(u * (float)n);
    k = k < n ? k : n - 1;
    uint32_t i = gldu(guide + k);
    while (i + 1 < n && u > gldf(cdf + i)) ++i;
    while (i > 0 && u <= gldf(cdf + i - 1)) --i;
    return i;
}

// The same search from a guide record g = (i, cdf[i-1], cdf[i], cdf[i+1]) of entry
// k = floor(u n): the search of cdf_lookup with the first forward and backward comparisons
// taken from the record.  It ends within one step of i almost always and then reads no CDF
// entry.  (After a forward step the backward condition u <= cdf[i-1] is false.)
__device__ __forceinline__ uint32_t cdf_search(const float *__restrict__ cdf, uint32_t n, float u, float4 g) {
    uint32_t i = __float_as_uint(g.x);
    if (i + 1 < n && u > g.z) {
        ++i;
        if (i + 1 < n && u > g.w) {
            ++i;
            while (i + 1 < n && u > gldf(cdf + i)) ++i;
        }
    } else if (i > 0 && u <= g.y) {
        --i;
        while (i > 0 && u <= gldf(cdf + i - 1)) --i;
    }
    return i;
}
__device__ __forceinline__ float4 guide_record(const float4 *__restrict__ guide, uint32_t n, float u) {
    uint32_t k = (uint32_t)(u * (float)n);
    k = k < n ? k : n - 1;
    return gld4(guide + k);
}
__device__ __forceinline__ uint32_t cdf_lookup(const float *__restrict__ cdf, const float4 *__restrict__ guide,
                                               uint32_t n, float u) {
    return cdf_search(cdf, n, u, guide_record(guide, n, u));
}

// two independent lookups with their guide records loaded together
__device__ __forceinline__ void cdf_lookup2(const float *__restrict__ cdf, const float4 *__restrict__ guide, uint32_t n,
                                            float ua, float ub, uint32_t &ia, uint32_t &ib) {
    const float4 ga = guide_record(guide, n, ua), gb = guide_record(guide, n, ub);
    ia = cdf_search(cdf, n, ua, ga);
    ib = cdf_search(cdf, n, ub, gb);
}

// point on element e at parameter(s) from the sampler; fills everything but p
template <int D> __device__ __forceinline__ void load_attr(const DevScene &S, uint32_t e, float *v /* D*D */, BP<D> &b,
                                                          float *bvv, float *alv, float &area) {
    b.elem = e;
    if constexpr (D == 2) {
        const float4 f0 = gld4(S.attr + 3 * e), f1 = gld4(S.attr + 3 * e + 1), f2 = gld4(S.attr + 3 * e + 2);
        v[0] = f0.x; v[1] = f0.y; v[2] = f0.z; v[3] = f0.w;
        b.n[0] = f1.x; b.n[1] = f1.y;
        bvv[0] = f1.z; bvv[1] = f1.w;
        alv[0] = f2.x; alv[1] = f2.y;
        b.type = (int)f2.z;
        area = f2.w;
    } else {
        const float4 f0 = gld4(S.attr + 5 * e), f1 = gld4(S.attr + 5 * e + 1), f2 = gld4(S.attr + 5 * e + 2),
                     f3 = gld4(S.attr + 5 * e + 3), f4 = gld4(S.attr + 5 * e + 4);
        v[0] = f0.x; v[1] = f0.y; v[2] = f0.z; v[3] = f0.w; v[4] = f1.x; v[5] = f1.y; v[6] = f1.z; v[7] = f1.w; v[8] = f2.x;
        b.n[0] = f2.y; b.n[1] = f2.z; b.n[2] = f2.w;
        bvv[0] = f3.x; bvv[1] = f3.y; bvv[2] = f3.z;
        b.type = (int)f3.w;
        alv[0] = f4.x; alv[1] = f4.y; alv[2] = f4.z;
        area = f4.w;
    }
}

template <int D> struct Sample {
    BP<D> y;
    float inv_pdf;
};

template <int D> __device__ __forceinline__ void uniform_in_element(const float *v, const float *bvv, const float *alv,
                                                                   BP<D> &b, Rng &rng) {
    if constexpr (D == 2) {
        const float u2 = rng.uniform();
#pragma unroll
        for (int k = 0; k < 2; ++k) b.p[k] = v[k] + u2 * (v[2 + k] - v[k]);          // interpolate (scene.cuh:394-396)
        b.bv = bvv[0] + u2 * (bvv[1] - bvv[0]);
        b.alpha = alv[0] + u2 * (alv[1] - alv[0]);
        b.bw[0] = 1.0f - u2; b.bw[1] = u2;
    } else {
        const float u2 = rng.uniform();
        const float u3 = rng.uniform();
        const float s = sqrtf(u2);                                                     // scene.cuh:667-671
        const float w0 = 1.0f - s, w1 = s * (1.0f - u3), w2 = u3 * s;
#pragma unroll
        for (int k = 0; k < 3; ++k) b.p[k] = w0 * v[k] + w1 * v[3 + k] + w2 * v[6 + k];
        b.bv = w0 * bvv[0] + w1 * bvv[1] + w2 * bvv[2];
        b.alpha = w0 * alv[0] + w1 * alv[1] + w2 * alv[2];
        b.bw[0] = w0; b.bw[1] = w1; b.bw[2] = w2;
    }
}

// scene.cuh:441-462 / 734-758
template <int D> __device__ __forceinline__ Sample<D> sample_uniform(const DevScene &S, Rng &rng) {
    const float u1 = rng.uniform();
    const uint32_t e = cdf_lookup(S.area_cdf, S.area_guide, S.n, u1);
    float v[D * D], bvv[D], alv[D], area;
    Sample<D> r;
    load_attr<D>(S, e, v, r.y, bvv, alv, area);
    uniform_in_element<D>(v, bvv, alv, r.y, rng);
    r.inv_pdf = S.total_area;
    return r;
}

// The geometric part of a uniform boundary sample (element, vertex weights, point, normal):
// the same draws and arithmetic as sample_uniform, reading only the element's vertices and
// normal (2 of 3 attribute rows in 2D, 3 of 5 in 3D).  complete_uniform adds the boundary
// value, Robin coefficient and type, so complete_uniform(S, uniform_geometry(S, rng)) ==
// sample_uniform(S, rng).y bit for bit.
template <int D> struct GeomSample {
    uint32_t elem;
    float bw[D];
    float p[D], n[D];
};

// The geometry of a uniform sample of element e from the sample's remaining uniforms: u[1]
// (and u[2] in 3D) if `u` is given, otherwise drawn from rng after the element's attribute
// loads are issued (the draws then overlap the loads).
// In 2D, NORMAL = false leaves n unset and skips its load (candidates whose normal is read
// only for the one RIS keeps).
template <int D, bool NORMAL = true>
__device__ __forceinline__ GeomSample<D> uniform_geometry_of(const DevScene &S, uint32_t e, const float *u, Rng *rng) {
    GeomSample<D> g;
    g.elem = e;
    if constexpr (D == 2) {
        const float4 f0 = gld4(S.attr + 3 * e);
        const float v[4] = {f0.x, f0.y, f0.z, f0.w};
        if constexpr (NORMAL) {
            const float4 f1 = gld4(S.attr + 3 * e + 1);
            g.n[0] = f1.x; g.n[1] = f1.y;
        } else {
            g.n[0] = 0.f; g.n[1] = 0.f;
        }
        const float u2 = u ? u[1] : rng->uniform();
#pragma unroll
        for (int k = 0; k < 2; ++k) g.p[k] = v[k] + u2 * (v[2 + k] - v[k]);            // as uniform_in_element
        g.bw[0] = 1.0f - u2; g.bw[1] = u2;
    } else {
        const float4 f0 = gld4(S.geo + 4 * e), f1 = gld4(S.geo + 4 * e + 1), f2 = gld4(S.geo + 4 * e + 2);
        const float v[9] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x};
        g.n[0] = f2.y; g.n[1] = f2.z; g.n[2] = f2.w;
        const float u2 = u ? u[1] : rng->uniform();
        const float u3 = u ? u[2] : rng->uniform();
        const float sq = sqrtf(u2);
        const float w0 = 1.0f - sq, w1 = sq * (1.0f - u3), w2 = u3 * sq;
#pragma unroll
        for (int k = 0; k < 3; ++k) g.p[k] = w0 * v[k] + w1 * v[3 + k] + w2 * v[6 + k];
        g.bw[0] = w0; g.bw[1] = w1; g.bw[2] = w2;
    }
    return g;
}

// 3D: the element from the fused record of entry floor(u1 n_fused), whose geometry rows are
// those of element i or i + 1 (bit 31 of the index word), the one covering more of the entry;
// the search usually ends there, otherwise the element's own geometry record is read.  The
// same draws and arithmetic as uniform_geometry_of.
template <int D> __device__ __forceinline__ GeomSample<D> uniform_geometry(const DevScene &S, Rng &rng) {
    const float u1 = rng.uniform();
    if constexpr (D == 2) {
        const uint32_t e = cdf_lookup(S.area_cdf, S.area_guide, S.n, u1);
        return uniform_geometry_of<D>(S, e, nullptr, &rng);
    } else {
        uint32_t k = (uint32_t)(u1 * (float)S.n_fused);
        k = k < S.n_fused ? k : S.n_fused - 1;
        const float4 *rec = S.area_fused + 4 * (size_t)k;
        const float4 g = gld4(rec);
        float4 f0 = gld4(rec + 1), f1 = gld4(rec + 2), f2 = gld4(rec + 3);
        const uint32_t w = __float_as_uint(g.x), i = w & 0x7FFFFFFFu;
        const uint32_t e = cdf_search(S.area_cdf, S.n, u1, make_float4(__uint_as_float(i), g.y, g.z, g.w));
        if (e != i + (w >> 31)) {
            f0 = gld4(S.geo + 4 * e); f1 = gld4(S.geo + 4 * e + 1); f2 = gld4(S.geo + 4 * e + 2);
        }
        GeomSample<D> gs;
        gs.elem = e;
        const float v[9] = {f0.x, f0.y, f0.z, f0.w, f1.x, f1.y, f1.z, f1.w, f2.x};
        gs.n[0] = f2.y; gs.n[1] = f2.z; gs.n[2] = f2.w;
        const float u2 = rng.uniform();
        const float u3 = rng.uniform();
        const float sq = sqrtf(u2);
        const float w0 = 1.0f - sq, w1 = sq * (1.0f - u3), w2 = u3 * sq;
#pragma unroll
        for (int k2 = 0; k2 < 3; ++k2) gs.p[k2] = w0 * v[k2] + w1 * v[3 + k2] + w2 * v[6 + k2];
        gs.bw[0] = w0; gs.bw[1] = w1; gs.bw[2] = w2;
        return gs;
    }
}

template <int D> __device__ __forceinline__ BP<D> complete_uniform(const DevScene &S, const GeomSample<D> &g) {
    BP<D> b;
    b.elem = g.elem;
#pragma unroll
    for (int k = 0; k < D; ++k) { b.p[k] = g.p[k]; b.n[k] = g.n[k]; b.bw[k] = g.bw[k]; }
    if constexpr (D == 2) {
        const float4 f1 = gld4(S.attr + 3 * g.elem + 1), f2 = gld4(S.attr + 3 * g.elem + 2);
        const float u2 = g.bw[1];
        b.bv = f1.z + u2 * (f1.w - f1.z);
        b.alpha = f2.x + u2 * (f2.y - f2.x);
        b.type = (int)f2.z;
    } else {
        const float4 f3 = gld4(S.attr + 5 * g.elem + 3), f4 = gld4(S.attr + 5 * g.elem + 4);
        b.bv = g.bw[0] * f3.x + g.bw[1] * f3.y + g.bw[2] * f3.z;
        b.alpha = g.bw[0] * f4.x + g.bw[1] * f4.y + g.bw[2] * f4.z;
        b.type = (int)f3.w;
    }
    return b;
}

// scene.cuh:586-622 / 891-918
template <int D> __device__ __forceinline__ Sample<D> sample_bv(const DevScene &S, Rng &rng) {
    const float u1 = rng.uniform();
    const uint32_t e = cdf_lookup(S.bv_cdf, S.bv_guide, S.n, u1);
    float v[D * D], bvv[D], alv[D], area;
    Sample<D> r;
    load_attr<D>(S, e, v, r.y, bvv, alv, area);
    if constexpr (D == 2) {
        const float u2 = rng.uniform();
        const float ba = fabsf(bvv[0]), bb = fabsf(bvv[1]);
        float interp;
        if (fabsf(ba - bb) < 10 * kFltEps) interp = u2;
        else interp = (-ba + sqrtf(ba * ba + u2 * (bb * bb - ba * ba))) / (bb - ba);
#pragma unroll
        for (int k = 0; k < 2; ++k) r.y.p[k] = v[k] + interp * (v[2 + k] - v[k]);
        r.y.bv = bvv[0] + interp * (bvv[1] - bvv[0]);
        r.y.alpha = alv[0] + interp * (alv[1] - alv[0]);
        r.y.bw[0] = 1.0f - interp; r.y.bw[1] = interp;
        r.inv_pdf = S.total_bv / (ba + interp * (bb - ba));
    } else {
        const float pmf = e == 0 ? gldf(S.bv_cdf + e) : gldf(S.bv_cdf + e) - gldf(S.bv_cdf + e - 1);
        uniform_in_element<D>(v, bvv, alv, r.y, rng);
        r.inv_pdf = area / pmf;
    }
    return r;
}

// ----------------------------------------------------------------------------------------
// 4-wide BVH all-hit line query
// ----------------------------------------------------------------------------------------
template <int D> struct LineHit {
    uint32_t count;
    uint32_t id;      // element id of the picked hit
    float t, a, b;    // line parameter and element parameters (2D: a; 3D: barycentrics of b, c)
};

// Running uniform pick over a query's hits: count them and keep the hit with the smallest
// triple32(pick ^ id * 0x9E3779B9) (ties: smaller id), which is uniform over the hits for a
// uniform pick and independent of the traversal order.  Selects only, no lane branches.
template <int D>
__device__ __forceinline__ void hit_update(bool hit, uint32_t id, uint32_t skip, uint32_t pick, float t, float a,
                                           float b, LineHit<D> &h, uint32_t &best) {
    const bool ok = hit && id != skip;
    const uint32_t key = triple32(pick ^ (id * 0x9E3779B9u));
    const bool take = ok && (key < best || (key == best && id < h.id));
    h.count += ok ? 1u : 0u;
    best = take ? key : best;
    h.id = take ? id : h.id;
    h.t = take ? t : h.t;
    h.a = take ? a : h.a;
    h.b = take ? b : h.b;
}

// Geometry record of primitive `prim` (leaf order): 2D one float4, 3D three.
template <int D> struct PrimRec {
    float4 g[D == 2 ? 1 : 3];
};
template <int D> __device__ __forceinline__ PrimRec<D> prim_load(const DevScene &S, uint32_t prim) {
    PrimRec<D> r;
    if constexpr (D == 2) {
        r.g[0] = gld4(S.prims + prim);
    } else {
        r.g[0] = gld4(S.prims + 3 * prim);
        r.g[1] = gld4(S.prims + 3 * prim + 1);
        r.g[2] = gld4(S.prims + 3 * prim + 2);
    }
    return r;
}

// line o + t d against a primitive record, solved as the reference solves it
// (scene.cuh:414-436, 707-730): the matrix [d, a - b, a - c] (2D: [d, a - b]) is inverted as
// adjugate / det, with 1 / det an IEEE (correctly rounded) division, and the inverse is
// applied to a - o.  Every product is rounded on its own (contraction is off in this file)
// and sums of three terms run left to right, so t, a and b are the same FP32 values a
// non-contracted CPU evaluation of that algebra gives, and the hit predicate -- parameters
// in the closed element, |t| > 10 FLT_EPSILON, no NaN (scene.cuh:431-432, 725-727) -- decides
// exactly as it does there, grazing and vertex cases included.
template <int D>
__device__ __forceinline__ bool prim_eval(const PrimRec<D> &r, const float *o, const float *d, float &t, float &a,
                                          float &b) {
    constexpr float eps = 10 * kFltEps;
    if constexpr (D == 2) {
        const float4 g = r.g[0];                            // a, a - b
        const float m00 = d[0], m10 = d[1], m01 = g.z, m11 = g.w;
        const float det = m00 * m11 - m10 * m01;
        const float invdet = 1.0f / det;
        const float i00 = m11 * invdet, i10 = -m10 * invdet, i01 = -m01 * invdet, i11 = m00 * invdet;
        const float r0 = g.x - o[0], r1 = g.y - o[1];
        t = i00 * r0 + i01 * r1;
        a = i10 * r0 + i11 * r1;
        b = 0.f;
        return a >= 0.0f && a <= 1.0f && fabsf(t) > eps;   // ordered comparisons reject NaN
    } else {
        const float4 g0 = r.g[0], g1 = r.g[1], g2 = r.g[2];   // a, a - b, a - c
        // m = [d, a - b, a - c] by rows
        const float m[3][3] = {{d[0], g0.w, g1.z}, {d[1], g1.x, g1.w}, {d[2], g1.y, g2.x}};
        // cofactor (i, j) with cyclic indices, as a fixed-size 3x3 inverse forms it
        auto cof = [&](int i, int j) {
            const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
            return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
        };
        const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
        const float det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
        const float invdet = 1.0f / det;
        const float r0 = g0.x - o[0], r1 = g0.y - o[1], r2 = g0.z - o[2];
        float tab[3];
#pragma unroll
        for (int i = 0; i < 3; ++i) {
            // row i of the inverse is column i of the adjugate: cof(j, i) / det
            const float i0 = (i == 0 ? c00 : cof(0, i)) * invdet;
            const float i1 = (i == 0 ? c10 : cof(1, i)) * invdet;
            const float i2 = (i == 0 ? c20 : cof(2, i)) * invdet;
            tab[i] = i0 * r0 + i1 * r1 + i2 * r2;
        }
        t = tab[0]; a = tab[1]; b = tab[2];
        // the reference's !(a < 0 || b < 0 || a + b > 1 || |t| <= eps || isnan(t, a, b)): every
        // comparison with a NaN is false, so the ordered form below rejects NaN too
        return a >= 0.0f && b >= 0.0f && a + b <= 1.0f && fabsf(t) > eps;
    }
}

// line o + t d against primitive `prim` (leaf order)
template <int D> __device__ __forceinline__ bool prim_test(const DevScene &S, uint32_t prim, const float *o,
                                                           const float *d, float &t, float &a, float &b) {
    return prim_eval<D>(prim_load<D>(S, prim), o, d, t, a, b);
}

// Primitive boxes are inflated at build time by 4e-6 x |coordinate| (scene_build.cpp) and the
// quantized child boxes contain them; that is about 10 times the rounding of the slab
// parameters below (three roundings); a line meeting an element therefore always meets every
// box on the element's path to the root, and the plain comparison tmin <= tmax decides.

// Traversal state of one lane's query: the line's slab parameters, the current entry
// (internal node >= 0, leaf < 0, or kTravEmpty), the stack's top and bottom as element
// offsets (depth * stride; entries [sb, spo), the bottom moves only when another lane takes
// the bottom entry) and the range of pending primitives [lp, le).
constexpr int kTravEmpty = (int)0x80000000;
template <int D> struct Trav {
    float inv[D], oi[D];
    int cur;
    uint32_t spo, sb;
    uint32_t lp, le;

    __device__ __forceinline__ void begin(const float *o, const float *d) {
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float dk = fabsf(d[k]) < 1e-30f ? copysignf(1e-30f, d[k]) : d[k];   // finite slab parameters
            inv[k] = 1.0f / dk;
            oi[k] = o[k] * inv[k];
        }
        cur = 0; spo = 0; sb = 0; lp = 0; le = 0;
    }
    // a leaf entry becomes the pending primitive range and the next entry is popped
    __device__ __forceinline__ void settle(const uint32_t *stack, uint32_t stride) {
        const bool leaf = lp >= le && cur < 0 && cur != kTravEmpty;
        const uint32_t v = ~(uint32_t)cur;
        const bool pop = leaf && spo != sb;
        const int top = (int)stack[spo != sb ? spo - stride : 0u];
        lp = leaf ? v >> 4 : lp;
        le = leaf ? (v >> 4) + (v & 15u) + 1u : le;
        cur = leaf ? (pop ? top : kTravEmpty) : cur;
        spo = pop ? spo - stride : spo;
    }
    __device__ __forceinline__ bool want_prim() const { return lp < le; }
    // A lane with pending primitives may still descend: the node step only touches `cur` and
    // the stack, and a leaf it reaches waits in `cur` until the range is used up, so the
    // stack sees the same depth-first order as without the overlap.
    __device__ __forceinline__ bool want_node() const { return cur >= 0; }
    __device__ __forceinline__ bool done() const { return lp >= le && cur == kTravEmpty; }

    // four slab tests on the current node; the hit children are pushed and the top entry is
    // popped as the current one.  Every child is written to the slot above the top (the spare
    // slot above STACK entries takes the write past the last push), only pushes move the top,
    // so no lane branches.  A node's
    // pushes peak one entry above the depth-first bound of "continue with one hit child, push
    // the others"; the scene builds count that entry (stack_need = bound + 1).
    template <int STACK>
    __device__ __forceinline__ void node_step(const float4 *nodes, uint32_t *stack, uint32_t stride) {
        node_step<STACK>(nodes, stack, stride, inv, oi);
    }
    // the same with the slab parameters of another lane's line (a lane working on a subtree
    // taken from that lane's query); RAY: only the ray t >= 0 is wanted
    template <int STACK, bool RAY = false>
    __device__ __forceinline__ void node_step(const float4 *nodes, uint32_t *stack, uint32_t stride, const float *inv,
                                              const float *oi) {
        node_apply<STACK, RAY>(node_load(nodes), stack, stride, inv, oi);
    }
    // the current node's quantized record (kQNodeWords layout of scene_build.hpp): node box
    // corner lo and per-axis step s, 8-bit child bounds, child entries (0x80000000: none)
    struct NodeRec {
        float lo[D], s[D];
        uint32_t qlo[D], qhi[D];
        float4 chf;
    };
    __device__ __forceinline__ NodeRec node_load(const float4 *nodes) const {
        NodeRec r;
        if constexpr (D == 2) {
            const float4 *np = nodes + (size_t)cur * 3;
            const float4 a = gld4(np), b = gld4(np + 1);
            r.chf = gld4(np + 2);
            r.lo[0] = a.x; r.lo[1] = a.y; r.s[0] = a.z; r.s[1] = a.w;
            r.qlo[0] = __float_as_uint(b.x); r.qlo[1] = __float_as_uint(b.y);
            r.qhi[0] = __float_as_uint(b.z); r.qhi[1] = __float_as_uint(b.w);
        } else {
            const float4 *np = nodes + (size_t)cur * 4;
            const float4 a = gld4(np), b = gld4(np + 1), c = gld4(np + 2);
            r.chf = gld4(np + 3);
            r.lo[0] = a.x; r.lo[1] = a.y; r.lo[2] = a.z; r.s[0] = a.w;
            r.s[1] = b.x; r.s[2] = b.y; r.qlo[0] = __float_as_uint(b.z); r.qlo[1] = __float_as_uint(b.w);
            r.qlo[2] = __float_as_uint(c.x);
            r.qhi[0] = __float_as_uint(c.y); r.qhi[1] = __float_as_uint(c.z); r.qhi[2] = __float_as_uint(c.w);
        }
        return r;
    }
    // slab parameters of child bound lo + q s: q (s inv) + (lo inv - o inv).  They grow with q
    // where inv >= 0 and fall where inv < 0 (the rounded FMA is monotone in q), so the near
    // bound of every child along an axis is its lower bound for inv >= 0 and its upper bound
    // otherwise: one selection per node and axis instead of a min and a max per child.
    // RAY (a query that counts only hits with t > 0) also skips children whose slab interval
    // ends before t = 0: a counted hit has t > 10 FLT_EPSILON and lies in its (inflated) box,
    // so that box's interval reaches past 0.
    template <int STACK, bool RAY = false>
    __device__ __forceinline__ void node_apply(const NodeRec &r, uint32_t *stack, uint32_t stride, const float *inv,
                                               const float *oi) {
        const int ch[4] = {(int)__float_as_uint(r.chf.x), (int)__float_as_uint(r.chf.y), (int)__float_as_uint(r.chf.z),
                           (int)__float_as_uint(r.chf.w)};
        float A[D], B[D];
        uint32_t qn[D], qf[D];
#pragma unroll
        for (int k = 0; k < D; ++k) {
            A[k] = r.s[k] * inv[k];
            B[k] = __builtin_fmaf(r.lo[k], inv[k], -oi[k]);
            const bool up = !(inv[k] < 0.f);
            qn[k] = up ? r.qlo[k] : r.qhi[k];
            qf[k] = up ? r.qhi[k] : r.qlo[k];
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            float tmin = 0.f, tmax = 0.f;
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float n = (float)((qn[k] >> (8 * c)) & 0xFFu), f = (float)((qf[k] >> (8 * c)) & 0xFFu);
                const float t1 = __builtin_fmaf(n, A[k], B[k]), t2 = __builtin_fmaf(f, A[k], B[k]);
                tmin = k == 0 && !RAY ? t1 : fmaxf(tmin, t1);
                tmax = k == 0 ? t2 : fminf(tmax, t2);
            }
            // no overflow test: scene creation rejects BVHs whose bound (the pushes included)
            // exceeds STACK entries, so spo stays within them
            const bool push = ch[c] != kTravEmpty && tmin <= tmax;
            stack[spo] = (uint32_t)ch[c];
            spo = push ? spo + stride : spo;
        }
        // the last hit child (or the entry below, if none was hit) becomes the current entry
        const bool pop = spo != sb;
        const int top = (int)stack[pop ? spo - stride : 0u];
        spo = pop ? spo - stride : spo;
        cur = pop ? top : kTravEmpty;
    }
};

// Depth-first traversal of the 4-wide BVH for the infinite line o + t d, calling
// on_hit(hit, prim, t, a, b) for every primitive tested, hit = the reference predicate.
//
// Wave-scheduled: every lane holds one current entry (internal node or leaf) and a range of
// pending primitives; hit children are pushed on the lane's LDS stack and the top is popped
// (`stack`, entries `stride` words apart, STACK entries plus a spare; the builder bounds the
// depth), leaves included.  Each iteration the wave runs ONE kind of step, chosen by ballot:
// a primitive test for the lanes with pending primitives, or a node step (four slab tests)
// for the lanes whose current entry is an internal node (pending primitives or not) --
// whichever more lanes want.  Lanes therefore test primitives together instead of one at a
// time inside divergent per-child branches.
constexpr int kPrimsPerRound = 2;
template <int D, int STACK, typename OnHit>
__device__ __forceinline__ void traverse(const DevScene &S, const float *o, const float *d, uint32_t *stack,
                                         uint32_t stride, OnHit &&on_hit) {
    Trav<D> tr;
    tr.begin(o, d);
    while (true) {
        tr.settle(stack, stride);
        const bool want_leaf = tr.want_prim();
        const bool want_node = tr.want_node();
        const uint64_t bl = __ballot(want_leaf);
        const uint64_t bn = __ballot(want_node);
        if ((bl | bn) == 0) break;
        if (__popcll(bl) >= __popcll(bn)) {
            // up to kPrimsPerRound primitives of the lane's leaf per round: fewer rounds of the
            // settle / ballot bookkeeping per leaf
            if (want_leaf) {
#pragma unroll
                for (int u = 0; u < kPrimsPerRound; ++u) {
                    if (u > 0 && tr.lp >= tr.le) break;
                    float t, a, b;
                    const bool hit = prim_test<D>(S, tr.lp, o, d, t, a, b);
                    on_hit(hit, tr.lp, t, a, b);
                    ++tr.lp;
                }
            }
        } else if (want_node) {
            tr.template node_step<STACK>(S.nodes, stack, stride);
        }
    }
}

// Lane `src4 / 4`'s value of v (ds_bpermute; every lane of the query loop is active when it
// runs, so the source lane is too).
__device__ __forceinline__ float lane_read(float v, int src4) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src4, __float_as_int(v)));
}
__device__ __forceinline__ uint32_t lane_read(uint32_t v, int src4) {
    return (uint32_t)__builtin_amdgcn_ds_bpermute(src4, (int)v);
}

// LDS words per lane of a line-query kernel: the traversal stack (STACK entries and the spare
// slot) and the query's result slots (a hit count and a 64-bit (key, leaf) minimum).
template <int STACK> constexpr int line_query_lds_words() { return STACK + 1 + 3; }

// A steal round runs once at least this many lanes of the wave have finished their work.
constexpr int kStealMinIdle = 16;

// All hits of the line o + t d: the count and the hit with the smallest key
// triple32(pick ^ id * 0x9E3779B9), a uniform choice among the hits for a uniform pick, with
// the wave's lanes sharing the work of all its queries.  The lines of one wave need very
// different numbers of steps, and with one line per lane the wave runs until
// its longest query is done.  Here a lane whose work is done takes the bottom entry (the
// subtree nearest the root) of another lane's stack and traverses it with that lane's line,
// read from the owner lane's registers; it adds its hit count and its smallest
// (key, leaf index) to the owner's result slots in LDS with atomics.  Every primitive of the
// line's hit set is still tested exactly once, the count is a sum and the pick a minimum, so
// the result is the same whichever lane does which part.  The keys are distinct for distinct
// elements (triple32 and x -> pick ^ x * 0x9E3779B9 are bijections of 32-bit words), so the
// minimum over (key, leaf) picks the same hit as the (key, id) order of hit_update; its t and
// barycentrics are recomputed at the end by the owner with the same primitive test.
//
// LDS layout (`stack` = block base + threadIdx.x, entries `stride` words apart): the stacks,
// STACK + 1 entries per lane; then one count word per lane; then one 64-bit (key, leaf) word
// per lane (stride is a multiple of 64, so it is 8-byte aligned).  The active lanes of the
// wave at the call share the work among themselves.
//
// PICK = false counts only the hits with t > 0 (the crossings of the ray o + t d, t > 0) and
// picks nothing; pick and skip are unused.
template <int D, int STACK, bool PICK = true>
__device__ __forceinline__ LineHit<D> line_query_ws(const DevScene &S, const float *o, const float *d, uint32_t pick,
                                                    uint32_t skip, uint32_t *stack, uint32_t stride) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    uint32_t *const block_base = stack - tid;
    uint32_t *const wave_col = block_base + (tid - lane);                          // lane 0's stack column
    uint32_t *const res_cnt = block_base + (STACK + 1) * stride + (tid - lane);
    uint64_t *const res_key = reinterpret_cast<uint64_t *>(block_base + (STACK + 2) * stride) + (tid - lane);
    res_cnt[lane] = 0u;
    res_key[lane] = ~0ull;
    Trav<D> tr;
    tr.begin(o, d);
    const int self4 = (int)lane * 4;
    int own4 = self4;            // ds_bpermute address of the lane whose line this lane traverses
    uint32_t cnt = 0u;           // hits of the current work
    uint64_t best = ~0ull;       // smallest (key << 32 | leaf) of the current work
    const uint64_t active = __ballot(true);
    // Lanes share work only when they all walk the same scene: a batch launch numbers its
    // work items entry by entry, so the wave at an entry boundary can hold two scenes, and a
    // stack entry is a node index of its own scene.
    const uint64_t np_ = (uint64_t)(uintptr_t)S.nodes;
    const uint32_t np_lo = (uint32_t)np_, np_hi = (uint32_t)(np_ >> 32);
    const bool one_scene = __ballot(np_lo != (uint32_t)__builtin_amdgcn_readfirstlane((int)np_lo) ||
                                    np_hi != (uint32_t)__builtin_amdgcn_readfirstlane((int)np_hi)) == 0ull;
    while (true) {
        tr.settle(stack, stride);
        bool want_leaf = tr.want_prim();
        bool want_node = tr.want_node();
        uint64_t bl = __ballot(want_leaf);
        uint64_t bn = __ballot(want_node);
        // after settle a lane that wants neither step has no work left
        const uint64_t bi = active & ~(bl | bn);
        const uint64_t bd = one_scene && __popcll(bi) >= kStealMinIdle ? __ballot(tr.spo != tr.sb) : 0ull;
        if (bd != 0ull) {
            const bool idle = !(want_leaf || want_node);
            // idle lanes hand their results to the owners, then the i-th idle lane takes the
            // bottom stack entry of the i-th lane that has one
            if (idle && cnt != 0u) {
                atomicAdd(&res_cnt[own4 >> 2], cnt);
                if constexpr (PICK) atomicMin((unsigned long long *)&res_key[own4 >> 2], (unsigned long long)best);
                cnt = 0u;
                best = ~0ull;
            }
            // the i-th donor posts its lane in lane i's spare stack slot (never a live entry:
            // it only takes the write past the last push), the i-th idle lane reads it
            const uint32_t np = (uint32_t)min(__popcll(bi), __popcll(bd));
            const uint64_t mine = idle ? bi : bd;
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(mine >> 32),
                                                            __builtin_amdgcn_mbcnt_lo((uint32_t)mine, 0u));
            const bool given = !idle && ((bd >> lane) & 1ull) && rank < np;
            uint32_t *const mail = wave_col + (size_t)STACK * stride;
            if (given) mail[rank] = lane;
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            __builtin_amdgcn_wave_barrier();
            const int src = idle && rank < np ? (int)mail[rank] : -1;
            const int s4 = src >= 0 ? src * 4 : self4;
            const uint32_t dsb = lane_read(tr.sb, s4);
            const int down4 = __builtin_amdgcn_ds_bpermute(s4, own4);
            __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
            if (src >= 0) {
                tr.cur = (int)wave_col[(uint32_t)src + dsb];
                tr.spo = 0u;
                tr.sb = 0u;
                own4 = down4;
            }
            if (given) tr.sb += stride;
            // the thieves start on their entries in this round
            tr.settle(stack, stride);
            want_leaf = tr.want_prim();
            want_node = tr.want_node();
            bl = __ballot(want_leaf);
            bn = __ballot(want_node);
        }
        if ((bl | bn) == 0ull) break;
        if (__popcll(bl) >= __popcll(bn)) {
            float oo[D], dd[D];
#pragma unroll
            for (int k = 0; k < D; ++k) { oo[k] = lane_read(o[k], own4); dd[k] = lane_read(d[k], own4); }
            const uint32_t pk = PICK ? lane_read(pick, own4) : 0u, sk = PICK ? lane_read(skip, own4) : 0u;
            if (want_leaf) {
#pragma unroll
                for (int u = 0; u < kPrimsPerRound; ++u) {
                    if (u > 0 && tr.lp >= tr.le) break;
                    float t, a, b;
                    const PrimRec<D> pr = prim_load<D>(S, tr.lp);
                    const bool hit = prim_eval<D>(pr, oo, dd, t, a, b);
                    if constexpr (PICK) {
                        // the 3D record holds the element id
                        const uint32_t id = D == 3 ? __float_as_uint(pr.g[D == 3 ? 2 : 0].y) : gldu(S.prim_id + tr.lp);
                        const bool ok = hit && id != sk;
                        const uint64_t v = ((uint64_t)triple32(pk ^ (id * 0x9E3779B9u)) << 32) | tr.lp;
                        cnt += ok ? 1u : 0u;
                        best = ok && v < best ? v : best;
                    } else {
                        cnt += hit && t > 0.f ? 1u : 0u;
                    }
                    ++tr.lp;
                }
            }
        } else {
            float inv[D], oi[D];
#pragma unroll
            for (int k = 0; k < D; ++k) { inv[k] = lane_read(tr.inv[k], own4); oi[k] = lane_read(tr.oi[k], own4); }
            if (want_node) tr.template node_step<STACK, !PICK>(S.nodes, stack, stride, inv, oi);
        }
    }
    if (cnt != 0u) {
        atomicAdd(&res_cnt[own4 >> 2], cnt);
        if constexpr (PICK) atomicMin((unsigned long long *)&res_key[own4 >> 2], (unsigned long long)best);
    }
    __builtin_amdgcn_fence(__ATOMIC_ACQ_REL, "wavefront");
    LineHit<D> h;
    h.count = res_cnt[lane];
    h.id = 0xFFFFFFFFu; h.t = 0.f; h.a = 0.f; h.b = 0.f;
    if (PICK && h.count != 0u) {
        const uint32_t leaf = (uint32_t)res_key[lane];
        const PrimRec<D> pr = prim_load<D>(S, leaf);
        (void)prim_eval<D>(pr, o, d, h.t, h.a, h.b);
        h.id = D == 3 ? __float_as_uint(pr.g[D == 3 ? 2 : 0].y) : gldu(S.prim_id + leaf);
    }
    return h;
}

// scene.cuh:471-478 (2D), 769-782 (3D): a uniformly oriented line direction, evaluated as the
// reference's CPU build evaluates it: the FP32 angle times M_PI in double, cos and sin in
// double (wobtrig::sincos_small), each rounded once to FP32.  Over all 2^24 angles a uniform
// can give, those floats are the C library's (checked exhaustively on the host), so the
// direction is the CPU path's to the bit.
constexpr double kPiD = 3.14159265358979323846;
template <int D> __device__ __forceinline__ void line_dir(Rng &rng, float *d) {
    if constexpr (D == 2) {
        const float theta = rng.uniform();
        double s, c;
        wobtrig::sincos_small((double)theta * kPiD, s, c);
        d[0] = (float)c; d[1] = (float)s;
    } else {
        d[2] = rng.uniform();
        const float phi = 2.f * rng.uniform();
        double s, c;
        wobtrig::sincos_small((double)phi * kPiD, s, c);
        const float r = sqrtf(1.f - d[2] * d[2]);
        d[0] = (float)c * r; d[1] = (float)s * r;
    }
}

// scene.cuh:488-535 (2D), 791-840 (3D): the sample at the picked hit of the line x + t d,
// inv_pdf = N pi |r| / |rhat.n| (2D) or N 2pi |r|^2 / |rhat.n| (3D); no hit gives a zero
// point with inv_pdf 0.
template <int D>
__device__ __forceinline__ Sample<D> line_finish(const DevScene &S, const float *x, const float *d, const LineHit<D> &h) {
    Sample<D> r;
    if (h.count == 0) {                                                                 // scene.cuh:488-495
        r.y = bp_zero<D>();
        r.inv_pdf = 0.f;
        return r;
    }
    float v[D * D], bvv[D], alv[D], area;
    load_attr<D>(S, h.id, v, r.y, bvv, alv, area);
#pragma unroll
    for (int k = 0; k < D; ++k) r.y.p[k] = x[k] + h.t * d[k];
    if constexpr (D == 2) {
        const float dx = r.y.p[0] - v[0], dy = r.y.p[1] - v[1];
        const float interp = sqrtf(dx * dx + dy * dy) / area;                           // scene.cuh:505
        r.y.bv = bvv[0] + interp * (bvv[1] - bvv[0]);
        r.y.alpha = alv[0] + interp * (alv[1] - alv[0]);
        r.y.bw[0] = 1.0f - interp; r.y.bw[1] = interp;
    } else {
        // barycentric coordinates of the hit point (scene.cuh:45-69, 808-810)
        float v0[3], v1[3], v2[3];
#pragma unroll
        for (int k = 0; k < 3; ++k) { v0[k] = v[3 + k] - v[k]; v1[k] = v[6 + k] - v[k]; v2[k] = r.y.p[k] - v[k]; }
        const float d00 = dotv<3>(v0, v0), d01 = dotv<3>(v0, v1), d11 = dotv<3>(v1, v1);
        const float d20 = dotv<3>(v2, v0), d21 = dotv<3>(v2, v1);
        const float denom = d00 * d11 - d01 * d01;
        const float l1 = (d11 * d20 - d01 * d21) / denom;
        const float l2 = (d00 * d21 - d01 * d20) / denom;
        const float l0 = 1.0f - (l1 + l2);
        r.y.bv = l0 * bvv[0] + l1 * bvv[1] + l2 * bvv[2];
        r.y.alpha = l0 * alv[0] + l1 * alv[1] + l2 * alv[2];
        r.y.bw[0] = l0; r.y.bw[1] = l1; r.y.bw[2] = l2;
    }
    constexpr float eps = 10 * kFltEps;
    float rv[D];
#pragma unroll
    for (int k = 0; k < D; ++k) rv[k] = r.y.p[k] - x[k];
    const float z = dotv<D>(rv, rv), nr = sqrtf(z);
    float rh[D];
#pragma unroll
    for (int k = 0; k < D; ++k) rh[k] = z > 0.f ? rv[k] / nr : rv[k];
    const float cosv = fmaxf(fabsf(dotv<D>(rh, r.y.n)), eps);
    if constexpr (D == 2) r.inv_pdf = (float)h.count * kPi * nr / cosv;
    else r.inv_pdf = (float)h.count * 2.f * kPi * z / cosv;
    return r;
}

// scene.cuh:464-535 (2D), 760-840 (3D): a uniformly oriented line through x, all hits, one
// hit taken uniformly.
template <int D, int STACK>
__device__ __forceinline__ Sample<D> line_sample(const DevScene &S, const float *x, uint32_t skip, Rng &rng,
                                                 uint32_t *stack, uint32_t stride) {
    float d[D];
    line_dir<D>(rng, d);
    const uint32_t pick = rng.next_u32();
    const LineHit<D> h = line_query_ws<D, STACK>(S, x, d, pick, skip, stack, stride);
    return line_finish<D>(S, x, d, h);
}

// ----------------------------------------------------------------------------------------
// Volume sources
// ----------------------------------------------------------------------------------------
template <int D> struct VolSample {
    float p[D];
    float value;
    float inv_pdf;
};

// The inside test's occupancy grid at p: 0 outside, 1 inside, 2 unknown (p casts its own ray).
// An outside or inside cell lies more than a cell from every element's inflated box and the
// rays from two of its points agreed, so the parity is the same at each of its points; near
// cells, and points within a cell outside the grid, are unknown; points further out are
// outside the closed boundary.
template <int D> __device__ __forceinline__ uint32_t inside_grid_cell(const VolSrc &v, const float *p) {
    if (v.gres == 0u) return 2u;
    const float R = (float)v.gres;
    uint32_t idx = 0u, mul = 1u;
    bool out = false, edge = false;
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float sk = (p[k] - v.glo[k]) * v.ginv[k];
        if (!(sk >= -1.f && sk < R + 1.f)) out = true;
        else if (!(sk >= 0.f && sk < R)) edge = true;
        else { idx += (uint32_t)sk * mul; mul *= v.gres; }
    }
    if (out) return 0u;
    if (edge) return 2u;
    return (uint32_t)v.gflag[idx];
}

// Crossings of the ray o + t r, t > 0, with the boundary along a fixed direction r; odd
// means inside a closed boundary (a counting line query, the wave's lanes sharing the work).
// Kept out of line: it is the only traversal of a volume sample, and inlining it at every
// volume-term call site multiplies the kernels' code.
template <int D, int STACK>
__device__ __noinline__ uint32_t ray_crossings(const float4 *nodes, const float4 *prims, float o0, float o1, float o2,
                                               uint32_t *stack, uint32_t stride) {
    DevScene S{};
    S.nodes = nodes;
    S.prims = prims;
    const float o[3] = {o0, o1, o2};
    float d[D];
    if constexpr (D == 2) { d[0] = 0.58778525f; d[1] = 0.80901699f; }
    else { d[0] = 0.48218104f; d[1] = 0.66748405f; d[2] = 0.56749094f; }
    return line_query_ws<D, STACK, false>(S, o, d, 0u, 0u, stack, stride).count;
}

// scene.cuh:100-120: a cell with probability |cache| / sum |cache|, a uniform point in it,
// inv_pdf = cell volume / cell probability.
template <int D> __device__ __forceinline__ VolSample<D> sample_grid(const VolSrc &v, Rng &rng) {
    VolSample<D> r;
    const float unif = rng.uniform();
    const uint32_t idx = cdf_lookup(v.cdf, v.guide, v.n, unif);
    const float dx = v.domain_size / (float)v.grid_res;
    uint32_t u[3] = {idx % v.grid_res, (idx / v.grid_res) % v.grid_res, idx / (v.grid_res * v.grid_res)};
    if constexpr (D == 2) u[1] = idx / v.grid_res;
#pragma unroll
    for (int k = 0; k < D; ++k) r.p[k] = ((float)u[k] + 0.5f) * dx - (v.domain_size / 2);   // utils.cuh:92-99
#pragma unroll
    for (int k = 0; k < D; ++k) {
        const float uu = rng.uniform();
        r.p[k] += (uu - 0.5f) * v.domain_size / (float)v.grid_res;
    }
    r.value = v.cache[idx];
    const float pdf = idx == 0 ? v.cdf[idx] : v.cdf[idx] - v.cdf[idx - 1];
    float cellv = 1.f;
#pragma unroll
    for (int k = 0; k < D; ++k) cellv *= v.domain_size / (float)v.grid_res;
    r.inv_pdf = cellv / pdf;
    return r;
}

// Gaussian mixture: component j from the mixture CDF, Box-Muller normals from (0, 1]
// uniforms, value f(p), inv_pdf = 1 / mixture density(p); with the mask, value 0 on the far
// side of the boundary.
template <int D, int STACK>
__device__ __forceinline__ VolSample<D> sample_gauss(const DevScene &S, Rng &rng, bool interior, uint32_t *stack,
                                                     uint32_t stride) {
    const VolSrc &v = S.vol;
    VolSample<D> r;
    const float u = rng.uniform();
    uint32_t j = 0;
    while (j + 1 < v.n && u > v.cdf[j]) ++j;
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const float r1 = sqrtf(-2.f * logf(u1));
    // the angles' cos and sin in double, rounded once, as the oracle evaluates them
    double sn, cs;
    wobtrig::sincos_small(2.0 * kPiD * (double)u2, sn, cs);
    float z[3];
    z[0] = r1 * (float)cs;
    z[1] = r1 * (float)sn;
    if constexpr (D == 3) {
        const float u3 = rng.uniform(), u4 = rng.uniform();
        double s2, c2;
        wobtrig::sincos_small(2.0 * kPiD * (double)u4, s2, c2);
        z[2] = sqrtf(-2.f * logf(u3)) * (float)c2;
    }
    const float4 cj = v.comp[2 * j];
    const float cjv[3] = {cj.x, cj.y, cj.z};
#pragma unroll
    for (int k = 0; k < D; ++k) r.p[k] = cjv[k] + cj.w * z[k];
    float f = 0.f, pdf = 0.f;
    for (uint32_t i = 0; i < v.n; ++i) {
        const float4 a = v.comp[2 * i], b = v.comp[2 * i + 1];
        const float av[3] = {a.x, a.y, a.z};
        float d2 = 0.f;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            const float dk = r.p[k] - av[k];
            d2 += dk * dk;
        }
        const float e = expf(-d2 * b.z);
        f += b.x * e;
        pdf += b.y * e;
    }
    r.value = f;
    r.inv_pdf = 1.0f / pdf;
    if (v.mask) {
        uint32_t c = inside_grid_cell<D>(v, r.p);
        if (c == 2u)
            c = ray_crossings<D, STACK>(S.nodes, S.prims, r.p[0], r.p[1], D == 3 ? r.p[D - 1] : 0.f, stack, stride) & 1u;
        if ((c != 0u) != interior) r.value = 0.f;
    }
    return r;
}

// No volume source: every volume term is zero and draws nothing.
struct NoVol {
    template <int D> __device__ __forceinline__ float term(const BP<D> &, const Kern<D> &, Rng &) const { return 0.f; }
    template <int D> __device__ __forceinline__ float init_domain(const float *, const Kern<D> &, Rng &) const {
        return 0.f;
    }
    template <int D> __device__ __forceinline__ float init_boundary(const float *, const Kern<D> &, Rng &) const {
        return 0.f;
    }
    template <int D>
    __device__ __forceinline__ void init_derivative(const float *, const Kern<D> &, Rng &, float *g) const {
#pragma unroll
        for (int i = 0; i < D; ++i) g[i] = 0.f;
    }
};

// Volume terms of bie_models.cuh:79-154.  The reference draws ONE sample per call and adds its
// contribution num_volume_samples times; with `indep` every addend has its own sample.
template <int D, int STACK> struct VolCtx {
    const DevScene *S;
    uint32_t nvs;
    uint32_t indep;
    bool interior;
    uint32_t *stack;
    uint32_t stride;

    __device__ __forceinline__ VolSample<D> sample(Rng &rng) const {
        if (S->vol.kind == kVolGrid) return sample_grid<D>(S->vol, rng);
        return sample_gauss<D, STACK>(*S, rng, interior, stack, stride);
    }
    // :79-103
    __device__ __forceinline__ float term(const BP<D> &bp, const Kern<D> &k, Rng &rng) const {
        if (nvs == 0) return 0.f;
        VolSample<D> v;
        float res = 0.f;
        for (uint32_t i = 0; i < nvs; ++i) {
            if (i == 0 || indep) v = sample(rng);
            if (bp.type == kDirichlet) res += k.G(bp.p, v.p) * v.value * v.inv_pdf;
            else if (bp.type == kNeumann) res += k.dG_dny(bp.p, v.p, bp.n) * v.value * v.inv_pdf;
            else res += (k.dG_dny(bp.p, v.p, bp.n) + bp.alpha * k.G(v.p, bp.p)) * v.value * v.inv_pdf;
        }
        return res / (float)nvs;
    }
    // :121-135 (domain point)
    __device__ __forceinline__ float init_domain(const float *x, const Kern<D> &k, Rng &rng) const {
        if (nvs == 0) return 0.f;
        VolSample<D> v;
        float res = 0.f;
        for (uint32_t i = 0; i < nvs; ++i) {
            if (i == 0 || indep) v = sample(rng);
            res -= k.G(v.p, x) * v.value * v.inv_pdf;
        }
        return res / (float)nvs;
    }
    // :105-119 (boundary point)
    __device__ __forceinline__ float init_boundary(const float *x, const Kern<D> &k, Rng &rng) const {
        if (nvs == 0) return 0.f;
        VolSample<D> v;
        float res = 0.f;
        for (uint32_t i = 0; i < nvs; ++i) {
            if (i == 0 || indep) v = sample(rng);
            res -= k.G(x, v.p) * v.value * v.inv_pdf;
        }
        return res / (float)nvs;
    }
    // :137-154
    __device__ __forceinline__ void init_derivative(const float *x, const Kern<D> &k, Rng &rng, float *g) const {
#pragma unroll
        for (int i = 0; i < D; ++i) g[i] = 0.f;
        if (nvs == 0) return;
        VolSample<D> v;
        for (uint32_t i = 0; i < nvs; ++i) {
            if (i == 0 || indep) v = sample(rng);
            float dg[D];
            k.dG_dx(v.p, x, dg);
            const float w = v.value * v.inv_pdf;
#pragma unroll
            for (int c = 0; c < D; ++c) g[c] = g[c] - w * dg[c];
        }
#pragma unroll
        for (int c = 0; c < D; ++c) g[c] = g[c] / (float)nvs;
    }
};

// ----------------------------------------------------------------------------------------
// BIE models (bie_models.cuh).  interior_angle is 0.5 for every sample the scene produces
// (scene.cuh:455, 510, 615, 750, 815, 910), so integral_free_term is 0.5 for both domains.
// ----------------------------------------------------------------------------------------
template <int D, int BIE> struct Model {
    Kern<D> k;
    float dsign;   // domain_type_coefficient (+1 interior, -1 exterior)
    float c;       // first_kind_equation_scaling_constant
    static constexpr float ift = 0.5f;

    __device__ __forceinline__ float solution_domain_unknown(const float *x, const BP<D> &y, float inv) const {
        if constexpr (BIE == kSL) return k.G(y.p, x) * inv;                                       // :228-231
        else if constexpr (BIE == kDL) return -k.dG_dny(y.p, x, y.n) * inv;                       // :341-344
        else {
            if (y.type == kDirichlet) return dsign * inv * k.G(y.p, x);                           // :429-435
            if (y.type == kNeumann) return -dsign * inv * k.dG_dny(y.p, x, y.n);
            return -dsign * inv * (k.dG_dny(y.p, x, y.n) + y.alpha * k.G(y.p, x));
        }
    }
    __device__ __forceinline__ void gradient_domain_unknown(const float *x, const BP<D> &y, float inv, float *g) const {
        if constexpr (BIE == kSL) {                                                               // :240-243
            k.dG_dx(y.p, x, g);
#pragma unroll
            for (int i = 0; i < D; ++i) g[i] = g[i] * inv;
        } else if constexpr (BIE == kDL) {                                                        // :353-356
            k.d2G_dxdny(y.p, x, y.n, g);
#pragma unroll
            for (int i = 0; i < D; ++i) g[i] = -(g[i] * inv);
        } else {                                                                                  // :452-462
            if (y.type == kDirichlet) {
                k.dG_dx(y.p, x, g);
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = (dsign * inv) * g[i];
            } else if (y.type == kNeumann) {
                k.d2G_dxdny(y.p, x, y.n, g);
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = (-dsign * inv) * g[i];
            } else {
                float h[D];
                k.d2G_dxdny(y.p, x, y.n, g);
                k.dG_dx(y.p, x, h);
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = (-dsign * inv) * (g[i] + y.alpha * h[i]);
            }
        }
    }
    // known terms; `vt` supplies the volume terms (NoVol: none) and draws their samples in the
    // reference's order (the domain-point init term first, then the term at y)
    template <typename VT>
    __device__ __forceinline__ float solution_domain_known(const float *x, const BP<D> &y, float inv, const VT &vt,
                                                           Rng &rng) const {
        const float vi = vt.init_domain(x, k, rng);                                               // :233-238, :346-351, :442
        if constexpr (BIE != kDirect) return vi;
        else {
            const float b = y.bv + vt.term(y, k, rng);
            if (y.type == kDirichlet) return vi + -dsign * inv * k.dG_dny(y.p, x, y.n) * b;     // :443-445
            return vi + dsign * inv * k.G(y.p, x) * b;                                           // :446-448
        }
    }
    template <typename VT>
    __device__ __forceinline__ void gradient_domain_known(const float *x, const BP<D> &y, float inv, const VT &vt,
                                                          Rng &rng, float *g) const {
        float gv[D];
        vt.init_derivative(x, k, rng, gv);                                                        // :245-250, :358-363, :468
        if constexpr (BIE != kDirect) {
#pragma unroll
            for (int i = 0; i < D; ++i) g[i] = gv[i];
        } else {                                                                                  // :469-477
            const float b = y.bv + vt.term(y, k, rng);
            if (y.type == kDirichlet) {
                k.d2G_dxdny(y.p, x, y.n, g);
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = gv[i] + ((-dsign * inv) * g[i]) * b;
            } else {
                k.dG_dx(y.p, x, g);
#pragma unroll
                for (int i = 0; i < D; ++i) g[i] = gv[i] + ((dsign * inv) * g[i]) * b;
            }
        }
    }
    __device__ __forceinline__ float solution_boundary_unknown(const BP<D> &x, const BP<D> &y, float inv) const {
        if constexpr (BIE == kSL) {                                                               // :252-257
            if (x.type == kDirichlet) return 0.f;
            return k.G(y.p, x.p) * inv;
        } else if constexpr (BIE == kDL) {                                                        // :365-369
            return 0.f;
        } else {                                                                                  // :480-493
            if (x.type == kDirichlet) return 0.f;
            if (y.type == kDirichlet) return dsign / ift * inv * k.G(y.p, x.p);
            if (y.type == kNeumann) return -dsign / ift * inv * k.dG_dny(y.p, x.p, y.n);
            return -dsign / ift * inv * (k.dG_dny(y.p, x.p, y.n) + y.alpha * k.G(y.p, x.p));
        }
    }
    template <typename VT>
    __device__ __forceinline__ float solution_boundary_known(const BP<D> &x, const BP<D> &y, float inv, const VT &vt,
                                                             Rng &rng) const {
        if constexpr (BIE == kSL) {                                                               // :259-266
            if (x.type == kDirichlet) return x.bv;
            return vt.init_boundary(x.p, k, rng);
        } else if constexpr (BIE == kDL) {                                                        // :371-377
            return x.bv;
        } else {                                                                                  // :495-511
            if (x.type == kDirichlet) return x.bv;
            const float vi = vt.init_boundary(x.p, k, rng);
            const float b = y.bv + vt.term(y, k, rng);
            if (y.type == kDirichlet) return vi + -dsign / ift * inv * k.dG_dny(y.p, x.p, y.n) * b;
            return vi + dsign / ift * inv * k.G(y.p, x.p) * b;
        }
    }
    // Fredholm unknown after the first-kind splitting has been resolved (split not taken)
    __device__ __forceinline__ float fredholm_unknown_kernel(const BP<D> &x, const BP<D> &y, float inv) const {
        if constexpr (BIE == kDL) {                                                               // :386-387
            return dsign / (1.f - ift) * inv * k.dG_dny(y.p, x.p, y.n);
        } else if constexpr (BIE == kSL) {                                                        // :290-299
            if (x.type == kDirichlet) return -2.f * c * inv * k.G(y.p, x.p);
            if (x.type == kNeumann) return -dsign / ift * inv * k.dG_dny(x.p, y.p, x.n);
            return -dsign / ift * inv * (k.dG_dny(x.p, y.p, x.n) + x.alpha * k.G(y.p, x.p));
        } else {                                                                                  // :535-555
            if (x.type == kDirichlet) {
                if (y.type == kDirichlet) return -c / ift * inv * k.G(y.p, x.p);
                if (y.type == kNeumann) return c / ift * inv * k.dG_dny(y.p, x.p, y.n);
                return c / ift * inv * (k.dG_dny(y.p, x.p, y.n) + y.alpha * k.G(y.p, x.p));
            }
            if (y.type == kDirichlet) return dsign / ift * inv * k.G(y.p, x.p);
            if (y.type == kNeumann) return -dsign / ift * inv * k.dG_dny(y.p, x.p, y.n);
            return -dsign / ift * inv * (k.dG_dny(y.p, x.p, y.n) + y.alpha * k.G(y.p, x.p));
        }
    }
    // bie_models.cuh:268-300 (SL), 379-388 (DL), 513-556 (Direct).  Backward form: may set y = x.
    __device__ __forceinline__ float fredholm_unknown_backward(const BP<D> &x, BP<D> &y, float inv, Rng &rng) const {
        if constexpr (BIE != kDL) {
            if (x.type == kDirichlet) {
                if (rng.uniform() * (1.0f + c) < 1.0f) { y = x; return 1.0f + c; }
                inv *= (1.0f + c) / c;
            }
        }
        return fredholm_unknown_kernel(x, y, inv);
    }
    // forward form (is_backward_sampling = false): the split looks at y and may set x = y
    __device__ __forceinline__ float fredholm_unknown_forward(BP<D> &x, const BP<D> &y, float inv, Rng &rng) const {
        if constexpr (BIE != kDL) {
            if (y.type == kDirichlet) {
                if (rng.uniform() * (1.0f + c) < 1.0f) { x = y; return 1.0f + c; }
                inv *= (1.0f + c) / c;
            }
        }
        return fredholm_unknown_kernel(x, y, inv);
    }
    // bie_models.cuh:301-316 (SL), 389-399 (DL), 557-587 (Direct); the term at x is drawn
    // before the term at y
    template <typename VT>
    __device__ __forceinline__ float fredholm_known(const BP<D> &x, const BP<D> &y, float inv, const VT &vt,
                                                    Rng &rng) const {
        if constexpr (BIE == kSL) {
            const float a = x.bv + vt.term(x, k, rng);
            if (x.type == kDirichlet) return 2.f * c * a;
            return dsign / ift * a;
        } else if constexpr (BIE == kDL) {
            return dsign * 1.f / (1.f - ift) * (x.bv + vt.term(x, k, rng));
        } else {
            const bool xd = x.type == kDirichlet;
            const float a = xd ? x.bv + vt.term(x, k, rng) : 0.f;
            const float b = y.bv + vt.term(y, k, rng);
            if (xd) {
                if (y.type == kDirichlet) return dsign * c * (a + dsign / ift * inv * k.dG_dny(y.p, x.p, y.n) * b);
                return dsign * c * (a - dsign / ift * inv * k.G(y.p, x.p) * b);
            }
            if (y.type == kDirichlet) return -dsign / ift * inv * k.dG_dny(y.p, x.p, y.n) * b;
            return dsign / ift * inv * k.G(y.p, x.p) * b;
        }
    }
    // without a volume source (deterministic parity kernels)
    __device__ __forceinline__ float fredholm_known(const BP<D> &x, const BP<D> &y, float inv) const {
        Rng unused;
        return fredholm_known(x, y, inv, NoVol{}, unused);
    }
};

// ----------------------------------------------------------------------------------------
// Estimators (estimator.cuh)
// ----------------------------------------------------------------------------------------
struct WalkParams {
    uint32_t path_length;
    uint32_t nrc;
    uint32_t exclude_origin;   // WOB_EST_EXCLUDE_ORIGIN_ELEMENT
    uint32_t cv;               // WOB_EST_CONTROL_VARIATE (single-layer backward walks)
    uint32_t cv_samples;       // samples of the domain-level control-variate integral
    uint32_t density;          // WOB_EST_BOUNDARY_DENSITY: boundary estimates return the walk value
    uint32_t hcv;              // WOB_EST_HARMONIC_CV: the walk carries the residual data of u_h
    __device__ __forceinline__ uint32_t skip(uint32_t elem) const { return exclude_origin ? elem : 0xFFFFFFFFu; }
};

// Control-variate value V_r at a sampled point: linear interpolation of the per-vertex table
// of value function r (remaining path length r = 1 .. cv_R).
template <int D> __device__ __forceinline__ float cv_value(const DevScene &S, uint32_t r, const BP<D> &b) {
    if (b.elem == 0xFFFFFFFFu || r == 0 || r > S.cv_R) return 0.f;
    const float *t = S.cv_table + ((size_t)(r - 1) * S.n + b.elem) * D;
    float v = b.bw[0] * t[0];
#pragma unroll
    for (int k = 1; k < D; ++k) v += b.bw[k] * t[k];
    return v;
}

// Harmonic control variate u_h and its gradient at p.  The sources are the same for every
// lane of a wave, so their records are read through the constant address space (scalar
// loads).  Terms are added in source order after c0; phi uses the hardware reciprocal square
// root (3D) or base-2 logarithm (2D).
__device__ __forceinline__ float4 hcv_source(const float4 *src, uint32_t j) {
    const gf4_t v = *(const __attribute__((address_space(4))) gf4_t *)(src + j);
    return make_float4(v.x, v.y, v.z, v.w);
}
template <int D> __device__ __forceinline__ float hcv_value(const DevScene &S, const float *p) {
    float acc = S.hcv_c0;
    for (uint32_t j = 0; j < S.hcv_m; ++j) {
        const float4 s = hcv_source(S.hcv_src, j);
        if constexpr (D == 3) {
            const float d0 = p[0] - s.x, d1 = p[1] - s.y, d2 = p[2] - s.z;
            const float z = d0 * d0 + d1 * d1 + d2 * d2;
            acc += s.w * __builtin_amdgcn_rsqf(z);
        } else {
            const float d0 = p[0] - s.x, d1 = p[1] - s.y;
            const float z = d0 * d0 + d1 * d1;
            acc += s.z * __builtin_amdgcn_logf(z);
        }
    }
    return acc;
}
// grad u_h: 3D sum_j -w_j (p - s_j) |p - s_j|^-3; 2D (2 / ln 2) sum_j w_j (p - s_j) / |p - s_j|^2
template <int D> __device__ __forceinline__ void hcv_gradient(const DevScene &S, const float *p, float *g) {
#pragma unroll
    for (int k = 0; k < D; ++k) g[k] = 0.f;
    for (uint32_t j = 0; j < S.hcv_m; ++j) {
        const float4 s = hcv_source(S.hcv_src, j);
        if constexpr (D == 3) {
            const float d[3] = {p[0] - s.x, p[1] - s.y, p[2] - s.z};
            const float z = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
            const float ri = __builtin_amdgcn_rsqf(z);
            const float f = -s.w * (ri * ri * ri);
#pragma unroll
            for (int k = 0; k < 3; ++k) g[k] += f * d[k];
        } else {
            const float d[2] = {p[0] - s.x, p[1] - s.y};
            const float z = d[0] * d[0] + d[1] * d[1];
            const float f = s.z * __builtin_amdgcn_rcpf(z);
#pragma unroll
            for (int k = 0; k < 2; ++k) g[k] += f * d[k];
        }
    }
    if constexpr (D == 2) {
#pragma unroll
        for (int k = 0; k < 2; ++k) g[k] = g[k] * 2.8853900817779268f;   // 2 / ln 2
    }
}

// The residual datum at a boundary point: the point's datum minus u_h's datum of the same
// type (u_h at Dirichlet points, du_h/dn at Neumann points, du_h/dn + alpha u_h at Robin points).
// Out of line, with the sources passed as scalars: the walks call it only with the control
// variate on, and inlined its registers would count against every walk.
template <int D>
__device__ __noinline__ float hcv_residual_datum(const float4 *src, uint32_t m, float c0, float p0, float p1, float p2,
                                                 float n0, float n1, float n2, int type, float alpha, float bv) {
    DevScene S{};
    S.hcv_src = src;
    S.hcv_m = m;
    S.hcv_c0 = c0;
    const float p[3] = {p0, p1, p2}, n[3] = {n0, n1, n2};
    if (type == kDirichlet) return bv - hcv_value<D>(S, p);
    float g[D];
    hcv_gradient<D>(S, p, g);
    float dn = g[0] * n[0];
#pragma unroll
    for (int k = 1; k < D; ++k) dn += g[k] * n[k];
    if (type == kNeumann) return bv - dn;
    return bv - (dn + alpha * hcv_value<D>(S, p));
}
template <int D> __device__ __forceinline__ BP<D> hcv_residual(const DevScene &S, const BP<D> &y) {
    BP<D> r = y;
    r.bv = hcv_residual_datum<D>(S.hcv_src, S.hcv_m, S.hcv_c0, y.p[0], y.p[1], D == 3 ? y.p[D - 1] : 0.f, y.n[0],
                                 y.n[1], D == 3 ? y.n[D - 1] : 0.f, y.type, y.alpha, y.bv);
    return r;
}

// resampled importance sampling, scene.cuh:542-584 (target by the type of y_prev,
// estimator.cuh:286-328); candidates are uniform samples.  With cv_r > 0 it also returns
// ksum = sum_i s_i V_{cv_r}(z_i), s_i the signed kernel whose magnitude is the target.
template <int D, int BIE>
__device__ __forceinline__ Sample<D> ris(const DevScene &S, const Model<D, BIE> &m, const BP<D> &yp, uint32_t nrc,
                                         Rng &rng, uint32_t cv_r, float &ksum) {
    // the target needs only the candidate's position; the boundary data of the chosen one are
    // read once, after the loop
    float w_sum = 0.f;
    ksum = 0.f;
    const int tp = yp.type;
    // each kernel evaluated at most once per lane, so a wave whose lanes hold all three types
    // runs G and dG/dn once each (a Robin branch of its own would run them a second time)
    auto signed_kernel = [&](const float *y) -> float {
        float g = 0.f, dg = 0.f;
        if (tp != kNeumann) g = m.k.G(y, yp.p);
        if (tp != kDirichlet) dg = m.k.dG_dny(yp.p, y, yp.n);
        return tp == kDirichlet ? g : tp == kNeumann ? dg : dg + yp.alpha * g;
    };
    GeomSample<D> pick;
    pick.elem = 0xFFFFFFFFu;
#pragma unroll
    for (int k = 0; k < D; ++k) { pick.p[k] = 0.f; pick.n[k] = 0.f; pick.bw[k] = 0.f; }
    auto update = [&](const GeomSample<D> &xi, float r) {
        const float sk = signed_kernel(xi.p);
        const float w = fabsf(sk) * S.total_area;
        if (cv_r) {
            BP<D> c;
            c.elem = xi.elem;
#pragma unroll
            for (int k = 0; k < D; ++k) c.bw[k] = xi.bw[k];
            ksum += sk * cv_value<D>(S, cv_r, c);
        }
        w_sum += w;
        const bool take = r <= w / w_sum;
        pick.elem = take ? xi.elem : pick.elem;
#pragma unroll
        for (int k = 0; k < D; ++k) {
            pick.p[k] = take ? xi.p[k] : pick.p[k];
            pick.n[k] = take ? xi.n[k] : pick.n[k];
            pick.bw[k] = take ? xi.bw[k] : pick.bw[k];
        }
    };
    // 2D: candidates two at a time.  A candidate's D uniforms and its reservoir uniform are
    // drawn in the reference's order (the draws do not depend on the data), then both element
    // lookups and attribute loads are in flight together before the reservoir takes them in
    // order.  (In 3D the second candidate's registers cost more occupancy than the overlap
    // gains.)
    uint32_t i = 0;
    for (; D == 2 && i + 2 <= nrc; i += 2) {
        float ua[D], ub[D];
#pragma unroll
        for (int k = 0; k < D; ++k) ua[k] = rng.uniform();
        const float ra = rng.uniform();
#pragma unroll
        for (int k = 0; k < D; ++k) ub[k] = rng.uniform();
        const float rb = rng.uniform();
        uint32_t ea, eb;
        cdf_lookup2(S.area_cdf, S.area_guide, S.n, ua[0], ub[0], ea, eb);
        const GeomSample<D> xa = uniform_geometry_of<D, false>(S, ea, ua, nullptr);
        const GeomSample<D> xb = uniform_geometry_of<D, false>(S, eb, ub, nullptr);
        update(xa, ra);
        update(xb, rb);
    }
    for (; i < nrc; ++i) {
        const GeomSample<D> xa = uniform_geometry<D>(S, rng);
        update(xa, rng.uniform());
    }
    if constexpr (D == 2) {   // the kept candidate's normal (the pair loop skips candidates' normals)
        if (pick.elem != 0xFFFFFFFFu) {
            const float4 f1 = gld4(S.attr + 3 * pick.elem + 1);
            pick.n[0] = f1.x;
            pick.n[1] = f1.y;
        }
    }
    Sample<D> r;
    r.y = pick.elem == 0xFFFFFFFFu ? bp_zero<D>() : complete_uniform<D>(S, pick);
    r.inv_pdf = w_sum / fmaxf(fabsf(signed_kernel(r.y.p)) * (float)nrc, 10 * kFltEps);
    return r;
}

// estimator.cuh:232-357
// With the control variate (single-layer walks), zero-mean terms
//   D_m = W_m E[fu_m V_{r}(y_{m+1}) | y_m, candidates] - W_{m+1} V_{r}(y_{m+1}),  r = L - m - 1,
// are added for m <= L - 2, where the conditional mean is exact given the RIS candidates:
//   Dirichlet y_m (first-kind split): V_r(y_m) - 2c |G| / M sum_i G(z_i, y_m) V_r(z_i)
//   Neumann / Robin y_m:              -(dsign / ift) |G| / M sum_i s_i V_r(z_i).
// The walk's expectation is unchanged for any table V; with V equal to the value functions
// of the walk the path-choice randomness cancels.
// With the harmonic control variate (single and double layer), the known term at y_m uses the
// residual datum of the scene's harmonic function u_h (hcv_residual), and the domain estimate
// adds u_h(x): u - u_h solves the same problem with the residual data, so for any u_h harmonic
// in the domain the expectation is unchanged (up to the walk's truncation, now of the
// residual's series), and the variance falls with the size of the residual data.
template <int D, int BIE, int FL, int STACK, typename VT>
__device__ __forceinline__ float walk_backward(const DevScene &S, const Model<D, BIE> &m, const WalkParams &wp,
                                               const VT &vt, const BP<D> &x0, Rng &rng, uint32_t *stack,
                                               uint32_t stride) {
    float spc = 0.f, W = 1.f;
    BP<D> yp = x0;
    for (uint32_t it = 0; it < wp.path_length; ++it) {
        if (it == wp.path_length - 1) W = 0.5f * W;
        if constexpr (FL == FL_LINE_LINE) {
            Sample<D> q = line_sample<D, STACK>(S, yp.p, wp.skip(yp.elem), rng, stack, stride);
            if (wp.hcv) spc += W * m.fredholm_known(hcv_residual<D>(S, yp), q.y, q.inv_pdf, vt, rng);
            else spc += W * m.fredholm_known(yp, q.y, q.inv_pdf, vt, rng);
            W *= m.fredholm_unknown_backward(yp, q.y, q.inv_pdf, rng);
            yp = q.y;
        } else if constexpr (FL == FL_LINE_BV) {
            // with the control variate the known term's sample is uniform (the residual data need
            // not vanish where the data do) and both points carry residual data
            if (wp.hcv) {
                const Sample<D> q = sample_uniform<D>(S, rng);
                spc += W * m.fredholm_known(hcv_residual<D>(S, yp), hcv_residual<D>(S, q.y), q.inv_pdf, vt, rng);
            } else {
                const Sample<D> q = sample_bv<D>(S, rng);
                spc += W * m.fredholm_known(yp, q.y, q.inv_pdf, vt, rng);
            }
            Sample<D> q = line_sample<D, STACK>(S, yp.p, wp.skip(yp.elem), rng, stack, stride);
            W = W * m.fredholm_unknown_backward(yp, q.y, q.inv_pdf, rng);
            yp = q.y;
        } else if constexpr (BIE == kSL) {
            // SL ignores y here (estimator.cuh:278-281)
            if (wp.hcv) spc += W * m.fredholm_known(hcv_residual<D>(S, yp), yp, 0.f, vt, rng);
            else spc += W * m.fredholm_known(yp, yp, 0.f, vt, rng);
            const uint32_t cv_r = (wp.cv && it + 1 < wp.path_length) ? wp.path_length - it - 1 : 0u;
            float ksum = 0.f;
            Sample<D> q;
            if (wp.nrc == 1) {
                q = sample_uniform<D>(S, rng);
                if (cv_r) {
                    const float sk = yp.type == kDirichlet ? m.k.G(q.y.p, yp.p)
                                     : yp.type == kNeumann ? m.k.dG_dny(yp.p, q.y.p, yp.n)
                                                           : m.k.dG_dny(yp.p, q.y.p, yp.n) + yp.alpha * m.k.G(q.y.p, yp.p);
                    ksum = sk * cv_value<D>(S, cv_r, q.y);
                }
            } else {
                q = ris<D, BIE>(S, m, yp, wp.nrc, rng, cv_r, ksum);
            }
            const float fu = m.fredholm_unknown_backward(yp, q.y, q.inv_pdf, rng);
            if (cv_r) {
                const float scale = S.total_area / (float)wp.nrc;
                const float e = yp.type == kDirichlet ? cv_value<D>(S, cv_r, yp) - 2.f * m.c * scale * ksum
                                                      : -(m.dsign / m.ift) * scale * ksum;
                spc += W * e - (W * fu) * cv_value<D>(S, cv_r, q.y);
            }
            W = W * fu;
            yp = q.y;
        } else {
            Sample<D> q = sample_uniform<D>(S, rng);
            if (wp.hcv) spc += W * m.fredholm_known(hcv_residual<D>(S, yp), hcv_residual<D>(S, q.y), q.inv_pdf, vt, rng);
            else spc += W * m.fredholm_known(yp, q.y, q.inv_pdf, vt, rng);
            q = sample_uniform<D>(S, rng);
            W = W * m.fredholm_unknown_backward(yp, q.y, q.inv_pdf, rng);
            yp = q.y;
        }
        if (W == 0.0f) break;
    }
    return spc;
}

// first samples of the backward estimators (estimator.cuh:158-168, 199-209)
// (with the control variate the known terms' sample y1 is uniform, as in walk_backward)
template <int D, int FL, int STACK>
__device__ __forceinline__ void first_samples(const DevScene &S, const float *x, uint32_t skip, Sample<D> &y,
                                              Sample<D> &y1, Rng &rng, uint32_t *stack, uint32_t stride,
                                              bool hcv = false) {
    if constexpr (FL == FL_LINE_LINE) {
        y = line_sample<D, STACK>(S, x, skip, rng, stack, stride);
        y1 = y;
    } else if constexpr (FL == FL_LINE_BV) {
        y = line_sample<D, STACK>(S, x, skip, rng, stack, stride);
        y1 = hcv ? sample_uniform<D>(S, rng) : sample_bv<D>(S, rng);
    } else {
        y = sample_uniform<D>(S, rng);
        y1 = sample_uniform<D>(S, rng);
    }
}

// one domain path: out = [u, grad u]
template <int D, int BIE, int EST, int FL, int STACK, typename VT>
__device__ __forceinline__ void domain_path(const DevScene &S, const Model<D, BIE> &m, const WalkParams &wp,
                                            const VT &vt, const float *x, Rng &rng, uint32_t *stack, uint32_t stride,
                                            float *out) {
    if constexpr (EST == kBackward) {                                                          // estimator.cuh:154-191
        Sample<D> y, y1;
        first_samples<D, FL, STACK>(S, x, 0xFFFFFFFFu, y, y1, rng, stack, stride, wp.hcv != 0);
        if constexpr (BIE == kDirect) {
            if (wp.hcv) y1.y = hcv_residual<D>(S, y1.y);   // the known terms' datum
        }
        const float known = m.solution_domain_known(x, y1.y, y1.inv_pdf, vt, rng);
        float gk[D], gu[D];
        m.gradient_domain_known(x, y1.y, y1.inv_pdf, vt, rng, gk);
        const float unk = m.solution_domain_unknown(x, y.y, y.inv_pdf);
        m.gradient_domain_unknown(x, y.y, y.inv_pdf, gu);
        const float spc = walk_backward<D, BIE, FL, STACK>(S, m, wp, vt, y.y, rng, stack, stride);
        out[0] = known + spc * unk;
#pragma unroll
        for (int k = 0; k < D; ++k) out[1 + k] = gk[k] + spc * gu[k];
        if (wp.hcv) {   // u = u_h + the walk's estimate of u - u_h (residual data, the same source)
            float gh[D];
            hcv_gradient<D>(S, x, gh);
            out[0] += hcv_value<D>(S, x);
#pragma unroll
            for (int k = 0; k < D; ++k) out[1 + k] += gh[k];
        }
        if constexpr (BIE == kSL && FL == FL_UNIFORM) {
            if (wp.cv) {
                // I(x) - unk V_L(y0): I(x) = |G| / M0 sum_j G(z_j, x) V_L(z_j) from an independent
                // Philox stream, so a zero table leaves every other draw and the result unchanged
                const float v0 = cv_value<D>(S, wp.path_length, y.y);
                Rng r2 = rng;
                r2.block = 0; r2.left = 0; r2.stream = rng.stream ^ 0x80000000u;
                float iu = 0.f, ig[D];
#pragma unroll
                for (int k = 0; k < D; ++k) ig[k] = 0.f;
                for (uint32_t j = 0; j < wp.cv_samples; ++j) {
                    const Sample<D> z = sample_uniform<D>(S, r2);
                    const float v = cv_value<D>(S, wp.path_length, z.y);
                    iu += m.k.G(z.y.p, x) * v;
                    float g[D];
                    m.k.dG_dx(z.y.p, x, g);
#pragma unroll
                    for (int k = 0; k < D; ++k) ig[k] += g[k] * v;
                }
                const float sc = wp.cv_samples ? S.total_area / (float)wp.cv_samples : 0.f;
                out[0] += sc * iu - unk * v0;
#pragma unroll
                for (int k = 0; k < D; ++k) out[1 + k] += sc * ig[k] - gu[k] * v0;
            }
        }
    } else {                                                                                   // estimator.cuh:405-454
        // With the harmonic control variate the start point is uniform (the residual data need
        // not vanish where the data do, so the data's own density would miss part of them) and
        // the known term reads the residual datum.
        Sample<D> y = sample_uniform<D>(S, rng);
        out[0] = m.solution_domain_known(x, y.y, y.inv_pdf, vt, rng);
        float g[D], gu[D];
        m.gradient_domain_known(x, y.y, y.inv_pdf, vt, rng, g);
        if (FL == FL_FWD_LINE && !wp.hcv) y = sample_bv<D>(S, rng);
        else y = sample_uniform<D>(S, rng);
        const Sample<D> y2 = sample_uniform<D>(S, rng);
        float W = y.inv_pdf * m.fredholm_known(wp.hcv ? hcv_residual<D>(S, y.y) : y.y, y2.y, y2.inv_pdf, vt, rng);
        BP<D> yc = y.y;
        for (uint32_t it = 0; it < wp.path_length; ++it) {
            if (it == wp.path_length - 1) W *= 0.5f;
            out[0] += W * m.solution_domain_unknown(x, yc, 1.0f);
            m.gradient_domain_unknown(x, yc, 1.0f, gu);
#pragma unroll
            for (int k = 0; k < D; ++k) g[k] += W * gu[k];
            BP<D> yp = yc;
            Sample<D> q;
            if constexpr (FL == FL_FWD_LINE) q = line_sample<D, STACK>(S, yp.p, wp.skip(yp.elem), rng, stack, stride);
            else q = sample_uniform<D>(S, rng);
            yc = q.y;
            W = m.fredholm_unknown_forward(yc, yp, q.inv_pdf, rng) * W;
        }
#pragma unroll
        for (int k = 0; k < D; ++k) out[1 + k] = g[k];
        if (wp.hcv) {   // u = u_h + the walk's estimate of u - u_h
            float gh[D];
            hcv_gradient<D>(S, x, gh);
            out[0] += hcv_value<D>(S, x);
#pragma unroll
            for (int k = 0; k < D; ++k) out[1 + k] += gh[k];
        }
    }
}

// one boundary path (estimator.cuh:193-222, forward :456-496)
template <int D, int BIE, int EST, int FL, int STACK, typename VT>
__device__ __forceinline__ float boundary_path(const DevScene &S, const Model<D, BIE> &m, const WalkParams &wp,
                                               const VT &vt, const BP<D> &x, Rng &rng, uint32_t *stack,
                                               uint32_t stride) {
    if constexpr (EST == kBackward) {
        if (wp.density) return walk_backward<D, BIE, FL, STACK>(S, m, wp, vt, x, rng, stack, stride);
        Sample<D> y, y1;
        first_samples<D, FL, STACK>(S, x.p, wp.skip(x.elem), y, y1, rng, stack, stride);
        float res = m.solution_boundary_known(x, y1.y, y1.inv_pdf, vt, rng);
        const float unk = m.solution_boundary_unknown(x, y.y, y.inv_pdf);
        const float spc = walk_backward<D, BIE, FL, STACK>(S, m, wp, vt, y.y, rng, stack, stride);
        return res + spc * unk;
    } else {
        Sample<D> y = sample_uniform<D>(S, rng);
        float res = m.solution_boundary_known(x, y.y, y.inv_pdf, vt, rng);
        if constexpr (FL == FL_FWD_LINE) y = sample_bv<D>(S, rng);
        else y = sample_uniform<D>(S, rng);
        const Sample<D> y2 = sample_uniform<D>(S, rng);
        float W = y.inv_pdf * m.fredholm_known(y.y, y2.y, y2.inv_pdf, vt, rng);
        BP<D> yc = y.y;
        for (uint32_t it = 0; it < wp.path_length; ++it) {
            if (it == wp.path_length - 1) W *= 0.5f;
            res += m.solution_boundary_unknown(x, yc, 1.0f) * W;
            BP<D> yp = yc;
            Sample<D> q;
            if constexpr (FL == FL_FWD_LINE) q = line_sample<D, STACK>(S, yp.p, wp.skip(yp.elem), rng, stack, stride);
            else q = sample_uniform<D>(S, rng);
            yc = q.y;
            W = m.fredholm_unknown_forward(yc, yp, q.inv_pdf, rng) * W;
        }
        return res;
    }
}

}  // namespace wobdev
