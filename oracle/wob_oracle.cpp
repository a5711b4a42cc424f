// CPU oracle for the walk-on-boundary hot path.
//
// TEST INFRASTRUCTURE ONLY.  Nothing in the product (wobtoolbox_amd/) links, loads or
// calls this file.  It is used by tests/ as the checker, by __graft_entry__.smoke() as the
// checker, and by bench.py's cpu_baseline leg (OpenMP, one thread per evaluation point,
// sequential paths: the Thrust-OMP execution model of the reference's CPU build).
//
// It restates, in FP32 with the reference's operation order, the following reference code
// (paths relative to the reference repository root):
//   wob/wob/equation_kernel.cuh:67-145   Laplace fundamental solutions (2D, 3D)
//   wob/wob/bie_models.cuh:53-588        BIE models: single layer, double layer, direct
//   wob/wob/scene.cuh:30-43              binary_search over a CDF
//   wob/wob/scene.cuh:45-69              barycentric coordinates (igl formula)
//   wob/wob/scene.cuh:78-120             grid volume cache sampling
//   wob/wob/scene.cuh:264-299            area / boundary-value CDFs with a Kahan sum
//   wob/wob/scene.cuh:378-436, 646-730   segment / triangle elements and line intersection
//   wob/wob/scene.cuh:441-622, 734-918   boundary samplers (uniform, line intersection,
//                                        resampled importance sampling, boundary value)
//   wob/wob/estimator.cuh:113-504        backward and forward estimators (the walk)
//   wob/wob/utils.cuh:148-166            Kahan summation
//   wob/wob/runtime_utils.cuh:89-103     per-point mt19937 seeding (reference RNG mode)
// plus an FP64 twin of the kernels (closed forms) and a brute-force all-hit line query.
//
// Deviations that cannot be pinned against the reference (its lbvh, curand and Eigen
// sources are absent) are explicit options:
//   pick_mode 0 ("reference"): hits are collected in BVH traversal order into a 64-entry
//     buffer and buf[min(N-1, floor(N*u))] is taken, u drawn after the query only when
//     N > 0, count clamped to 64 (scene.cuh:481-502).
//   pick_mode 1 ("matched"): the hit with the smallest triple32(pick ^ id*0x9E3779B9)
//     is taken, where pick is a 32-bit draw made before the query; this rule does not
//     depend on traversal order, so the GPU path can be compared path by path.
//   rng_mode 0: Philox4x32-10 (Salmon et al., SC'11) with key = seed and counter =
//     (block, path, point, stream).  rng_mode 1: per-point std::mt19937 seeded from a
//     default std::mt19937_64 as in runtime_utils.cuh:89-103 (CPU baseline only).
// The CDFs are kept in the caller's element order (the reference keeps BVH leaf order);
// the sampled distribution is the same.

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <random>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr float kFltEps = std::numeric_limits<float>::epsilon();
constexpr float kPi = (float)M_PI;

// ------------------------------------------------------------------------------------
// Philox4x32-10 (Salmon, Moraes, Dror, Shaw: "Parallel Random Numbers: As Easy as 1, 2, 3")
// ------------------------------------------------------------------------------------
static inline void philox4x32_10(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
    uint32_t c0 = ctr_in[0], c1 = ctr_in[1], c2 = ctr_in[2], c3 = ctr_in[3];
    uint32_t k0 = key_in[0], k1 = key_in[1];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        const uint32_t n0 = hi1 ^ c1 ^ k0;
        const uint32_t n1 = lo1;
        const uint32_t n2 = hi0 ^ c3 ^ k1;
        const uint32_t n3 = lo0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

// 24-bit uniform in (0, 1]: the reference's device uniform is curand_uniform, also (0, 1].
static inline float u32_to_unit(uint32_t x) { return (float)((x >> 8) + 1u) * (1.0f / 16777216.0f); }

static inline uint32_t triple32(uint32_t x) {
    x ^= x >> 17; x *= 0xED5AD4BBu;
    x ^= x >> 11; x *= 0xAC4C1B51u;
    x ^= x >> 15; x *= 0x31848BABu;
    x ^= x >> 14;
    return x;
}

struct Rng {
    int mode = 0;
    // philox stream
    uint32_t key[2]{0, 0};
    uint32_t ctr[4]{0, 0, 0, 0};
    uint32_t buf[4]{0, 0, 0, 0};
    int idx = 4;
    // reference mode
    std::mt19937 mt;
    // mode 2: a fixed value (deterministic weight evaluation)
    float fixed_u = 0.5f;

    void init_philox(uint64_t seed, uint32_t path, uint32_t point, uint32_t stream) {
        mode = 0;
        key[0] = (uint32_t)seed; key[1] = (uint32_t)(seed >> 32);
        ctr[0] = 0; ctr[1] = path; ctr[2] = point; ctr[3] = stream;
        idx = 4;
    }
    uint32_t next_u32() {
        if (mode == 1) return (uint32_t)mt();
        if (idx == 4) {
            philox4x32_10(ctr, key, buf);
            ctr[0] += 1;
            idx = 0;
        }
        return buf[idx++];
    }
    // utils.cuh:51-56 (CPU build: 1 - U[0,1)) in reference mode, otherwise (0, 1] from Philox.
    float uniform() {
        if (mode == 1) return 1.0f - std::uniform_real_distribution<float>()(mt);
        if (mode == 2) return fixed_u;
        return u32_to_unit(next_u32());
    }
};

// ------------------------------------------------------------------------------------
// Small fixed-size vectors with Eigen's evaluation order (component-wise ops; dot is a
// left-to-right sum; norm = sqrt(squaredNorm); normalized = v / norm when norm > 0).
// ------------------------------------------------------------------------------------
template <int D, typename S = float> struct Vec {
    S v[D];
    S &operator[](int i) { return v[i]; }
    const S &operator[](int i) const { return v[i]; }
    static Vec zero() { Vec r; for (int i = 0; i < D; ++i) r.v[i] = 0; return r; }
};
template <int D, typename S> Vec<D, S> operator+(const Vec<D, S> &a, const Vec<D, S> &b) { Vec<D, S> r; for (int i = 0; i < D; ++i) r[i] = a[i] + b[i]; return r; }
template <int D, typename S> Vec<D, S> operator-(const Vec<D, S> &a, const Vec<D, S> &b) { Vec<D, S> r; for (int i = 0; i < D; ++i) r[i] = a[i] - b[i]; return r; }
template <int D, typename S> Vec<D, S> operator*(S s, const Vec<D, S> &a) { Vec<D, S> r; for (int i = 0; i < D; ++i) r[i] = s * a[i]; return r; }
template <int D, typename S> Vec<D, S> operator*(const Vec<D, S> &a, S s) { Vec<D, S> r; for (int i = 0; i < D; ++i) r[i] = a[i] * s; return r; }
template <int D, typename S> Vec<D, S> operator/(const Vec<D, S> &a, S s) { Vec<D, S> r; for (int i = 0; i < D; ++i) r[i] = a[i] / s; return r; }
template <int D, typename S> S dot(const Vec<D, S> &a, const Vec<D, S> &b) { S s = a[0] * b[0]; for (int i = 1; i < D; ++i) s = s + a[i] * b[i]; return s; }
template <int D, typename S> S sqnorm(const Vec<D, S> &a) { return dot(a, a); }
template <int D, typename S> S norm(const Vec<D, S> &a) { return std::sqrt(sqnorm(a)); }
template <int D, typename S> Vec<D, S> normalized(const Vec<D, S> &a) {
    S z = sqnorm(a);
    if (z > S(0)) return a / (S)std::sqrt(z);
    return a;
}
template <typename S> Vec<3, S> cross(const Vec<3, S> &a, const Vec<3, S> &b) {
    Vec<3, S> r;
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
    return r;
}

enum BoundaryType { Dirichlet = 0, Neumann = 1, Robin = 2 };
enum BIEModelType { IndirectSingleLayer = 0, IndirectDoubleLayer = 1, Direct = 2 };
enum DomainType { InteriorDomain = 0, ExteriorDomain = 1 };
enum EstimatorType { BackwardEstimator = 0, ForwardEstimator = 1 };
enum ProblemType { DirichletProblem = 0, NeumannProblem = 1, RobinProblem = 2, MixedProblem = 3 };

// points.hpp:32-39
template <int D> struct BPoint {
    Vec<D> p, n;
    float interior_angle;
    float boundary_value;
    int boundary_type;
    float robin_alpha;
    int elem = -1;   // element the point lies on (-1: none / unknown); used only by exclude_origin
    static BPoint zero() { BPoint b; b.p = Vec<D>::zero(); b.n = Vec<D>::zero(); b.interior_angle = 0; b.boundary_value = 0; b.boundary_type = Dirichlet; b.robin_alpha = 0; b.elem = -1; return b; }
};

// ------------------------------------------------------------------------------------
// equation_kernel.cuh:67-145 (FP32, reference operation order)
// ------------------------------------------------------------------------------------
template <int D> struct Kernel;
template <> struct Kernel<2> {
    float eps;
    float G(const Vec<2> &y, const Vec<2> &x) const {                                    // :74-76
        return -1.0f / (2.0f * kPi) * std::log(std::max(norm(y - x), eps));
    }
    Vec<2> dG_dx(const Vec<2> &y, const Vec<2> &x) const {                               // :78-81
        const Vec<2> r = y - x;
        return (1.0f / (2.0f * kPi) * normalized(r)) / std::max(norm(r), eps);
    }
    float dG_dny(const Vec<2> &y, const Vec<2> &x, const Vec<2> &ny) const {             // :83-86
        const Vec<2> r = y - x;
        return -1.0f / (2.0f * kPi) * dot(normalized(r), ny) / std::max(norm(r), eps);
    }
    Vec<2> d2G_dxdny(const Vec<2> &y, const Vec<2> &x, const Vec<2> &ny) const {         // :88-92
        const Vec<2> r = y - x;
        const Vec<2> rh = normalized(r);
        return 1.0f / (2.0f * kPi * std::max(sqnorm(r), eps)) * (ny - (2.0f * dot(rh, ny)) * rh);
    }
    float d2G_dnxdny(const Vec<2> &y, const Vec<2> &x, const Vec<2> &ny, const Vec<2> &nx) const { // :94-100
        const Vec<2> r = y - x;
        const Vec<2> rh = normalized(r);
        return 1.0f / (2.0f * kPi * std::max(sqnorm(r), eps)) * (dot(nx, ny) - 2.0f * dot(rh, nx) * dot(rh, ny));
    }
};
template <> struct Kernel<3> {
    float eps;
    float G(const Vec<3> &y, const Vec<3> &x) const {                                    // :115-117
        return 1.0f / (4.0f * kPi * std::max(norm(y - x), eps));
    }
    Vec<3> dG_dx(const Vec<3> &y, const Vec<3> &x) const {                               // :119-122
        const Vec<3> r = y - x;
        return (1.0f / (4.0f * kPi) * normalized(r)) / std::max(sqnorm(r), eps);
    }
    float dG_dny(const Vec<3> &y, const Vec<3> &x, const Vec<3> &ny) const {             // :124-127
        const Vec<3> r = y - x;
        return -1.0f / (4.0f * kPi) * dot(normalized(r), ny) / std::max(sqnorm(r), eps);
    }
    Vec<3> d2G_dxdny(const Vec<3> &y, const Vec<3> &x, const Vec<3> &ny) const {         // :129-134
        const Vec<3> r = y - x;
        const Vec<3> rh = normalized(r);
        return 1.0f / (4.0f * kPi * std::max(norm(r) * sqnorm(r), eps)) * (ny - (3.0f * dot(rh, ny)) * rh);
    }
    float d2G_dnxdny(const Vec<3> &y, const Vec<3> &x, const Vec<3> &ny, const Vec<3> &nx) const { // :136-142
        const Vec<3> r = y - x;
        const Vec<3> rh = normalized(r);
        return 1.0f / (4.0f * kPi * std::max(norm(r) * sqnorm(r), eps)) * (dot(nx, ny) - 3.0f * dot(rh, nx) * dot(rh, ny));
    }
};

// FP64 closed forms (textbook Laplace Green's functions), same eps clamping as the reference.
template <int D> struct Kernel64;
template <> struct Kernel64<2> {
    double eps;
    double G(const Vec<2, double> &y, const Vec<2, double> &x) const { return -std::log(std::max(norm(y - x), eps)) / (2.0 * M_PI); }
    Vec<2, double> dG_dx(const Vec<2, double> &y, const Vec<2, double> &x) const { auto r = y - x; return (1.0 / (2.0 * M_PI * std::max(norm(r), eps))) * normalized(r); }
    double dG_dny(const Vec<2, double> &y, const Vec<2, double> &x, const Vec<2, double> &ny) const { auto r = y - x; return -dot(normalized(r), ny) / (2.0 * M_PI * std::max(norm(r), eps)); }
    Vec<2, double> d2G_dxdny(const Vec<2, double> &y, const Vec<2, double> &x, const Vec<2, double> &ny) const { auto r = y - x; auto rh = normalized(r); return (1.0 / (2.0 * M_PI * std::max(sqnorm(r), eps))) * (ny - (2.0 * dot(rh, ny)) * rh); }
    double d2G_dnxdny(const Vec<2, double> &y, const Vec<2, double> &x, const Vec<2, double> &ny, const Vec<2, double> &nx) const { auto r = y - x; auto rh = normalized(r); return (dot(nx, ny) - 2.0 * dot(rh, nx) * dot(rh, ny)) / (2.0 * M_PI * std::max(sqnorm(r), eps)); }
};
template <> struct Kernel64<3> {
    double eps;
    double G(const Vec<3, double> &y, const Vec<3, double> &x) const { return 1.0 / (4.0 * M_PI * std::max(norm(y - x), eps)); }
    Vec<3, double> dG_dx(const Vec<3, double> &y, const Vec<3, double> &x) const { auto r = y - x; return (1.0 / (4.0 * M_PI * std::max(sqnorm(r), eps))) * normalized(r); }
    double dG_dny(const Vec<3, double> &y, const Vec<3, double> &x, const Vec<3, double> &ny) const { auto r = y - x; return -dot(normalized(r), ny) / (4.0 * M_PI * std::max(sqnorm(r), eps)); }
    Vec<3, double> d2G_dxdny(const Vec<3, double> &y, const Vec<3, double> &x, const Vec<3, double> &ny) const { auto r = y - x; auto rh = normalized(r); return (1.0 / (4.0 * M_PI * std::max(norm(r) * sqnorm(r), eps))) * (ny - (3.0 * dot(rh, ny)) * rh); }
    double d2G_dnxdny(const Vec<3, double> &y, const Vec<3, double> &x, const Vec<3, double> &ny, const Vec<3, double> &nx) const { auto r = y - x; auto rh = normalized(r); return (dot(nx, ny) - 3.0 * dot(rh, nx) * dot(rh, ny)) / (4.0 * M_PI * std::max(norm(r) * sqnorm(r), eps)); }
};

// ------------------------------------------------------------------------------------
// Elements (scene.cuh:378-397, 646-678) and the scene (CDFs, BVH)
// ------------------------------------------------------------------------------------
template <int D> struct Element {
    Vec<D> v[D];           // a, b (, c)
    Vec<D> n;
    float bv[D];
    int type;
    float alpha[D];
    float area() const {
        if constexpr (D == 2) return norm(v[1] - v[0]);                     // scene.cuh:384
        else return norm(cross(v[2] - v[0], v[1] - v[0])) / 2.0f;           // scene.cuh:652-654
    }
    float boundary_value() const {
        if constexpr (D == 2) return area() * (std::abs(bv[0]) + std::abs(bv[1])) / 2.0f;                        // :386-392
        else return area() * (std::abs(bv[0]) + std::abs(bv[1]) + std::abs(bv[2])) / 3.0f;                       // :656-663
    }
};

template <int D> struct AABB { float lo[D], hi[D]; };

struct VolumeGrid {
    // scene.cuh:78-120 (grid volume cache with value-proportional sampling)
    bool present = false;
    unsigned grid_res = 0;
    float domain_size = 0;
    std::vector<float> cache, cdf;
};

// Source f(x) = sum_j A_j exp(-|x - c_j|^2 / (2 w_j^2)), sampled from the mixture with weights
// proportional to |A_j| w_j^D (no counterpart in the reference, whose only source is the grid).
// With `mask`, samples on the far side of the boundary (outside the domain) contribute 0.
struct VolumeGauss {
    bool present = false;
    bool mask = true;
    unsigned n = 0;
    std::vector<float> c, amp, inv2w2, pn, cdf, w;   // c: n * 3; pn: mixture density coefficient
    // Occupancy grid of the inside test (no counterpart in the reference): res^D cells over the
    // vertex box, each outside (0), inside (1) or near the boundary (2).  A cell is near when it
    // lies within one cell of an element's inflated box; otherwise the rays from two points of
    // the cell (its centre and the point a quarter cell towards its low corner) decide it when
    // their parities agree, and it is near when they do not.  A sample in an outside or inside
    // cell takes the cell's parity (no element meets the cell, so the parity is the same at
    // every point of it); a sample in a near cell, or within one cell outside the grid, casts
    // its own ray; a sample further out is outside.
    unsigned gres = 0;
    float glo[3] = {0.f, 0.f, 0.f}, gh[3] = {0.f, 0.f, 0.f}, ginv[3] = {0.f, 0.f, 0.f};
    std::vector<unsigned char> gflag;
};

template <int D> struct Scene {
    std::vector<Element<D>> elems;
    std::vector<float> area_cdf, bv_cdf;
    float total_area = 0, total_bv = 0;
    // binary LBVH (Karras 2012), internal nodes [0, n-1), leaves [n-1, 2n-1)
    struct Node { int left, right, obj; AABB<D> box; };
    std::vector<Node> nodes;
    VolumeGrid vol;
    VolumeGauss gauss;
    // harmonic control variate u_h(p) = hcv_c0 + sum_j w_j phi(|p - s_j|^2), phi(z) = z^(-1/2)
    // (3D) or log2 z (2D); hcv: m records (s_j, w_j) of D + 1 floats (no counterpart in the
    // reference)
    std::vector<float> hcv;
    float hcv_c0 = 0.f;
    unsigned hcv_m = 0;
};

// u_h at p, terms added in source order after c0 (FP32)
template <int D> static float hcv_value(const Scene<D> &s, const float *p) {
    float acc = s.hcv_c0;
    for (unsigned j = 0; j < s.hcv_m; ++j) {
        const float *r = &s.hcv[(size_t)j * (D + 1)];
        float z = 0.f;
        for (int k = 0; k < D; ++k) {
            const float d = p[k] - r[k];
            z = k == 0 ? d * d : z + d * d;
        }
        acc += r[D] * (D == 3 ? 1.0f / std::sqrt(z) : std::log2(z));
    }
    return acc;
}
// grad u_h: 3D sum_j -w_j (p - s_j) |p - s_j|^-3; 2D (2 / ln 2) sum_j w_j (p - s_j) / |p - s_j|^2
template <int D> static void hcv_gradient(const Scene<D> &s, const float *p, float *g) {
    for (int k = 0; k < D; ++k) g[k] = 0.f;
    for (unsigned j = 0; j < s.hcv_m; ++j) {
        const float *r = &s.hcv[(size_t)j * (D + 1)];
        float d[3], z = 0.f;
        for (int k = 0; k < D; ++k) {
            d[k] = p[k] - r[k];
            z = k == 0 ? d[k] * d[k] : z + d[k] * d[k];
        }
        float f;
        if (D == 3) {
            const float ri = 1.0f / std::sqrt(z);
            f = -r[D] * (ri * ri * ri);
        } else {
            f = r[D] * (1.0f / z);
        }
        for (int k = 0; k < D; ++k) g[k] += f * d[k];
    }
    if (D == 2)
        for (int k = 0; k < D; ++k) g[k] = g[k] * 2.8853900817779268f;
}

// the residual datum at y: y's datum minus u_h's datum of y's type
template <int D> static BPoint<D> hcv_residual(const Scene<D> &s, const BPoint<D> &y) {
    BPoint<D> r = y;
    if (y.boundary_type == Dirichlet) {
        r.boundary_value = y.boundary_value - hcv_value(s, y.p.v);
    } else {
        float g[3];
        hcv_gradient(s, y.p.v, g);
        float dn = g[0] * y.n[0];
        for (int k = 1; k < D; ++k) dn += g[k] * y.n[k];
        if (y.boundary_type == Neumann) r.boundary_value = y.boundary_value - dn;
        else r.boundary_value = y.boundary_value - (dn + y.robin_alpha * hcv_value(s, y.p.v));
    }
    return r;
}

// utils.cuh:148-166
struct Kahan {
    float sum = 0, c = 0;
    void add(float value) { float y = value - c; float t = sum + y; c = (t - sum) - y; sum = t; }
};

// scene.cuh:30-43
static inline unsigned binary_search(const float *a, unsigned n, float key) {
    unsigned low = 0, high = n - 1;
    while (low < high) {
        unsigned mid = (high + low) / 2;
        if (key <= a[mid]) high = mid; else low = mid + 1;
    }
    return high;
}

template <int D> static void build_cdfs(Scene<D> &s) {
    // scene.cuh:271-298 (element order instead of BVH leaf order)
    const size_t n = s.elems.size();
    s.area_cdf.resize(n); s.bv_cdf.resize(n);
    Kahan ka;
    for (size_t i = 0; i < n; ++i) { ka.add(s.elems[i].area()); s.area_cdf[i] = ka.sum; }
    s.total_area = ka.sum;
    for (size_t i = 0; i < n; ++i) s.area_cdf[i] /= s.total_area;
    Kahan kb;
    for (size_t i = 0; i < n; ++i) { kb.add(s.elems[i].boundary_value()); s.bv_cdf[i] = kb.sum; }
    s.total_bv = kb.sum;
    for (size_t i = 0; i < n; ++i) s.bv_cdf[i] /= s.total_bv;
}

// --- Karras LBVH over element centroids (mirrors the structure of the lbvh library the
//     reference uses: Morton codes, sorted leaves, radix-tree hierarchy, bottom-up boxes).
static inline uint32_t expand_bits10(uint32_t v) {
    v = (v * 0x00010001u) & 0xFF0000FFu;
    v = (v * 0x00000101u) & 0x0F00F00Fu;
    v = (v * 0x00000011u) & 0xC30C30C3u;
    v = (v * 0x00000005u) & 0x49249249u;
    return v;
}
static inline uint32_t expand_bits16(uint32_t v) {
    v &= 0xFFFFu;
    v = (v | (v << 8)) & 0x00FF00FFu;
    v = (v | (v << 4)) & 0x0F0F0F0Fu;
    v = (v | (v << 2)) & 0x33333333u;
    v = (v | (v << 1)) & 0x55555555u;
    return v;
}

template <int D> static AABB<D> elem_box(const Element<D> &e) {
    AABB<D> b;
    for (int k = 0; k < D; ++k) {
        float lo = e.v[0][k], hi = e.v[0][k];
        for (int j = 1; j < D; ++j) { lo = std::min(lo, e.v[j][k]); hi = std::max(hi, e.v[j][k]); }
        b.lo[k] = lo; b.hi[k] = hi;
    }
    return b;
}
template <int D> static AABB<D> merge(const AABB<D> &a, const AABB<D> &b) {
    AABB<D> r;
    for (int k = 0; k < D; ++k) { r.lo[k] = std::min(a.lo[k], b.lo[k]); r.hi[k] = std::max(a.hi[k], b.hi[k]); }
    return r;
}

template <int D> static void build_lbvh(Scene<D> &s) {
    const int n = (int)s.elems.size();
    std::vector<AABB<D>> boxes(n);
    AABB<D> world = elem_box(s.elems[0]);
    for (int i = 0; i < n; ++i) { boxes[i] = elem_box(s.elems[i]); world = merge(world, boxes[i]); }
    std::vector<std::pair<uint64_t, int>> keys(n);
    for (int i = 0; i < n; ++i) {
        uint32_t code = 0;
        for (int k = 0; k < D; ++k) {
            float c = 0.5f * (boxes[i].lo[k] + boxes[i].hi[k]);
            float ext = world.hi[k] - world.lo[k];
            float t = ext > 0 ? (c - world.lo[k]) / ext : 0.5f;
            const float scale = D == 3 ? 1024.0f : 65536.0f;
            uint32_t q = (uint32_t)std::min(std::max(t * scale, 0.0f), scale - 1.0f);
            code |= (D == 3 ? expand_bits10(q) : expand_bits16(q)) << (D - 1 - k);
        }
        keys[i] = {((uint64_t)code << 32) | (uint32_t)i, i};
    }
    std::sort(keys.begin(), keys.end());
    s.nodes.assign(std::max(1, 2 * n - 1), {});
    const int leaf0 = n - 1;
    for (int i = 0; i < n; ++i) {
        auto &L = s.nodes[leaf0 + i];
        L.left = L.right = -1; L.obj = keys[i].second; L.box = boxes[keys[i].second];
    }
    if (n == 1) return;
    auto delta = [&](int i, int j) -> int {
        if (j < 0 || j >= n) return -1;
        return __builtin_clzll(keys[i].first ^ keys[j].first);   // keys are unique (index in low bits)
    };
    std::vector<int> parent(2 * n - 1, -1);
    for (int i = 0; i < n - 1; ++i) {
        const int d = (delta(i, i + 1) - delta(i, i - 1)) >= 0 ? 1 : -1;
        const int dmin = delta(i, i - d);
        int lmax = 2;
        while (delta(i, i + lmax * d) > dmin) lmax *= 2;
        int l = 0;
        for (int t = lmax / 2; t >= 1; t /= 2)
            if (delta(i, i + (l + t) * d) > dmin) l += t;
        const int j = i + l * d;
        const int dnode = delta(i, j);
        int s_ = 0;
        int t = l;
        do {
            t = (t + 1) / 2;
            if (delta(i, i + (s_ + t) * d) > dnode) s_ += t;
        } while (t > 1);
        const int gamma = i + s_ * d + std::min(d, 0);
        const int left = (std::min(i, j) == gamma) ? leaf0 + gamma : gamma;
        const int right = (std::max(i, j) == gamma + 1) ? leaf0 + gamma + 1 : gamma + 1;
        s.nodes[i].left = left; s.nodes[i].right = right; s.nodes[i].obj = -1;
        parent[left] = i; parent[right] = i;
    }
    // bottom-up boxes: post-order from the root (iterative)
    std::vector<int> order; order.reserve(n - 1);
    std::vector<int> st{0};
    while (!st.empty()) {
        int v = st.back(); st.pop_back();
        if (v >= leaf0) continue;
        order.push_back(v);
        st.push_back(s.nodes[v].left); st.push_back(s.nodes[v].right);
    }
    for (auto it = order.rbegin(); it != order.rend(); ++it)
        s.nodes[*it].box = merge(s.nodes[s.nodes[*it].left].box, s.nodes[s.nodes[*it].right].box);
}

// Conservative line-vs-box test for an infinite line (box inflated so that the
// primitive test, not the box test, decides every hit).
template <int D> static inline bool line_hits_box(const AABB<D> &b, const Vec<D> &o, const Vec<D> &d) {
    double tmin = -INFINITY, tmax = INFINITY;
    for (int k = 0; k < D; ++k) {
        const double pad = 1e-5 * (std::abs((double)b.lo[k]) + std::abs((double)b.hi[k])) + 1e-6;
        const double lo = b.lo[k] - pad, hi = b.hi[k] + pad;
        if (d[k] == 0.0f) {
            if (o[k] < lo || o[k] > hi) return false;
            continue;
        }
        double t1 = (lo - o[k]) / d[k], t2 = (hi - o[k]) / d[k];
        if (t1 > t2) std::swap(t1, t2);
        tmin = std::max(tmin, t1); tmax = std::min(tmax, t2);
        if (tmin > tmax) return false;
    }
    return true;
}

// scene.cuh:414-436 (2D) / 707-730 (3D): solve with the adjugate inverse, accept with the
// reference's closed-interval tests and |t| > 10*FLT_EPSILON.
static inline bool intersect(const Vec<2> &o, const Vec<2> &d, const Element<2> &e, Vec<2> &p, float &t_out) {
    constexpr float eps = 10 * kFltEps;
    // mat = [d, a - b]; ta = inverse(mat) * (a - o)
    const Vec<2> c1 = e.v[0] - e.v[1];
    const float m00 = d[0], m10 = d[1], m01 = c1[0], m11 = c1[1];
    const float det = m00 * m11 - m10 * m01;
    const float invdet = 1.0f / det;
    const float i00 = m11 * invdet, i10 = -m10 * invdet, i01 = -m01 * invdet, i11 = m00 * invdet;
    const Vec<2> rhs = e.v[0] - o;
    const float t = i00 * rhs[0] + i01 * rhs[1];
    const float a = i10 * rhs[0] + i11 * rhs[1];
    if (a < 0.0f || a > 1.0f || std::abs(t) <= eps || std::isnan(t) || std::isnan(a)) return false;
    p = o + t * d;
    t_out = t;
    return true;
}
static inline bool intersect(const Vec<3> &o, const Vec<3> &d, const Element<3> &e, Vec<3> &p, float &t_out) {
    constexpr float eps = 10 * kFltEps;
    // mat = [d, a - b, a - c]; tab = inverse(mat) * (a - o); inverse = adjugate / det
    float m[3][3];
    const Vec<3> c1 = e.v[0] - e.v[1], c2 = e.v[0] - e.v[2];
    for (int r = 0; r < 3; ++r) { m[r][0] = d[r]; m[r][1] = c1[r]; m[r][2] = c2[r]; }
    auto cof = [&](int i, int j) {
        const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
        return m[i1][j1] * m[i2][j2] - m[i1][j2] * m[i2][j1];
    };
    const float c00 = cof(0, 0), c10 = cof(1, 0), c20 = cof(2, 0);
    const float det = c00 * m[0][0] + c10 * m[1][0] + c20 * m[2][0];
    const float invdet = 1.0f / det;
    float inv[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) inv[i][j] = cof(j, i) * invdet;
    const Vec<3> rhs = e.v[0] - o;
    float tab[3];
    for (int i = 0; i < 3; ++i) tab[i] = inv[i][0] * rhs[0] + inv[i][1] * rhs[1] + inv[i][2] * rhs[2];
    if (tab[1] < 0.0f || tab[2] < 0.0f || tab[1] + tab[2] > 1.0f || std::abs(tab[0]) <= eps ||
        std::isnan(tab[0]) || std::isnan(tab[1]) || std::isnan(tab[2]))
        return false;
    p = o + tab[0] * d;
    t_out = tab[0];
    return true;
}

template <int D> struct Hit { int obj; Vec<D> p; float t; };

// All hits of the infinite line o + t d.  use_bvh=0: brute force over every element.
template <int D>
static unsigned collect_hits(const Scene<D> &s, const Vec<D> &o, const Vec<D> &d, bool use_bvh, Hit<D> *buf,
                             unsigned cap, int skip = -1) {
    unsigned count = 0;
    auto test = [&](int obj) {
        if (obj == skip) return;
        Vec<D> p; float t;
        if (intersect(o, d, s.elems[obj], p, t)) {
            if (count < cap) buf[count] = {obj, p, t};
            ++count;
        }
    };
    const int n = (int)s.elems.size();
    if (!use_bvh) {
        for (int i = 0; i < n; ++i) test(i);
        return count;
    }
    if (n == 1) { test(0); return count; }
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    const int leaf0 = n - 1;
    while (sp > 0) {
        const int v = stack[--sp];
        const auto &N = s.nodes[v];
        const int ch[2] = {N.left, N.right};
        for (int c = 0; c < 2; ++c) {
            const int u = ch[c];
            if (!line_hits_box(s.nodes[u].box, o, d)) continue;
            if (u >= leaf0) test(s.nodes[u].obj);
            else if (sp < 64) stack[sp++] = u;
        }
    }
    return count;
}

// ------------------------------------------------------------------------------------
// Samplers (scene.cuh:441-622 for 2D, 734-918 for 3D)
// ------------------------------------------------------------------------------------
template <int D> struct Sampled { BPoint<D> y; float inv_pdf; };

struct Options {
    int pick_mode = 1;      // 0: reference buffer pick, 1: matched min-hash pick
    bool use_bvh = true;
    bool exclude_origin = false;   // ignore the element the line's origin lies on
    bool harmonic_cv = false;      // single/double-layer walks: the residual of the scene's harmonic u_h
};

template <int D> static BPoint<D> point_on_element_uniform(const Element<D> &e, Rng &rng) {
    BPoint<D> r;
    r.elem = -1;   // the callers set the element index
    r.interior_angle = 0.5f;
    r.boundary_type = e.type;
    r.n = e.n;
    if constexpr (D == 2) {
        const float u2 = rng.uniform();
        for (int k = 0; k < 2; ++k) r.p[k] = e.v[0][k] + u2 * (e.v[1][k] - e.v[0][k]);
        r.boundary_value = e.bv[0] + u2 * (e.bv[1] - e.bv[0]);
        r.robin_alpha = e.alpha[0] + u2 * (e.alpha[1] - e.alpha[0]);
    } else {
        const float u2 = rng.uniform();
        const float u3 = rng.uniform();
        const float s = std::sqrt(u2);
        const float w0 = 1.0f - s, w1 = s * (1.0f - u3), w2 = u3 * s;     // scene.cuh:667-671
        for (int k = 0; k < 3; ++k) r.p[k] = w0 * e.v[0][k] + w1 * e.v[1][k] + w2 * e.v[2][k];
        r.boundary_value = w0 * e.bv[0] + w1 * e.bv[1] + w2 * e.bv[2];
        r.robin_alpha = w0 * e.alpha[0] + w1 * e.alpha[1] + w2 * e.alpha[2];
    }
    return r;
}

template <int D> static Sampled<D> sample_uniform(const Scene<D> &s, Rng &rng) {
    const float u1 = rng.uniform();
    const unsigned idx = binary_search(s.area_cdf.data(), (unsigned)s.elems.size(), u1);
    Sampled<D> r{point_on_element_uniform(s.elems[idx], rng), s.total_area};
    r.y.elem = (int)idx;
    return r;
}

template <int D> static Sampled<D> sample_boundary_value(const Scene<D> &s, Rng &rng) {
    const float u1 = rng.uniform();
    const unsigned idx = binary_search(s.bv_cdf.data(), (unsigned)s.elems.size(), u1);
    const Element<D> &e = s.elems[idx];
    if constexpr (D == 2) {
        // scene.cuh:596-621
        const float u2 = rng.uniform();
        const float ba = std::abs(e.bv[0]), bb = std::abs(e.bv[1]);
        float interp;
        if (std::abs(ba - bb) < 10 * kFltEps) interp = u2;
        else interp = (-ba + std::sqrt(ba * ba + u2 * (bb * bb - ba * ba))) / (bb - ba);
        BPoint<2> r;
        for (int k = 0; k < 2; ++k) r.p[k] = e.v[0][k] + interp * (e.v[1][k] - e.v[0][k]);
        r.n = e.n; r.interior_angle = 0.5f;
        r.boundary_value = e.bv[0] + interp * (e.bv[1] - e.bv[0]);
        r.boundary_type = e.type;
        r.robin_alpha = e.alpha[0] + interp * (e.alpha[1] - e.alpha[0]);
        r.elem = (int)idx;
        return {r, s.total_bv / (ba + interp * (bb - ba))};
    } else {
        // scene.cuh:896-917
        const float pmf = idx == 0 ? s.bv_cdf[idx] : s.bv_cdf[idx] - s.bv_cdf[idx - 1];
        BPoint<3> r = point_on_element_uniform(e, rng);
        r.elem = (int)idx;
        return {r, e.area() / pmf};
    }
}

// scene.cuh:464-535 (2D) and 760-840 (3D)
template <int D> static Sampled<D> sample_line_intersection(const Scene<D> &s, const Vec<D> &x, Rng &rng, const Options &opt,
                                                            int origin_elem = -1, unsigned *count_out = nullptr) {
    const int skip = opt.exclude_origin ? origin_elem : -1;
    Vec<D> dir;
    if constexpr (D == 2) {
        const float theta = rng.uniform();
        dir[0] = (float)std::cos(theta * M_PI);
        dir[1] = (float)std::sin(theta * M_PI);
    } else {
        dir[2] = rng.uniform();
        const float phi = 2.f * rng.uniform();
        dir[0] = (float)std::cos(phi * M_PI);
        dir[1] = (float)std::sin(phi * M_PI);
        const float r = std::sqrt(1.f - dir[2] * dir[2]);
        dir[0] *= r; dir[1] *= r;
    }
    uint32_t pick_bits = 0;
    if (opt.pick_mode == 1) pick_bits = rng.next_u32();

    constexpr unsigned kBuf = 64;
    Hit<D> buf[kBuf];
    int obj = -1;
    Vec<D> p;
    unsigned count;
    if (opt.pick_mode == 1) {
        // order-independent uniform pick: smallest hash of the element id
        Hit<D> all[1024];
        count = collect_hits(s, x, dir, opt.use_bvh, all, 1024, skip);
        uint32_t best = 0xFFFFFFFFu;
        for (unsigned i = 0; i < std::min(count, 1024u); ++i) {
            const uint32_t h = triple32(pick_bits ^ ((uint32_t)all[i].obj * 0x9E3779B9u));
            if (obj < 0 || h < best || (h == best && all[i].obj < obj)) { best = h; obj = all[i].obj; p = all[i].p; }
        }
    } else {
        count = collect_hits(s, x, dir, opt.use_bvh, buf, kBuf, skip);
        if (count > 0) {
            count = std::min(count, kBuf);
            const unsigned k = std::min(count - 1, (unsigned)(count * rng.uniform()));
            obj = buf[k].obj; p = buf[k].p;
        }
    }
    if (count_out) *count_out = count;
    if (count == 0) {
        // scene.cuh:488-495: default point, zero inverse pdf
        BPoint<D> r = BPoint<D>::zero();
        r.interior_angle = 0.5f;
        Sampled<D> out{r, 0.0f};
        // inv_pdf = 0 * ... (scene.cuh:531-533 with count 0)
        return out;
    }
    const Element<D> &e = s.elems[obj];
    BPoint<D> r;
    r.p = p; r.n = e.n; r.interior_angle = 0.5f; r.boundary_type = e.type; r.elem = obj;
    if constexpr (D == 2) {
        const float interp = norm(p - e.v[0]) / e.area();                        // scene.cuh:505
        r.boundary_value = e.bv[0] + interp * (e.bv[1] - e.bv[0]);
        r.robin_alpha = e.alpha[0] + interp * (e.alpha[1] - e.alpha[0]);
    } else {
        // barycentric_coordinates (scene.cuh:45-69, igl formula)
        const Vec<3> v0 = e.v[1] - e.v[0], v1 = e.v[2] - e.v[0], v2 = p - e.v[0];
        const float d00 = v0[0] * v0[0] + v0[1] * v0[1] + v0[2] * v0[2];
        const float d01 = v0[0] * v1[0] + v0[1] * v1[1] + v0[2] * v1[2];
        const float d11 = v1[0] * v1[0] + v1[1] * v1[1] + v1[2] * v1[2];
        const float d20 = v2[0] * v0[0] + v2[1] * v0[1] + v2[2] * v0[2];
        const float d21 = v2[0] * v1[0] + v2[1] * v1[1] + v2[2] * v1[2];
        const float denom = d00 * d11 - d01 * d01;
        const float l1 = (d11 * d20 - d01 * d21) / denom;
        const float l2 = (d00 * d21 - d01 * d20) / denom;
        const float l0 = 1.0f - (l1 + l2);
        r.boundary_value = l0 * e.bv[0] + l1 * e.bv[1] + l2 * e.bv[2];
        r.robin_alpha = l0 * e.alpha[0] + l1 * e.alpha[1] + l2 * e.alpha[2];
    }
    constexpr float eps = 10 * kFltEps;
    const Vec<D> rv = r.p - x;
    float inv_pdf;
    if constexpr (D == 2)
        inv_pdf = count * kPi * norm(rv) / std::max(std::abs(dot(normalized(rv), r.n)), eps);          // :532-533
    else
        inv_pdf = count * 2.f * kPi * sqnorm(rv) / std::max(std::abs(dot(normalized(rv), r.n)), eps);  // :837-838
    return {r, inv_pdf};
}

// ------------------------------------------------------------------------------------
// Volume sources
// ------------------------------------------------------------------------------------
struct VolSample { bool valid; float p[3]; float value; float inv_pdf; };

static inline void idx_to_domain_point(unsigned idx, unsigned grid_res, float domain_size, int D, float *out) {
    // utils.cuh:92-99
    const float dx = domain_size / grid_res;
    unsigned u[3] = {idx % grid_res, (idx / grid_res) % grid_res, idx / (grid_res * grid_res)};
    if (D == 2) u[1] = idx / grid_res;
    for (int k = 0; k < D; ++k) out[k] = ((float)u[k] + 0.5f) * dx - (domain_size / 2);
}

// Cells per axis of the inside test's occupancy grid.
constexpr unsigned kInsideGridRes2 = 512, kInsideGridRes3 = 64;

// Fixed direction of the inside/outside ray (parity of the crossings with t > 0).
static const float kRay2[2] = {0.58778525f, 0.80901699f};
static const float kRay3[3] = {0.48218104f, 0.66748405f, 0.56749094f};

template <int D> static unsigned ray_crossings(const Scene<D> &s, const Vec<D> &o) {
    Vec<D> d;
    for (int k = 0; k < D; ++k) d[k] = D == 2 ? kRay2[k] : kRay3[k];
    unsigned count = 0;
    auto test = [&](int obj) {
        Vec<D> p; float t;
        if (intersect(o, d, s.elems[obj], p, t) && t > 0.f) ++count;
    };
    const int n = (int)s.elems.size();
    if (n == 1) { test(0); return count; }
    int stack[64];
    int sp = 0;
    stack[sp++] = 0;
    const int leaf0 = n - 1;
    while (sp > 0) {
        const auto &N = s.nodes[stack[--sp]];
        const int ch[2] = {N.left, N.right};
        for (int c = 0; c < 2; ++c) {
            const int u = ch[c];
            if (!line_hits_box(s.nodes[u].box, o, d)) continue;
            if (u >= leaf0) test(s.nodes[u].obj);
            else if (sp < 64) stack[sp++] = u;
        }
    }
    return count;
}

// The inside-test grid's verdict at p: 0 outside, 1 inside, 2 unknown (the caller casts p's ray).
template <int D> static int inside_grid_cell(const VolumeGauss &g, const float *p) {
    if (g.gres == 0) return 2;
    const float R = (float)g.gres;
    size_t idx = 0, mul = 1;
    bool out = false, edge = false;
    for (int k = 0; k < D; ++k) {
        const float sk = (p[k] - g.glo[k]) * g.ginv[k];
        if (!(sk >= -1.f && sk < R + 1.f)) out = true;
        else if (!(sk >= 0.f && sk < R)) edge = true;
        else { idx += (size_t)sk * mul; mul *= g.gres; }
    }
    if (out) return 0;
    if (edge) return 2;
    return g.gflag[idx];
}

template <int D> static void build_inside_grid(const Scene<D> &s, VolumeGauss &g, unsigned res) {
    g.gres = 0;
    g.gflag.clear();
    float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (const auto &e : s.elems)
        for (int j = 0; j < D; ++j)
            for (int k = 0; k < D; ++k) { lo[k] = std::min(lo[k], e.v[j][k]); hi[k] = std::max(hi[k], e.v[j][k]); }
    for (int k = 0; k < D; ++k)
        if (!(hi[k] > lo[k])) return;
    for (int k = 0; k < D; ++k) {
        g.glo[k] = lo[k];
        g.gh[k] = (hi[k] - lo[k]) / (float)res;
        g.ginv[k] = (float)res / (hi[k] - lo[k]);
    }
    size_t n = 1;
    for (int k = 0; k < D; ++k) n *= res;
    g.gres = res;
    g.gflag.assign(n, 255);
    for (const auto &e : s.elems) {
        int i0[3] = {0, 0, 0}, i1[3] = {0, 0, 0};
        for (int k = 0; k < D; ++k) {
            float bl = e.v[0][k], bh = e.v[0][k];
            for (int j = 1; j < D; ++j) { bl = std::min(bl, e.v[j][k]); bh = std::max(bh, e.v[j][k]); }
            const float pad = std::max(std::abs(bl), std::abs(bh)) * 4.0e-6f + 1.0e-7f;
            bl -= pad;
            bh += pad;
            i0[k] = std::max(0, (int)std::floor((bl - g.glo[k]) * g.ginv[k]) - 1);
            i1[k] = std::min((int)res - 1, (int)std::floor((bh - g.glo[k]) * g.ginv[k]) + 1);
        }
        const int z0 = D == 3 ? i0[2] : 0, z1 = D == 3 ? i1[2] : 0;
        for (int z = z0; z <= z1; ++z)
            for (int y = i0[1]; y <= i1[1]; ++y)
                for (int x = i0[0]; x <= i1[0]; ++x) g.gflag[(size_t)x + ((size_t)y + (size_t)z * res) * res] = 2;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (long long ci = 0; ci < (long long)n; ++ci) {
        if (g.gflag[(size_t)ci] == 2) continue;
        size_t rem = (size_t)ci;
        Vec<D> pc, pq;
        for (int k = 0; k < D; ++k) {
            const float ik = (float)(rem % res);
            rem /= res;
            pc[k] = g.glo[k] + (ik + 0.5f) * g.gh[k];
            pq[k] = g.glo[k] + (ik + 0.25f) * g.gh[k];
        }
        const unsigned a = ray_crossings(s, pc) & 1u, b = ray_crossings(s, pq) & 1u;
        g.gflag[(size_t)ci] = (unsigned char)(a == b ? a : 2u);
    }
}

template <int D> static VolSample sample_volume(const Scene<D> &s, Rng &rng, bool interior) {
    VolSample r{false, {0, 0, 0}, 0, 0};
    if (s.vol.present) {
        // scene.cuh:100-120 (sample_volme_cache_cache_value_sampling)
        const auto &v = s.vol;
        unsigned cells = 1;
        for (int k = 0; k < D; ++k) cells *= v.grid_res;
        const float unif = rng.uniform();
        const unsigned idx = binary_search(v.cdf.data(), cells, unif);
        idx_to_domain_point(idx, v.grid_res, v.domain_size, D, r.p);
        for (int k = 0; k < D; ++k) {
            const float u = rng.uniform();
            r.p[k] += (u - 0.5f) * v.domain_size / v.grid_res;
        }
        r.value = v.cache[idx];
        const float pdf = idx == 0 ? v.cdf[idx] : v.cdf[idx] - v.cdf[idx - 1];
        float cellv = 1.f;
        for (int k = 0; k < D; ++k) cellv *= v.domain_size / v.grid_res;
        r.inv_pdf = cellv / pdf;
        r.valid = true;
        return r;
    }
    if (!s.gauss.present) return r;
    const auto &g = s.gauss;
    const float u = rng.uniform();
    unsigned j = 0;
    while (j + 1 < g.n && u > g.cdf[j]) ++j;
    // Box-Muller normals from (0, 1] uniforms
    const float u1 = rng.uniform(), u2 = rng.uniform();
    const float r1 = std::sqrt(-2.f * std::log(u1));
    float z[3];
    z[0] = r1 * (float)std::cos(2.0 * M_PI * (double)u2);
    z[1] = r1 * (float)std::sin(2.0 * M_PI * (double)u2);
    if (D == 3) {
        const float u3 = rng.uniform(), u4 = rng.uniform();
        z[2] = std::sqrt(-2.f * std::log(u3)) * (float)std::cos(2.0 * M_PI * (double)u4);
    }
    for (int k = 0; k < D; ++k) r.p[k] = g.c[3 * j + k] + g.w[j] * z[k];
    float f = 0.f, pdf = 0.f;
    for (unsigned i = 0; i < g.n; ++i) {
        float d2 = 0.f;
        for (int k = 0; k < D; ++k) { const float dk = r.p[k] - g.c[3 * i + k]; d2 += dk * dk; }
        const float e = std::exp(-d2 * g.inv2w2[i]);
        f += g.amp[i] * e;
        pdf += g.pn[i] * e;
    }
    r.value = f;
    r.inv_pdf = 1.0f / pdf;
    if (g.mask) {
        int cell = inside_grid_cell<D>(g, r.p);
        if (cell == 2) {
            Vec<D> o;
            for (int k = 0; k < D; ++k) o[k] = r.p[k];
            cell = (int)(ray_crossings(s, o) & 1u);
        }
        if ((cell != 0) != interior) r.value = 0.f;
    }
    r.valid = true;
    return r;
}

// ------------------------------------------------------------------------------------
// BIE models (bie_models.cuh)
// ------------------------------------------------------------------------------------
template <int D> struct Model {
    int bie;            // BIEModelType
    float dsign;        // domain_type_coefficient: +1 interior, -1 exterior  (:66-71)
    bool interior;
    Kernel<D> k;
    float c;            // first_kind_equation_scaling_constant
    unsigned nvs;       // num_volume_samples
    bool indep = false; // draw num_volume_samples independent samples (the reference reuses one)
    const Scene<D> *scene;

    float ift(float ia) const { return interior ? ia : 1.f - ia; }   // :73-78

    Vec<D> vp(const VolSample &v) const { Vec<D> r; for (int k = 0; k < D; ++k) r[k] = v.p[k]; return r; }

    // :79-103 (one sample reused num_volume_samples times)
    float vol_term(const BPoint<D> &bp, Rng &rng) const {
        if (nvs == 0) return 0.f;
        VolSample v = sample_volume(*scene, rng, interior);
        float res = 0.f;
        Vec<D> q = vp(v);
        for (unsigned i = 0; i < nvs; ++i) {
            if (indep && i > 0) { v = sample_volume(*scene, rng, interior); q = vp(v); }
            if (bp.boundary_type == Dirichlet) res += k.G(bp.p, q) * v.value * v.inv_pdf;
            else if (bp.boundary_type == Neumann) res += k.dG_dny(bp.p, q, bp.n) * v.value * v.inv_pdf;
            else res += (k.dG_dny(bp.p, q, bp.n) + bp.robin_alpha * k.G(q, bp.p)) * v.value * v.inv_pdf;
        }
        return res / nvs;
    }
    // :105-119 and :121-135
    float vol_init(const Vec<D> &x, bool boundary_form, Rng &rng) const {
        if (nvs == 0) return 0.f;
        VolSample v = sample_volume(*scene, rng, interior);
        float res = 0.f;
        Vec<D> q = vp(v);
        for (unsigned i = 0; i < nvs; ++i) {
            if (indep && i > 0) { v = sample_volume(*scene, rng, interior); q = vp(v); }
            if (boundary_form) res -= k.G(x, q) * v.value * v.inv_pdf;
            else res -= k.G(q, x) * v.value * v.inv_pdf;
        }
        return res / nvs;
    }
    // :137-154
    Vec<D> vol_init_derivative(const Vec<D> &x, Rng &rng) const {
        Vec<D> res = Vec<D>::zero();
        if (nvs == 0) return res;
        VolSample v = sample_volume(*scene, rng, interior);
        Vec<D> q = vp(v);
        for (unsigned i = 0; i < nvs; ++i) {
            if (indep && i > 0) { v = sample_volume(*scene, rng, interior); q = vp(v); }
            res = res - (v.value * v.inv_pdf) * k.dG_dx(q, x);
        }
        return res / (float)nvs;
    }

    // ---- domain terms
    float solution_domain_unknown(const Vec<D> &x, const BPoint<D> &y, float inv) const {
        if (bie == IndirectSingleLayer) return k.G(y.p, x) * inv;                                   // :228-231
        if (bie == IndirectDoubleLayer) return -k.dG_dny(y.p, x, y.n) * inv;                        // :341-344
        if (y.boundary_type == Dirichlet) return dsign * inv * k.G(y.p, x);                         // :429-430
        if (y.boundary_type == Neumann) return -dsign * inv * k.dG_dny(y.p, x, y.n);                // :431-432
        return -dsign * inv * (k.dG_dny(y.p, x, y.n) + y.robin_alpha * k.G(y.p, x));               // :433-435
    }
    Vec<D> gradient_domain_unknown(const Vec<D> &x, const BPoint<D> &y, float inv) const {
        if (bie == IndirectSingleLayer) return k.dG_dx(y.p, x) * inv;                               // :240-243
        if (bie == IndirectDoubleLayer) return -1.0f * (k.d2G_dxdny(y.p, x, y.n) * inv);            // :353-356
        if (y.boundary_type == Dirichlet) return (dsign * inv) * k.dG_dx(y.p, x);                   // :454-455
        if (y.boundary_type == Neumann) return (-dsign * inv) * k.d2G_dxdny(y.p, x, y.n);           // :456-457
        return (-dsign * inv) * (k.d2G_dxdny(y.p, x, y.n) + y.robin_alpha * k.dG_dx(y.p, x));      // :458-461
    }
    float solution_domain_known(const Vec<D> &x, const BPoint<D> &y, float inv, Rng &rng) const {
        if (bie != Direct) return vol_init(x, false, rng);                                          // :233-238, :346-351
        float res = vol_init(x, false, rng);                                                        // :442
        if (y.boundary_type == Dirichlet)
            res += -dsign * inv * k.dG_dny(y.p, x, y.n) * (y.boundary_value + vol_term(y, rng));    // :443-445
        else
            res += dsign * inv * k.G(y.p, x) * (y.boundary_value + vol_term(y, rng));               // :446-448
        return res;
    }
    Vec<D> gradient_domain_known(const Vec<D> &x, const BPoint<D> &y, float inv, Rng &rng) const {
        if (bie != Direct) return vol_init_derivative(x, rng);                                      // :245-250, :358-363
        Vec<D> res = vol_init_derivative(x, rng);                                                   // :468
        if (y.boundary_type == Dirichlet)
            res = res + ((-dsign * inv) * k.d2G_dxdny(y.p, x, y.n)) * (y.boundary_value + vol_term(y, rng));  // :469-472
        else
            res = res + ((dsign * inv) * k.dG_dx(y.p, x)) * (y.boundary_value + vol_term(y, rng));           // :473-476
        return res;
    }
    // ---- boundary terms
    float solution_boundary_unknown(const BPoint<D> &x, const BPoint<D> &y, float inv) const {
        if (bie == IndirectSingleLayer) {                                                           // :252-257
            if (x.boundary_type == Dirichlet) return 0.f;
            return k.G(y.p, x.p) * inv;
        }
        if (bie == IndirectDoubleLayer) return 0.f;                                                 // :365-369
        if (x.boundary_type == Dirichlet) return 0.f;                                               // :482
        const float f = ift(x.interior_angle);
        if (y.boundary_type == Dirichlet) return dsign / f * inv * k.G(y.p, x.p);                   // :484-486
        if (y.boundary_type == Neumann) return -dsign / f * inv * k.dG_dny(y.p, x.p, y.n);          // :487-489
        return -dsign / f * inv * (k.dG_dny(y.p, x.p, y.n) + y.robin_alpha * k.G(y.p, x.p));       // :490-492
    }
    float solution_boundary_known(const BPoint<D> &x, const BPoint<D> &y, float inv, Rng &rng) const {
        if (bie == IndirectSingleLayer) {                                                           // :259-266
            if (x.boundary_type == Dirichlet) return x.boundary_value;
            return vol_init(x.p, true, rng);
        }
        if (bie == IndirectDoubleLayer) return x.boundary_value;                                    // :371-377
        if (x.boundary_type == Dirichlet) return x.boundary_value;                                  // :499
        float res = vol_init(x.p, true, rng);                                                       // :501
        const float f = ift(x.interior_angle);
        if (y.boundary_type == Dirichlet)
            res += -dsign / f * inv * k.dG_dny(y.p, x.p, y.n) * (y.boundary_value + vol_term(y, rng));   // :502-505
        else
            res += dsign / f * inv * k.G(y.p, x.p) * (y.boundary_value + vol_term(y, rng));              // :506-509
        return res;
    }
    // ---- Fredholm terms
    // u_split: the uniform used by the first-kind splitting (drawn by the caller only when needed).
    float fredholm_unknown(BPoint<D> &x, BPoint<D> &y, float inv, Rng &rng, bool backward) const {
        if (bie == IndirectDoubleLayer)                                                             // :379-388
            return dsign / (1.f - ift(x.interior_angle)) * inv * k.dG_dny(y.p, x.p, y.n);
        // first-kind splitting (SL :272-288, Direct :517-533)
        if (backward && x.boundary_type == Dirichlet) {
            if (rng.uniform() * (1.0f + c) < 1.0f) { y = x; return (1.0f + c); }
            inv *= (1.0f + c) / c;
        } else if (!backward && y.boundary_type == Dirichlet) {
            if (rng.uniform() * (1.0f + c) < 1.0f) { x = y; return (1.0f + c); }
            inv *= (1.0f + c) / c;
        }
        if (bie == IndirectSingleLayer) {
            if (x.boundary_type == Dirichlet) return -2.f * c * inv * k.G(y.p, x.p);                                // :290-291
            if (x.boundary_type == Neumann) return -dsign / ift(x.interior_angle) * inv * k.dG_dny(x.p, y.p, x.n);  // :292-294
            return -dsign / ift(x.interior_angle) * inv * (k.dG_dny(x.p, y.p, x.n) + x.robin_alpha * k.G(y.p, x.p)); // :295-299
        }
        const float f = ift(x.interior_angle);                                                      // Direct :535-555
        if (x.boundary_type == Dirichlet) {
            if (y.boundary_type == Dirichlet) return -c / f * inv * k.G(y.p, x.p);
            if (y.boundary_type == Neumann) return c / f * inv * k.dG_dny(y.p, x.p, y.n);
            return c / f * inv * (k.dG_dny(y.p, x.p, y.n) + y.robin_alpha * k.G(y.p, x.p));
        }
        if (y.boundary_type == Dirichlet) return dsign / f * inv * k.G(y.p, x.p);
        if (y.boundary_type == Neumann) return -dsign / f * inv * k.dG_dny(y.p, x.p, y.n);
        return -dsign / f * inv * (k.dG_dny(y.p, x.p, y.n) + y.robin_alpha * k.G(y.p, x.p));
    }
    float fredholm_known(const BPoint<D> &x, const BPoint<D> &y, float inv, Rng &rng) const {
        if (bie == IndirectSingleLayer) {                                                           // :301-316
            if (x.boundary_type == Dirichlet) return 2.f * c * (x.boundary_value + vol_term(x, rng));
            return dsign / ift(x.interior_angle) * (x.boundary_value + vol_term(x, rng));
        }
        if (bie == IndirectDoubleLayer)                                                             // :389-399
            return dsign * 1.f / (1.f - ift(x.interior_angle)) * (x.boundary_value + vol_term(x, rng));
        const float f = ift(x.interior_angle);                                                      // Direct :557-587
        if (x.boundary_type == Dirichlet) {
            if (y.boundary_type == Dirichlet) {
                const float a = x.boundary_value + vol_term(x, rng);
                return dsign * c * (a + dsign / f * inv * k.dG_dny(y.p, x.p, y.n) * (y.boundary_value + vol_term(y, rng)));
            }
            const float a = x.boundary_value + vol_term(x, rng);
            return dsign * c * (a - dsign / f * inv * k.G(y.p, x.p) * (y.boundary_value + vol_term(y, rng)));
        }
        if (y.boundary_type == Dirichlet)
            return -dsign / f * inv * k.dG_dny(y.p, x.p, y.n) * (y.boundary_value + vol_term(y, rng));
        return dsign / f * inv * k.G(y.p, x.p) * (y.boundary_value + vol_term(y, rng));
    }
};

// ------------------------------------------------------------------------------------
// Estimators (estimator.cuh)
// ------------------------------------------------------------------------------------
template <int D> struct Estimator {
    const Scene<D> *s;
    Model<D> m;
    int est;             // EstimatorType
    int problem;         // ProblemType
    unsigned L;          // path_length
    unsigned nrc;        // num_resampling_candidates
    Options opt;

    bool dl_dirichlet() const { return m.bie == IndirectDoubleLayer && problem == DirichletProblem; }
    bool direct_neumann() const { return m.bie == Direct && problem == NeumannProblem; }

    Sampled<D> ris(const BPoint<D> &yp, Rng &rng) const {
        // scene.cuh:542-584; target chosen by the type of y_prev (estimator.cuh:286-328)
        auto target = [&](const BPoint<D> &y) -> float {
            if (yp.boundary_type == Dirichlet) return std::abs(m.k.G(y.p, yp.p));
            if (yp.boundary_type == Neumann) return std::abs(m.k.dG_dny(yp.p, y.p, yp.n));
            return std::abs(m.k.dG_dny(yp.p, y.p, yp.n) + yp.robin_alpha * m.k.G(y.p, yp.p));
        };
        BPoint<D> res = BPoint<D>::zero();
        float w_sum = 0.f;
        unsigned cnt = 0;
        for (unsigned i = 0; i < nrc; ++i) {
            Sampled<D> xi = sample_uniform(*s, rng);
            const float w = target(xi.y) * xi.inv_pdf;
            w_sum += w;
            cnt += 1;
            if (rng.uniform() <= w / w_sum) res = xi.y;
        }
        const float inv = w_sum / std::max(target(res) * cnt, 10 * kFltEps);
        return {res, inv};
    }

    // estimator.cuh:232-357
    float walk(const BPoint<D> &x0, Rng &rng) const {
        float spc = 0.f, W = 1.f;
        BPoint<D> yp = x0;
        BPoint<D> y = BPoint<D>::zero();
        float inv = 0.f;
        for (unsigned it = 0; it < L; ++it) {
            if (it == L - 1) W = 0.5f * W;
            if (dl_dirichlet()) {
                Sampled<D> q = sample_line_intersection(*s, yp.p, rng, opt, yp.elem);
                y = q.y; inv = q.inv_pdf;
                if (opt.harmonic_cv) spc += W * m.fredholm_known(hcv_residual(*s, yp), y, inv, rng);
                else spc += W * m.fredholm_known(yp, y, inv, rng);
                W *= m.fredholm_unknown(yp, y, inv, rng, true);
                yp = y;
            } else if (direct_neumann()) {
                // with the control variate the known term's sample is uniform (the residual data
                // need not vanish where the data do) and both points carry residual data
                Sampled<D> q = opt.harmonic_cv ? sample_uniform(*s, rng) : sample_boundary_value(*s, rng);
                y = q.y; inv = q.inv_pdf;
                if (opt.harmonic_cv) spc += W * m.fredholm_known(hcv_residual(*s, yp), hcv_residual(*s, y), inv, rng);
                else spc += W * m.fredholm_known(yp, y, inv, rng);
                q = sample_line_intersection(*s, yp.p, rng, opt, yp.elem);
                y = q.y; inv = q.inv_pdf;
                W = W * m.fredholm_unknown(yp, y, inv, rng, true);
                yp = y;
            } else if (m.bie == IndirectSingleLayer) {
                if (opt.harmonic_cv) spc += W * m.fredholm_known(hcv_residual(*s, yp), y, inv, rng);
                else spc += W * m.fredholm_known(yp, y, inv, rng);
                Sampled<D> q = nrc == 1 ? sample_uniform(*s, rng) : ris(yp, rng);
                y = q.y; inv = q.inv_pdf;
                W = W * m.fredholm_unknown(yp, y, inv, rng, true);
                yp = y;
            } else {
                Sampled<D> q = sample_uniform(*s, rng);
                y = q.y; inv = q.inv_pdf;
                if (opt.harmonic_cv) spc += W * m.fredholm_known(hcv_residual(*s, yp), hcv_residual(*s, y), inv, rng);
                else spc += W * m.fredholm_known(yp, y, inv, rng);
                q = sample_uniform(*s, rng);
                y = q.y; inv = q.inv_pdf;
                W = W * m.fredholm_unknown(yp, y, inv, rng, true);
                yp = y;
            }
            if (W == 0.0f) break;
        }
        return spc;
    }

    void first_samples(const Vec<D> &x, Sampled<D> &y, Sampled<D> &y1, Rng &rng, int origin_elem = -1) const {
        if (dl_dirichlet()) { y = sample_line_intersection(*s, x, rng, opt, origin_elem); y1 = y; }
        else if (direct_neumann()) {
            y = sample_line_intersection(*s, x, rng, opt, origin_elem);
            y1 = opt.harmonic_cv ? sample_uniform(*s, rng) : sample_boundary_value(*s, rng);
        }
        else { y = sample_uniform(*s, rng); y1 = sample_uniform(*s, rng); }
    }

    // out: [u, grad u] (D + 1)
    void domain_path(const Vec<D> &x, Rng &rng, float *out) const {
        if (est == BackwardEstimator) {                                                             // :154-191
            Sampled<D> y, y1;
            first_samples(x, y, y1, rng);
            if (opt.harmonic_cv && m.bie == Direct) y1.y = hcv_residual(*s, y1.y);   // the known terms' datum
            out[0] = m.solution_domain_known(x, y1.y, y1.inv_pdf, rng);
            Vec<D> g = m.gradient_domain_known(x, y1.y, y1.inv_pdf, rng);
            const float spc = walk(y.y, rng);
            out[0] += spc * m.solution_domain_unknown(x, y.y, y.inv_pdf);
            const Vec<D> gu = m.gradient_domain_unknown(x, y.y, y.inv_pdf);
            for (int k = 0; k < D; ++k) out[1 + k] = g[k] + spc * gu[k];
            if (opt.harmonic_cv) {   // u = u_h + the walk's estimate of u - u_h
                float gh[3];
                hcv_gradient(*s, x.v, gh);
                out[0] += hcv_value(*s, x.v);
                for (int k = 0; k < D; ++k) out[1 + k] += gh[k];
            }
            return;
        }
        // forward :405-454.  With the harmonic control variate the start point is uniform
        // (the residual data need not vanish where the data do, so the data's own density
        // would miss part of them) and the known term reads the residual datum.
        Sampled<D> y = sample_uniform(*s, rng);
        out[0] = m.solution_domain_known(x, y.y, y.inv_pdf, rng);
        Vec<D> g = m.gradient_domain_known(x, y.y, y.inv_pdf, rng);
        const bool sl_neumann = m.bie == IndirectSingleLayer && problem == NeumannProblem;
        y = sl_neumann && !opt.harmonic_cv ? sample_boundary_value(*s, rng) : sample_uniform(*s, rng);
        Sampled<D> y2 = sample_uniform(*s, rng);
        float W = y.inv_pdf * m.fredholm_known(opt.harmonic_cv ? hcv_residual(*s, y.y) : y.y, y2.y, y2.inv_pdf, rng);
        BPoint<D> yc = y.y;
        for (unsigned it = 0; it < L; ++it) {
            if (it == L - 1) W *= 0.5f;
            out[0] += W * m.solution_domain_unknown(x, yc, 1.0f);
            const Vec<D> gu = m.gradient_domain_unknown(x, yc, 1.0f);
            for (int k = 0; k < D; ++k) g[k] += W * gu[k];
            BPoint<D> yp = yc;
            Sampled<D> q = sl_neumann ? sample_line_intersection(*s, yp.p, rng, opt, yp.elem) : sample_uniform(*s, rng);
            yc = q.y;
            W = m.fredholm_unknown(yc, yp, q.inv_pdf, rng, false) * W;
        }
        for (int k = 0; k < D; ++k) out[1 + k] = g[k];
        if (opt.harmonic_cv) {   // u = u_h + the walk's estimate of u - u_h
            float gh[3];
            hcv_gradient(*s, x.v, gh);
            out[0] += hcv_value(*s, x.v);
            for (int k = 0; k < D; ++k) out[1 + k] += gh[k];
        }
    }

    float boundary_path(const BPoint<D> &x, Rng &rng) const {
        if (est == BackwardEstimator) {                                                             // :193-222
            Sampled<D> y, y1;
            first_samples(x.p, y, y1, rng, x.elem);
            float res = m.solution_boundary_known(x, y1.y, y1.inv_pdf, rng);
            const float spc = walk(y.y, rng);
            res += spc * m.solution_boundary_unknown(x, y.y, y.inv_pdf);
            return res;
        }
        // forward :456-496
        Sampled<D> y = sample_uniform(*s, rng);
        float res = m.solution_boundary_known(x, y.y, y.inv_pdf, rng);
        const bool sl_neumann = m.bie == IndirectSingleLayer && problem == NeumannProblem;
        y = sl_neumann ? sample_boundary_value(*s, rng) : sample_uniform(*s, rng);
        Sampled<D> y2 = sample_uniform(*s, rng);
        float W = y.inv_pdf * m.fredholm_known(y.y, y2.y, y2.inv_pdf, rng);
        BPoint<D> yc = y.y;
        for (unsigned it = 0; it < L; ++it) {
            if (it == L - 1) W *= 0.5f;
            res += m.solution_boundary_unknown(x, yc, 1.0f) * W;
            BPoint<D> yp = yc;
            Sampled<D> q = sl_neumann ? sample_line_intersection(*s, yp.p, rng, opt, yp.elem) : sample_uniform(*s, rng);
            yc = q.y;
            W = m.fredholm_unknown(yc, yp, q.inv_pdf, rng, false) * W;
        }
        return res;
    }
};

}  // namespace orc

// ======================================================================================
// C ABI for ctypes (tests, smoke checker, CPU baseline)
// ======================================================================================
using namespace orc;

struct OrcScene {
    int dim;
    Scene<2> s2;
    Scene<3> s3;
};

struct orc_est_desc {
    int32_t bie_model, domain, estimator, problem;
    uint32_t path_length, num_resampling_candidates, num_volume_samples;
    float eps, first_kind_equation_scaling_constant;
    int32_t pick_mode, rng_mode, use_bvh, exclude_origin;
    int32_t independent_volume;
    int32_t boundary_density;   // boundary entry: return compute_sample_unknown_boundary (the walk)
    int32_t harmonic_cv;        // single/double-layer walks: the residual of the scene's harmonic u_h
};

extern "C" {

void orc_philox4x32_10(const uint32_t *ctr, const uint32_t *key, uint32_t *out) { philox4x32_10(ctr, key, out); }

uint32_t orc_triple32(uint32_t x) { return triple32(x); }

// out per sample: [G, dG_dx (D), dG_dny, d2G_dxdny (D), d2G_dnxdny]  (3 + 2D values)
void orc_eval_kernels_f32(int dim, int n, const float *y, const float *x, const float *ny, const float *nx, float eps,
                          float *out) {
    const int D = dim, W = 3 + 2 * D;
    for (int i = 0; i < n; ++i) {
        float *o = out + (size_t)i * W;
        if (D == 2) {
            Kernel<2> k{eps};
            Vec<2> Y{{y[2 * i], y[2 * i + 1]}}, X{{x[2 * i], x[2 * i + 1]}}, NY{{ny[2 * i], ny[2 * i + 1]}}, NX{{nx[2 * i], nx[2 * i + 1]}};
            o[0] = k.G(Y, X);
            auto g = k.dG_dx(Y, X); o[1] = g[0]; o[2] = g[1];
            o[3] = k.dG_dny(Y, X, NY);
            auto h = k.d2G_dxdny(Y, X, NY); o[4] = h[0]; o[5] = h[1];
            o[6] = k.d2G_dnxdny(Y, X, NY, NX);
        } else {
            Kernel<3> k{eps};
            Vec<3> Y{{y[3 * i], y[3 * i + 1], y[3 * i + 2]}}, X{{x[3 * i], x[3 * i + 1], x[3 * i + 2]}};
            Vec<3> NY{{ny[3 * i], ny[3 * i + 1], ny[3 * i + 2]}}, NX{{nx[3 * i], nx[3 * i + 1], nx[3 * i + 2]}};
            o[0] = k.G(Y, X);
            auto g = k.dG_dx(Y, X); o[1] = g[0]; o[2] = g[1]; o[3] = g[2];
            o[4] = k.dG_dny(Y, X, NY);
            auto h = k.d2G_dxdny(Y, X, NY); o[5] = h[0]; o[6] = h[1]; o[7] = h[2];
            o[8] = k.d2G_dnxdny(Y, X, NY, NX);
        }
    }
}

void orc_eval_kernels_f64(int dim, int n, const double *y, const double *x, const double *ny, const double *nx,
                          double eps, double *out) {
    const int D = dim, W = 3 + 2 * D;
    for (int i = 0; i < n; ++i) {
        double *o = out + (size_t)i * W;
        if (D == 2) {
            Kernel64<2> k{eps};
            Vec<2, double> Y{{y[2 * i], y[2 * i + 1]}}, X{{x[2 * i], x[2 * i + 1]}}, NY{{ny[2 * i], ny[2 * i + 1]}}, NX{{nx[2 * i], nx[2 * i + 1]}};
            o[0] = k.G(Y, X);
            auto g = k.dG_dx(Y, X); o[1] = g[0]; o[2] = g[1];
            o[3] = k.dG_dny(Y, X, NY);
            auto h = k.d2G_dxdny(Y, X, NY); o[4] = h[0]; o[5] = h[1];
            o[6] = k.d2G_dnxdny(Y, X, NY, NX);
        } else {
            Kernel64<3> k{eps};
            Vec<3, double> Y{{y[3 * i], y[3 * i + 1], y[3 * i + 2]}}, X{{x[3 * i], x[3 * i + 1], x[3 * i + 2]}};
            Vec<3, double> NY{{ny[3 * i], ny[3 * i + 1], ny[3 * i + 2]}}, NX{{nx[3 * i], nx[3 * i + 1], nx[3 * i + 2]}};
            o[0] = k.G(Y, X);
            auto g = k.dG_dx(Y, X); o[1] = g[0]; o[2] = g[1]; o[3] = g[2];
            o[4] = k.dG_dny(Y, X, NY);
            auto h = k.d2G_dxdny(Y, X, NY); o[5] = h[0]; o[6] = h[1]; o[7] = h[2];
            o[8] = k.d2G_dnxdny(Y, X, NY, NX);
        }
    }
}

// Elements: vertices n*D*D (a, b[, c]), normals n*D, boundary values n*D, types n, alphas n*D.
void *orc_scene_create(int dim, int n, const float *verts, const float *normals, const float *bvals,
                       const int32_t *types, const float *alphas) {
    if ((dim != 2 && dim != 3) || n <= 0) return nullptr;
    auto *S = new OrcScene();
    S->dim = dim;
    auto fill = [&](auto &scene, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        scene.elems.resize(n);
        for (int i = 0; i < n; ++i) {
            auto &e = scene.elems[i];
            for (int j = 0; j < D; ++j)
                for (int k = 0; k < D; ++k) e.v[j][k] = verts[((size_t)i * D + j) * D + k];
            for (int k = 0; k < D; ++k) e.n[k] = normals[(size_t)i * D + k];
            for (int j = 0; j < D; ++j) { e.bv[j] = bvals[(size_t)i * D + j]; e.alpha[j] = alphas ? alphas[(size_t)i * D + j] : 0.f; }
            e.type = types[i];
        }
        build_cdfs(scene);
        build_lbvh(scene);
    };
    if (dim == 2) fill(S->s2, std::integral_constant<int, 2>{});
    else fill(S->s3, std::integral_constant<int, 3>{});
    return S;
}

void orc_scene_destroy(void *p) { delete (OrcScene *)p; }

// totals[0] = total area, totals[1] = total boundary value
void orc_scene_totals(void *p, float *totals) {
    auto *S = (OrcScene *)p;
    if (S->dim == 2) { totals[0] = S->s2.total_area; totals[1] = S->s2.total_bv; }
    else { totals[0] = S->s3.total_area; totals[1] = S->s3.total_bv; }
}

void orc_scene_cdfs(void *p, float *area_cdf, float *bv_cdf) {
    auto *S = (OrcScene *)p;
    const auto &a = S->dim == 2 ? S->s2.area_cdf : S->s3.area_cdf;
    const auto &b = S->dim == 2 ? S->s2.bv_cdf : S->s3.bv_cdf;
    std::copy(a.begin(), a.end(), area_cdf);
    std::copy(b.begin(), b.end(), bv_cdf);
}

// Grid volume source (scene.cuh:206-240): cdf of |cache|, normalised, inclusive scan.
int orc_scene_set_volume_grid(void *p, unsigned grid_res, float domain_size, const float *cache) {
    auto *S = (OrcScene *)p;
    VolumeGrid &v = S->dim == 2 ? S->s2.vol : S->s3.vol;
    size_t cells = (size_t)grid_res * grid_res * (S->dim == 3 ? grid_res : 1);
    v.present = true; v.grid_res = grid_res; v.domain_size = domain_size;
    v.cache.assign(cache, cache + cells);
    v.cdf.resize(cells);
    float total = 0.f;
    for (size_t i = 0; i < cells; ++i) { v.cdf[i] = std::abs(cache[i]); total += v.cdf[i]; }
    if (total == 0.f) for (auto &c : v.cdf) c = 1.f / cells;
    else for (auto &c : v.cdf) c /= total;
    for (size_t i = 1; i < cells; ++i) v.cdf[i] += v.cdf[i - 1];
    return 0;
}

// Gaussian-mixture volume source (see VolumeGauss); n = 0 removes it.  The host-side
// constants are computed in double and rounded once.
int orc_scene_set_volume_gaussians(void *p, unsigned n, const float *centers, const float *widths, const float *amps,
                                   int mask) {
    auto *S = (OrcScene *)p;
    VolumeGauss &g = S->dim == 2 ? S->s2.gauss : S->s3.gauss;
    const int D = S->dim;
    g = VolumeGauss();
    if (n == 0) return 0;
    g.present = true; g.mask = mask != 0; g.n = n;
    g.c.assign((size_t)n * 3, 0.f);
    std::vector<double> wt(n);
    double tot = 0;
    for (unsigned j = 0; j < n; ++j) {
        for (int k = 0; k < D; ++k) g.c[3 * j + k] = centers[(size_t)j * D + k];
        wt[j] = std::abs((double)amps[j]) * std::pow((double)widths[j], D);
        tot += wt[j];
    }
    double acc = 0;
    for (unsigned j = 0; j < n; ++j) {
        const double w = widths[j], pj = wt[j] / tot;
        acc += pj;
        g.w.push_back(widths[j]);
        g.amp.push_back(amps[j]);
        g.inv2w2.push_back((float)(1.0 / (2.0 * w * w)));
        g.pn.push_back((float)(pj * std::pow(2.0 * M_PI * w * w, -0.5 * D)));
        g.cdf.push_back(j + 1 == n ? 1.0f : (float)acc);
    }
    if (g.mask) {
        if (D == 2) build_inside_grid(S->s2, g, kInsideGridRes2);
        else build_inside_grid(S->s3, g, kInsideGridRes3);
    }
    return 0;
}

// Inside/outside parity of a point (ray crossings with t > 0 along the fixed direction).
int orc_inside(void *p, const float *pt) {
    auto *S = (OrcScene *)p;
    if (S->dim == 2) { Vec<2> o{{pt[0], pt[1]}}; return (int)(ray_crossings(S->s2, o) & 1u); }
    Vec<3> o{{pt[0], pt[1], pt[2]}};
    return (int)(ray_crossings(S->s3, o) & 1u);
}

// Inside/outside verdict of a masked Gaussian source's volume sample at pt: the occupancy
// grid's cell, or the point's own ray where the grid leaves it open (-1: no masked source).
int orc_inside_grid(void *p, const float *pt, int *cell) {
    auto *S = (OrcScene *)p;
    if (S->dim == 2) {
        const VolumeGauss &g = S->s2.gauss;
        if (!g.present || !g.mask) return -1;
        *cell = inside_grid_cell<2>(g, pt);
        if (*cell != 2) return *cell;
        Vec<2> o{{pt[0], pt[1]}};
        return (int)(ray_crossings(S->s2, o) & 1u);
    }
    const VolumeGauss &g = S->s3.gauss;
    if (!g.present || !g.mask) return -1;
    *cell = inside_grid_cell<3>(g, pt);
    if (*cell != 2) return *cell;
    Vec<3> o{{pt[0], pt[1], pt[2]}};
    return (int)(ray_crossings(S->s3, o) & 1u);
}

// All-hit line query; returns the hit count, ids / t of the first `cap` hits (unsorted).
int orc_intersect_line(void *p, const float *origin, const float *dir, int use_bvh, int cap, int32_t *ids, float *ts) {
    auto *S = (OrcScene *)p;
    std::vector<uint32_t> dummy;
    if (S->dim == 2) {
        std::vector<Hit<2>> buf(std::max(cap, 1));
        Vec<2> o{{origin[0], origin[1]}}, d{{dir[0], dir[1]}};
        unsigned c = collect_hits(S->s2, o, d, use_bvh != 0, buf.data(), (unsigned)cap);
        for (unsigned i = 0; i < std::min(c, (unsigned)cap); ++i) { ids[i] = buf[i].obj; ts[i] = buf[i].t; }
        return (int)c;
    }
    std::vector<Hit<3>> buf(std::max(cap, 1));
    Vec<3> o{{origin[0], origin[1], origin[2]}}, d{{dir[0], dir[1], dir[2]}};
    unsigned c = collect_hits(S->s3, o, d, use_bvh != 0, buf.data(), (unsigned)cap);
    for (unsigned i = 0; i < std::min(c, (unsigned)cap); ++i) { ids[i] = buf[i].obj; ts[i] = buf[i].t; }
    return (int)c;
}

// orc_intersect_line over n lines (origins, dirs: n * dim floats), OpenMP over the lines:
// counts[i], and ids / t of the first `cap` hits of line i at [i * cap, (i + 1) * cap).
void orc_intersect_lines(void *p, int n, const float *origins, const float *dirs, int use_bvh, int cap, int threads,
                         int32_t *counts, int32_t *ids, float *ts) {
    auto *S = (OrcScene *)p;
    const int D = S->dim;
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(dynamic, 64)
    for (int i = 0; i < n; ++i)
        counts[i] = orc_intersect_line(p, origins + (size_t)i * D, dirs + (size_t)i * D, use_bvh, cap,
                                       ids + (size_t)i * cap, ts + (size_t)i * cap);
}

}  // extern "C"

template <int D> static Estimator<D> make_est(const Scene<D> &s, const orc_est_desc *d) {
    Estimator<D> e;
    e.s = &s;
    e.m.bie = d->bie_model;
    e.m.interior = d->domain == InteriorDomain;
    e.m.dsign = e.m.interior ? 1.0f : -1.0f;
    e.m.k.eps = d->eps;
    e.m.c = d->first_kind_equation_scaling_constant;
    e.m.nvs = d->num_volume_samples;
    e.m.indep = d->independent_volume != 0;
    e.m.scene = &s;
    e.est = d->estimator;
    e.problem = d->problem;
    e.L = d->path_length;
    e.nrc = d->num_resampling_candidates;
    e.opt.pick_mode = d->pick_mode;
    e.opt.use_bvh = d->use_bvh != 0;
    e.opt.exclude_origin = d->exclude_origin != 0;
    e.opt.harmonic_cv = d->harmonic_cv != 0;
    return e;
}

static std::vector<uint64_t> reference_seeds(uint64_t n) {
    // runtime_utils.cuh:89-93: default-constructed mt19937_64, one draw per point
    std::mt19937_64 g;
    std::vector<uint64_t> s(n);
    for (auto &v : s) v = g();
    return s;
}

template <int D>
static void run_domain(const Scene<D> &s, const orc_est_desc *d, const float *pts, const uint32_t *pt_index, int n,
                       unsigned paths, uint64_t seed, uint64_t path_base, unsigned stream, int threads, double *sum,
                       double *sumsq, float *path_out) {
    Estimator<D> e = make_est(s, d);
    std::vector<uint64_t> seeds;
    if (d->rng_mode == 1) {
        uint64_t mx = 0;
        for (int i = 0; i < n; ++i) mx = std::max<uint64_t>(mx, pt_index ? pt_index[i] : (uint64_t)i);
        seeds = reference_seeds(mx + 1);
    }
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const uint32_t gid = pt_index ? pt_index[i] : (uint32_t)i;
        Vec<D> x;
        for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
        Kahan ks[D + 1], kq[D + 1];
        Rng rng;
        if (d->rng_mode == 1) { rng.mode = 1; rng.mt.seed((std::mt19937::result_type)seeds[gid]); }
        for (unsigned j = 0; j < paths; ++j) {
            if (d->rng_mode == 0) rng.init_philox(seed, (uint32_t)(path_base + j), gid, stream);
            float out[D + 1];
            e.domain_path(x, rng, out);
            for (int k = 0; k <= D; ++k) { ks[k].add(out[k]); kq[k].add(out[k] * out[k]); }
            if (path_out)
                for (int k = 0; k <= D; ++k) path_out[((size_t)i * paths + j) * (D + 1) + k] = out[k];
        }
        for (int k = 0; k <= D; ++k) {
            sum[(size_t)i * (D + 1) + k] += ks[k].sum;
            sumsq[(size_t)i * (D + 1) + k] += kq[k].sum;
        }
    }
}

template <int D>
static void run_boundary(const Scene<D> &s, const orc_est_desc *d, const float *bpts, int n, unsigned paths,
                         uint64_t seed, uint64_t path_base, unsigned stream, int threads, double *sum, double *sumsq,
                         float *path_out) {
    Estimator<D> e = make_est(s, d);
    const int W = 2 * D + 4;
#ifdef _OPENMP
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < n; ++i) {
        const float *r = bpts + (size_t)i * W;
        BPoint<D> x;
        for (int k = 0; k < D; ++k) { x.p[k] = r[k]; x.n[k] = r[D + k]; }
        x.interior_angle = r[2 * D]; x.boundary_value = r[2 * D + 1]; x.boundary_type = (int)r[2 * D + 2]; x.robin_alpha = r[2 * D + 3];
        Kahan ks, kq;
        Rng rng;
        for (unsigned j = 0; j < paths; ++j) {
            rng.init_philox(seed, (uint32_t)(path_base + j), (uint32_t)i, stream);
            const float v = d->boundary_density ? e.walk(x, rng) : e.boundary_path(x, rng);
            ks.add(v); kq.add(v * v);
            if (path_out) path_out[(size_t)i * paths + j] = v;
        }
        sum[i] += ks.sum;
        sumsq[i] += kq.sum;
    }
}

extern "C" {

// Per-point estimates at the given points.  pt_index (optional) gives each point's global
// index used as the RNG counter word (defaults to the position in `pts`).  sum / sumsq
// (n * (D+1) doubles) are accumulated into.  path_out (optional): n * paths * (D+1) floats.
int orc_estimate_domain(void *p, const orc_est_desc *d, const float *pts, const uint32_t *pt_index, int n,
                        unsigned paths, uint64_t seed, uint64_t path_base, unsigned stream, int threads, double *sum,
                        double *sumsq, float *path_out) {
    auto *S = (OrcScene *)p;
    if (S->dim == 2) run_domain(S->s2, d, pts, pt_index, n, paths, seed, path_base, stream, threads, sum, sumsq, path_out);
    else run_domain(S->s3, d, pts, pt_index, n, paths, seed, path_base, stream, threads, sum, sumsq, path_out);
    return 0;
}

// Boundary points: records of 2D+4 floats [p (D), n (D), interior_angle, boundary_value, type, robin_alpha].
int orc_estimate_boundary(void *p, const orc_est_desc *d, const float *bpts, int n, unsigned paths, uint64_t seed,
                          uint64_t path_base, unsigned stream, int threads, double *sum, double *sumsq,
                          float *path_out) {
    auto *S = (OrcScene *)p;
    if (S->dim == 2) run_boundary(S->s2, d, bpts, n, paths, seed, path_base, stream, threads, sum, sumsq, path_out);
    else run_boundary(S->s3, d, bpts, n, paths, seed, path_base, stream, threads, sum, sumsq, path_out);
    return 0;
}

// Deterministic weight evaluation for parity with the device.  Inputs: n pairs of boundary
// records x, y (2D+4 floats each), inv_pdf, domain points xd (D floats) and one split
// uniform per pair (used only when the first-kind splitting applies).
// out per pair: [fredholm_known, fredholm_unknown, split_taken, solution_domain_unknown,
//                gradient_domain_unknown (D), solution_boundary_unknown] = 5 + D floats.
int orc_eval_weights(int dim, int bie, int domain, float eps, float fkesc, int n, const float *xb, const float *yb,
                     const float *inv_pdf, const float *xd, const float *u_split, float *out) {
    auto run = [&](auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        const int W = 2 * D + 4, O = 5 + D;
        Model<D> m;
        m.bie = bie; m.interior = domain == InteriorDomain; m.dsign = m.interior ? 1.f : -1.f;
        m.k.eps = eps; m.c = fkesc; m.nvs = 0; m.scene = nullptr;
        for (int i = 0; i < n; ++i) {
            auto rec = [&](const float *r) {
                BPoint<D> b;
                for (int k = 0; k < D; ++k) { b.p[k] = r[k]; b.n[k] = r[D + k]; }
                b.interior_angle = r[2 * D]; b.boundary_value = r[2 * D + 1]; b.boundary_type = (int)r[2 * D + 2]; b.robin_alpha = r[2 * D + 3];
                return b;
            };
            BPoint<D> x = rec(xb + (size_t)i * W), y = rec(yb + (size_t)i * W);
            Vec<D> X; for (int k = 0; k < D; ++k) X[k] = xd[(size_t)i * D + k];
            float *o = out + (size_t)i * O;
            Rng dummy;
            o[0] = m.fredholm_known(x, y, inv_pdf[i], dummy);
            // the split uniform is supplied by the caller (Rng mode 2 returns it)
            Rng fixed; fixed.mode = 2; fixed.fixed_u = u_split[i];
            BPoint<D> x2 = x, y2 = y;
            o[1] = m.fredholm_unknown(x2, y2, inv_pdf[i], fixed, true);
            const float split = (bie != IndirectDoubleLayer && x.boundary_type == Dirichlet &&
                                 u_split[i] * (1.0f + fkesc) < 1.0f) ? 1.f : 0.f;
            o[2] = split;
            o[3] = m.solution_domain_unknown(X, y, inv_pdf[i]);
            Vec<D> g = m.gradient_domain_unknown(X, y, inv_pdf[i]);
            for (int k = 0; k < D; ++k) o[4 + k] = g[k];
            o[4 + D] = m.solution_boundary_unknown(x, y, inv_pdf[i]);
        }
    };
    if (dim == 2) run(std::integral_constant<int, 2>{});
    else run(std::integral_constant<int, 3>{});
    return 0;
}

// Domain-representation terms of a model without volume source for domain points xd and
// boundary records yb: per pair [solution_domain_known, solution_domain_unknown,
// gradient_domain_known (dim), gradient_domain_unknown (dim)] (bie_models.cuh:228-250,
// 341-363, 427-478).
int orc_eval_domain_terms(int dim, int bie, int domain, float eps, int n, const float *xd, const float *yb,
                          const float *inv_pdf, float *out) {
    auto run = [&](auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        const int W = 2 * D + 4, O = 2 + 2 * D;
        Model<D> m;
        m.bie = bie; m.interior = domain == InteriorDomain; m.dsign = m.interior ? 1.f : -1.f;
        m.k.eps = eps; m.c = 1.f; m.nvs = 0; m.scene = nullptr;
        for (int i = 0; i < n; ++i) {
            const float *r = yb + (size_t)i * W;
            BPoint<D> y;
            for (int k = 0; k < D; ++k) { y.p[k] = r[k]; y.n[k] = r[D + k]; }
            y.interior_angle = r[2 * D]; y.boundary_value = r[2 * D + 1]; y.boundary_type = (int)r[2 * D + 2];
            y.robin_alpha = r[2 * D + 3];
            Vec<D> X; for (int k = 0; k < D; ++k) X[k] = xd[(size_t)i * D + k];
            Rng dummy;
            float *o = out + (size_t)i * O;
            o[0] = m.solution_domain_known(X, y, inv_pdf[i], dummy);
            o[1] = m.solution_domain_unknown(X, y, inv_pdf[i]);
            const Vec<D> gk = m.gradient_domain_known(X, y, inv_pdf[i], dummy);
            const Vec<D> gu = m.gradient_domain_unknown(X, y, inv_pdf[i]);
            for (int k = 0; k < D; ++k) { o[2 + k] = gk[k]; o[2 + D + k] = gu[k]; }
        }
    };
    if (dim == 2) run(std::integral_constant<int, 2>{});
    else run(std::integral_constant<int, 3>{});
    return 0;
}

// One Scene::sample_boundary_uniform draw (scene.cuh:441-462 / 734-758) from the Philox
// stream (seed, path, point, stream).  out_rec: 2D+4 floats, returns inv_pdf.
// Harmonic control variate: bind m records (s_j, w_j) of dim + 1 floats and c0 (m = 0 unbinds).
int orc_scene_set_harmonic_cv(void *p, unsigned m, const float *src, float c0) {
    auto *S = (OrcScene *)p;
    auto set = [&](auto &s) {
        const int D = S->dim;
        s.hcv.assign(src, src + (size_t)m * (D + 1));
        s.hcv_c0 = m ? c0 : 0.f;
        s.hcv_m = m;
    };
    if (S->dim == 2) set(S->s2); else set(S->s3);
    return 0;
}

// u_h and grad u_h at n points: out n * (dim + 1)
void orc_eval_harmonic_cv(void *p, int n, const float *pts, float *out) {
    auto *S = (OrcScene *)p;
    const int D = S->dim;
    for (int i = 0; i < n; ++i) {
        float g[3];
        const float *x = pts + (size_t)i * D;
        if (D == 2) { out[(size_t)i * 3] = hcv_value(S->s2, x); hcv_gradient(S->s2, x, g); }
        else { out[(size_t)i * 4] = hcv_value(S->s3, x); hcv_gradient(S->s3, x, g); }
        for (int k = 0; k < D; ++k) out[(size_t)i * (D + 1) + 1 + k] = g[k];
    }
}

// Training targets of the harmonic control variate: sample i = the uniform boundary point of
// the Philox stream (seed, path i, point 0, stream), target = its datum plus the walk's
// volume term there with nvs independent volume samples (interior mask), as a BoundaryPoint
// record [p, n, ia, target, type, alpha] (2 dim + 4 floats).
void orc_hcv_targets(void *p, int n, uint64_t seed, uint32_t stream, unsigned nvs, float eps, float *rec) {
    auto *S = (OrcScene *)p;
    auto run = [&](auto &s, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        Model<D> m;
        m.bie = IndirectDoubleLayer; m.interior = true; m.dsign = 1.0f; m.k.eps = eps; m.c = 1.f;
        m.nvs = nvs; m.indep = true; m.scene = &s;
        for (int i = 0; i < n; ++i) {
            Rng rng; rng.init_philox(seed, (uint32_t)i, 0u, stream);
            Sampled<D> y = sample_uniform(s, rng);
            const float t = y.y.boundary_value + m.vol_term(y.y, rng);
            float *r = rec + (size_t)i * (2 * D + 4);
            for (int k = 0; k < D; ++k) { r[k] = y.y.p[k]; r[D + k] = y.y.n[k]; }
            r[2 * D] = y.y.interior_angle; r[2 * D + 1] = t; r[2 * D + 2] = (float)y.y.boundary_type;
            r[2 * D + 3] = y.y.robin_alpha;
        }
    };
    if (S->dim == 2) run(S->s2, std::integral_constant<int, 2>{});
    else run(S->s3, std::integral_constant<int, 3>{});
}

float orc_sample_uniform(void *p, uint64_t seed, uint32_t path, uint32_t point, uint32_t stream, float *out_rec) {
    auto *S = (OrcScene *)p;
    Rng rng; rng.init_philox(seed, path, point, stream);
    auto emit = [&](auto &s, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        Sampled<D> r = sample_uniform(s, rng);
        for (int k = 0; k < D; ++k) { out_rec[k] = r.y.p[k]; out_rec[D + k] = r.y.n[k]; }
        out_rec[2 * D] = r.y.interior_angle; out_rec[2 * D + 1] = r.y.boundary_value;
        out_rec[2 * D + 2] = (float)r.y.boundary_type; out_rec[2 * D + 3] = r.y.robin_alpha;
        return r.inv_pdf;
    };
    if (S->dim == 2) return emit(S->s2, std::integral_constant<int, 2>{});
    return emit(S->s3, std::integral_constant<int, 3>{});
}

// VPL aggregation (apps/potential_flow_with_vpl.cu:44-71): for each point x_i,
//   sum[i][0] = sum_j c_j * solution_domain_unknown(x_i, y_j, inv_pdf),
//   sum[i][1..D] = sum_j c_j * gradient_domain_unknown(x_i, y_j, inv_pdf),
// the products in FP32 as the reference forms them, accumulated in FP64 (the reference
// Kahan-sums in FP32); abssum receives the sums of |term| (the scale of rounding errors).
int orc_vpl_aggregate(int dim, int bie, int domain, float eps, float inv_pdf, int n_vpl, const float *rec,
                      const float *contrib, int n_pts, const float *pts, int threads, double *sum, double *abssum) {
    auto run = [&](auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        const int W = 2 * D + 4;
        Model<D> m;
        m.bie = bie; m.interior = domain == InteriorDomain; m.dsign = m.interior ? 1.f : -1.f;
        m.k.eps = eps; m.c = 1.f; m.nvs = 0; m.scene = nullptr;
        std::vector<BPoint<D>> ys(n_vpl);
        for (int j = 0; j < n_vpl; ++j) {
            const float *r = rec + (size_t)j * W;
            BPoint<D> &b = ys[j];
            for (int k = 0; k < D; ++k) { b.p[k] = r[k]; b.n[k] = r[D + k]; }
            b.interior_angle = r[2 * D]; b.boundary_value = r[2 * D + 1]; b.boundary_type = (int)r[2 * D + 2];
            b.robin_alpha = r[2 * D + 3];
        }
#ifdef _OPENMP
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < n_pts; ++i) {
            Vec<D> x;
            for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
            double acc[D + 1] = {}, aacc[D + 1] = {};
            for (int j = 0; j < n_vpl; ++j) {
                const float c = contrib[j];
                const float u = c * m.solution_domain_unknown(x, ys[j], inv_pdf);
                const Vec<D> g = m.gradient_domain_unknown(x, ys[j], inv_pdf);
                acc[0] += u; aacc[0] += std::fabs(u);
                for (int k = 0; k < D; ++k) {
                    const float gk = c * g[k];
                    acc[1 + k] += gk; aacc[1 + k] += std::fabs(gk);
                }
            }
            for (int k = 0; k <= D; ++k) { sum[(size_t)i * (D + 1) + k] = acc[k]; abssum[(size_t)i * (D + 1) + k] = aacc[k]; }
        }
    };
    if (dim == 2) run(std::integral_constant<int, 2>{});
    else run(std::integral_constant<int, 3>{});
    return 0;
}

}  // extern "C"

// ---- closest points and walk on spheres (apps/simple_wob_and_wos.cu:48-79, 116-141) ----
// Brute force over the elements in caller order: the first element with the smallest squared
// distance wins.  Segment distance: the app's LineSqrD formula; triangle distance: the closest
// point on the triangle by Voronoi regions (Ericson, Real-Time Collision Detection, 5.1.5).
static float seg_dist2(const float *p, const Element<2> &e, float &t) {
    const float l0 = e.v[1][0] - e.v[0][0], l1 = e.v[1][1] - e.v[0][1];
    const float q0 = p[0] - e.v[0][0], q1 = p[1] - e.v[0][1];
    t = (q0 * l0 + q1 * l1) / (l0 * l0 + l1 * l1);
    t = std::min(std::max(t, 0.0f), 1.0f);
    const float e0 = (e.v[0][0] + l0 * t) - p[0], e1 = (e.v[0][1] + l1 * t) - p[1];
    return e0 * e0 + e1 * e1;
}

static float tri_dist2(const float *p, const Element<3> &e, float *w) {
    const float *a = &e.v[0][0], *b = &e.v[1][0], *c = &e.v[2][0];
    float ab[3], ac[3], ap[3], bp[3], cp[3];
    for (int k = 0; k < 3; ++k) { ab[k] = b[k] - a[k]; ac[k] = c[k] - a[k]; ap[k] = p[k] - a[k]; bp[k] = p[k] - b[k]; cp[k] = p[k] - c[k]; }
    auto dot3 = [](const float *x, const float *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    const float d1 = dot3(ab, ap), d2 = dot3(ac, ap);
    float wa, wb, wc;
    if (d1 <= 0.f && d2 <= 0.f) { wa = 1.f; wb = 0.f; wc = 0.f; }
    else {
        const float d3 = dot3(ab, bp), d4 = dot3(ac, bp), d5 = dot3(ab, cp), d6 = dot3(ac, cp);
        const float vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
        if (d3 >= 0.f && d4 <= d3) { wa = 0.f; wb = 1.f; wc = 0.f; }
        else if (d6 >= 0.f && d5 <= d6) { wa = 0.f; wb = 0.f; wc = 1.f; }
        else if (vc <= 0.f && d1 >= 0.f && d3 <= 0.f) { const float s = d1 / (d1 - d3); wa = 1.f - s; wb = s; wc = 0.f; }
        else if (vb <= 0.f && d2 >= 0.f && d6 <= 0.f) { const float s = d2 / (d2 - d6); wa = 1.f - s; wb = 0.f; wc = s; }
        else if (va <= 0.f && (d4 - d3) >= 0.f && (d5 - d6) >= 0.f) {
            const float s = (d4 - d3) / ((d4 - d3) + (d5 - d6)); wa = 0.f; wb = 1.f - s; wc = s;
        } else { const float den = 1.f / (va + vb + vc); wb = vb * den; wc = vc * den; wa = (1.f - wb) - wc; }
    }
    w[0] = wa; w[1] = wb; w[2] = wc;
    float e2 = 0.f;
    for (int k = 0; k < 3; ++k) { const float q = ((wa * a[k] + wb * b[k]) + wc * c[k]) - p[k]; e2 = e2 + q * q; }
    return e2;
}

template <int D> static float closest(const Scene<D> &s, const float *x, int &elem, float &value) {
    float best = std::numeric_limits<float>::infinity();
    elem = -1;
    float bw[3] = {0, 0, 0};
    for (size_t i = 0; i < s.elems.size(); ++i) {
        float w[3] = {0, 0, 0}, d2;
        if constexpr (D == 2) d2 = seg_dist2(x, s.elems[i], w[0]);
        else d2 = tri_dist2(x, s.elems[i], w);
        if (d2 < best) { best = d2; elem = (int)i; bw[0] = w[0]; bw[1] = w[1]; bw[2] = w[2]; }
    }
    const Element<D> &e = s.elems[elem];
    if constexpr (D == 2) value = (e.bv[1] - e.bv[0]) * bw[0] + e.bv[0];
    else value = (bw[0] * e.bv[0] + bw[1] * e.bv[1]) + bw[2] * e.bv[2];
    return best;
}

extern "C" {

int orc_closest_points(void *p, int n, const float *pts, int32_t *elem, float *dist, float *value) {
    auto *S = (OrcScene *)p;
    for (int i = 0; i < n; ++i) {
        int e; float v;
        const float d2 = S->dim == 2 ? closest(S->s2, pts + (size_t)i * 2, e, v) : closest(S->s3, pts + (size_t)i * 3, e, v);
        elem[i] = e; dist[i] = std::sqrt(d2); value[i] = v;
    }
    return 0;
}

// Walk on spheres with the device's streams: path j of point i uses (seed, path_base + j,
// i, stream); a 2D step draws u and moves by R (cos 2 pi u, sin 2 pi u); a 3D step draws z =
// 1 - 2u, then phi = 2 pi u'.  sum / sumsq: n doubles, accumulated into; path_out optional.
int orc_wos(void *p, const float *pts, int n, unsigned paths, uint64_t seed, uint64_t path_base, unsigned stream,
            float eps, unsigned max_steps, int threads, double *sum, double *sumsq, float *path_out) {
    auto *S = (OrcScene *)p;
    auto run = [&](auto &s, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
#ifdef _OPENMP
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for schedule(static)
#endif
        for (int i = 0; i < n; ++i) {
            Kahan ks, kq;
            for (unsigned j = 0; j < paths; ++j) {
                Rng rng; rng.init_philox(seed, (uint32_t)(path_base + j), (uint32_t)i, stream);
                float x[3] = {0, 0, 0};
                for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
                float val = 0.f;
                for (unsigned step = 0;; ++step) {
                    int e; float v;
                    const float R = std::sqrt(closest(s, x, e, v));
                    if (R < eps || step + 1 >= max_steps) { val = v; break; }
                    if constexpr (D == 2) {
                        const float th = 2.0f * rng.uniform();
                        x[0] = x[0] + R * (float)std::cos(th * M_PI);
                        x[1] = x[1] + R * (float)std::sin(th * M_PI);
                    } else {
                        const float z = 1.0f - 2.0f * rng.uniform();
                        const float ph = 2.0f * rng.uniform();
                        const float r = std::sqrt(std::max(1.0f - z * z, 0.0f));
                        x[0] = x[0] + R * (r * (float)std::cos(ph * M_PI));
                        x[1] = x[1] + R * (r * (float)std::sin(ph * M_PI));
                        x[2] = x[2] + R * z;
                    }
                }
                ks.add(val); kq.add(val * val);
                if (path_out) path_out[(size_t)i * paths + j] = val;
            }
            sum[i] += ks.sum; sumsq[i] += kq.sum;
        }
    };
    if (S->dim == 2) run(S->s2, std::integral_constant<int, 2>{});
    else run(S->s3, std::integral_constant<int, 3>{});
    return 0;
}

// WoBFunc (apps/simple_wob_and_wos.cu:81-113), the app's hand-written double-layer walk for
// interior Dirichlet problems: weight m * sign((y - x).n_y) from the first line sample, then
// path_length - 1 bounces multiplying -m * sign((y - y_prev).n_y), the last one halved; each
// visited point adds weight * g(y).  Same draws as the estimators (sample_boundary_line_
// intersection2 per bounce) on the Philox stream (seed, path_base + j, point i, stream).
// sum: n doubles accumulated into; path_out (optional): n * paths floats.
int orc_simple_wob(void *p, const float *pts, int n, unsigned paths, unsigned path_length, uint64_t seed,
                   uint64_t path_base, unsigned stream, int pick_mode, double *sum, float *path_out) {
    auto *S = (OrcScene *)p;
    auto run = [&](auto &s, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        Options opt; opt.pick_mode = pick_mode; opt.use_bvh = true;
        auto sgn = [](float v) { return v > 0 ? 1.0f : -1.0f; };
        for (int i = 0; i < n; ++i) {
            Kahan ks;
            Vec<D> x;
            for (int k = 0; k < D; ++k) x[k] = pts[(size_t)i * D + k];
            for (unsigned j = 0; j < paths; ++j) {
                Rng rng; rng.init_philox(seed, (uint32_t)(path_base + j), (uint32_t)i, stream);
                unsigned m = 0;
                Sampled<D> y = sample_line_intersection(s, x, rng, opt, -1, &m);
                float weight = (float)m * sgn(dot(y.y.p - x, y.y.n));
                float contrib = weight * y.y.boundary_value;
                for (unsigned b = 1; b < path_length; ++b) {
                    if (b == path_length - 1) weight *= 0.5f;
                    const Vec<D> yp = y.y.p;
                    y = sample_line_intersection(s, yp, rng, opt, y.y.elem, &m);
                    weight *= -(float)m * sgn(dot(y.y.p - yp, y.y.n));
                    contrib += weight * y.y.boundary_value;
                }
                ks.add(contrib);
                if (path_out) path_out[(size_t)i * paths + j] = contrib;
            }
            sum[i] += ks.sum;
        }
    };
    if (S->dim == 2) run(S->s2, std::integral_constant<int, 2>{});
    else run(S->s3, std::integral_constant<int, 3>{});
    return 0;
}

// Line-sampling helper for tests: one sample_boundary_line_intersection call from x with a
// Philox stream (seed, path, point, stream).  out_rec: 2D+4 floats, returns inv_pdf.
float orc_sample_line(void *p, const float *x, uint64_t seed, uint32_t path, uint32_t point, uint32_t stream,
                      int pick_mode, float *out_rec) {
    auto *S = (OrcScene *)p;
    Rng rng; rng.init_philox(seed, path, point, stream);
    Options opt; opt.pick_mode = pick_mode; opt.use_bvh = true;
    auto emit = [&](auto &s, auto dimtag) {
        constexpr int D = decltype(dimtag)::value;
        Vec<D> X; for (int k = 0; k < D; ++k) X[k] = x[k];
        Sampled<D> r = sample_line_intersection(s, X, rng, opt);
        for (int k = 0; k < D; ++k) { out_rec[k] = r.y.p[k]; out_rec[D + k] = r.y.n[k]; }
        out_rec[2 * D] = r.y.interior_angle; out_rec[2 * D + 1] = r.y.boundary_value;
        out_rec[2 * D + 2] = (float)r.y.boundary_type; out_rec[2 * D + 3] = r.y.robin_alpha;
        return r.inv_pdf;
    };
    if (S->dim == 2) return emit(S->s2, std::integral_constant<int, 2>{});
    return emit(S->s3, std::integral_constant<int, 3>{});
}

// orc_sample_line for points x_i (n * dim floats) on the streams (seed, path, i, stream):
// records (n * (2 dim + 4) floats) and inverse pdfs (n floats).
void orc_sample_lines(void *p, int n, const float *x, uint64_t seed, uint32_t path, uint32_t stream, int pick_mode,
                      float *out_rec, float *out_inv) {
    auto *S = (OrcScene *)p;
    const int D = S->dim;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i)
        out_inv[i] = orc_sample_line(p, x + (size_t)i * D, seed, path, (uint32_t)i, stream, pick_mode,
                                     out_rec + (size_t)i * (2 * D + 4));
}

}  // extern "C"
