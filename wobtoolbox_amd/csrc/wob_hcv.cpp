// C ABI of the learned harmonic control variate (include/wob_c.h): a combination of
// fundamental solutions u_h(x) = c0 + sum_j w_j phi(|x - s_j|^2) (phi(z) = z^(-1/2) in 3D,
// log2 z in 2D) with its sources on a sphere (circle) outside the scene (interior fits) or
// inside the body (exterior fits, with c0 = 0), fitted by FP64
// least squares to the walk's own known-term data at uniform boundary samples (the datum
// plus the volume term as the estimator samples it, bie_models.cuh:301-316, 389-399), each
// sample through the datum of its type (u_h, du_h/dn or du_h/dn + alpha u_h), and bound to
// the scene for WOB_EST_HARMONIC_CV walks.  The samples and the FP64 Gram matrix of the normal
// equations are formed on the device; the host scales the columns and solves by Cholesky.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "../../include/wob_c.h"
#include "scene_build.hpp"
#include "wob_capi_internal.hpp"
#include "wob_launch.hpp"

namespace {

constexpr uint32_t kMaxSources = 1024;

inline wob_status fail(wob_status s, const std::string &msg) { return wob_fail(s, msg); }

wob_status check_desc(const wob_scene *scene, const wob_hcv_desc *d) {
    if (!scene || !d) return fail(WOB_INVALID_ARG, "null scene or harmonic control-variate descriptor");
    if (d->n_sources == 0 || d->n_sources > kMaxSources) return fail(WOB_INVALID_ARG, "n_sources must be 1 .. 1024");
    if (d->flags & ~(WOB_HCV_EXTERIOR | WOB_HCV_ZERO_CHARGE))
        return fail(WOB_INVALID_ARG, "unknown harmonic control-variate flags");
    if ((d->flags & WOB_HCV_ZERO_CHARGE) && !(d->flags & WOB_HCV_EXTERIOR))
        return fail(WOB_INVALID_ARG, "WOB_HCV_ZERO_CHARGE needs WOB_HCV_EXTERIOR");
    if ((d->flags & WOB_HCV_ZERO_CHARGE) && d->n_sources < 2)
        return fail(WOB_INVALID_ARG, "a zero-charge fit needs n_sources >= 2");
    if ((d->flags & WOB_HCV_EXTERIOR) ? !(d->radius_scale > 0.f && d->radius_scale < 1.f) : !(d->radius_scale > 1.f))
        return fail(WOB_INVALID_ARG, (d->flags & WOB_HCV_EXTERIOR) ? "radius_scale must be in (0, 1) for an exterior fit"
                                                                    : "radius_scale must be > 1");
    if (d->n_train < 2 * (d->n_sources + 1)) return fail(WOB_INVALID_ARG, "n_train must be >= 2 (n_sources + 1)");
    if (!(d->ridge >= 0.f)) return fail(WOB_INVALID_ARG, "ridge must be >= 0");
    if (d->volume_samples > 0 && wob_scene_dev(scene).vol.kind == wobdev::kVolNone)
        return fail(WOB_INVALID_ARG, "volume_samples > 0 needs a volume source on the scene");
    if (d->volume_samples > 0 && !(d->eps > 0.f)) return fail(WOB_INVALID_ARG, "eps must be > 0");
    return WOB_OK;
}

// Distance from c to segment (a, b) (2D) or triangle (a, b, c3) (3D), in FP64: the closest
// point by the triangle's Voronoi regions (vertex, edge or face), as in Ericson's
// "Real-Time Collision Detection" 5.1.5.
double element_distance(int D, const float *v, const double *c) {
    double a[3] = {0, 0, 0}, b[3] = {0, 0, 0}, t3[3] = {0, 0, 0}, p[3] = {0, 0, 0};
    for (int k = 0; k < D; ++k) { a[k] = v[k]; b[k] = v[D + k]; p[k] = c[k]; if (D == 3) t3[k] = v[6 + k]; }
    auto sub = [](const double *x, const double *y, double *o) { for (int k = 0; k < 3; ++k) o[k] = x[k] - y[k]; };
    auto dot = [](const double *x, const double *y) { return x[0] * y[0] + x[1] * y[1] + x[2] * y[2]; };
    double q[3];
    if (D == 2) {
        double ab[3], ap[3];
        sub(b, a, ab);
        sub(p, a, ap);
        const double l2 = dot(ab, ab);
        const double t = l2 > 0 ? std::min(1.0, std::max(0.0, dot(ap, ab) / l2)) : 0.0;
        for (int k = 0; k < 3; ++k) q[k] = a[k] + t * ab[k];
    } else {
        double ab[3], ac[3], ap[3], bp[3], cp[3];
        sub(b, a, ab); sub(t3, a, ac); sub(p, a, ap);
        const double d1 = dot(ab, ap), d2 = dot(ac, ap);
        sub(p, b, bp);
        const double d3 = dot(ab, bp), d4 = dot(ac, bp);
        sub(p, t3, cp);
        const double d5 = dot(ab, cp), d6 = dot(ac, cp);
        const double vc = d1 * d4 - d3 * d2, vb = d5 * d2 - d1 * d6, va = d3 * d6 - d5 * d4;
        if (d1 <= 0 && d2 <= 0) for (int k = 0; k < 3; ++k) q[k] = a[k];
        else if (d3 >= 0 && d4 <= d3) for (int k = 0; k < 3; ++k) q[k] = b[k];
        else if (d6 >= 0 && d5 <= d6) for (int k = 0; k < 3; ++k) q[k] = t3[k];
        else if (vc <= 0 && d1 >= 0 && d3 <= 0) { const double t = d1 / (d1 - d3); for (int k = 0; k < 3; ++k) q[k] = a[k] + t * ab[k]; }
        else if (vb <= 0 && d2 >= 0 && d6 <= 0) { const double t = d2 / (d2 - d6); for (int k = 0; k < 3; ++k) q[k] = a[k] + t * ac[k]; }
        else if (va <= 0 && d4 - d3 >= 0 && d5 - d6 >= 0) {
            const double t = (d4 - d3) / ((d4 - d3) + (d5 - d6));
            for (int k = 0; k < 3; ++k) q[k] = b[k] + t * (t3[k] - b[k]);
        } else {
            const double den = 1.0 / (va + vb + vc), s = vb * den, t = vc * den;
            for (int k = 0; k < 3; ++k) q[k] = a[k] + s * ab[k] + t * ac[k];
        }
    }
    double r[3];
    sub(p, q, r);
    return std::sqrt(dot(r, r));
}

// Sources on a sphere (Fibonacci lattice) or a circle (equal angles) about the centre c of
// the scene's vertex box.  Interior fits: radius radius_scale * R, R the largest vertex
// distance from c, so every source lies outside the scene.  Exterior fits: radius
// radius_scale * R_in, R_in the smallest distance from c to an element, so the sources lie in
// the ball of radius R_in about c, inside a body that holds that ball.
std::vector<double> source_positions(const wobhost::HostScene &h, uint32_t m, double scale, bool exterior) {
    const int D = (int)h.dim;
    const int aw = D == 2 ? wobhost::kAttrWords2 : wobhost::kAttrWords3;
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (uint32_t e = 0; e < h.n; ++e)
        for (int v = 0; v < D; ++v)
            for (int k = 0; k < D; ++k) {
                const double x = h.elem_attr[(size_t)e * aw + v * D + k];
                lo[k] = std::min(lo[k], x);
                hi[k] = std::max(hi[k], x);
            }
    double c[3] = {0, 0, 0}, R2 = 0, Rin = INFINITY;
    for (int k = 0; k < D; ++k) c[k] = 0.5 * (lo[k] + hi[k]);
    for (uint32_t e = 0; e < h.n; ++e) {
        for (int v = 0; v < D; ++v) {
            double r2 = 0;
            for (int k = 0; k < D; ++k) {
                const double d = h.elem_attr[(size_t)e * aw + v * D + k] - c[k];
                r2 += d * d;
            }
            R2 = std::max(R2, r2);
        }
        if (exterior) Rin = std::min(Rin, element_distance(D, &h.elem_attr[(size_t)e * aw], c));
    }
    const double R = scale * (exterior ? Rin : std::sqrt(R2));
    std::vector<double> s((size_t)m * D);
    const double pi = 3.14159265358979323846;
    for (uint32_t j = 0; j < m; ++j) {
        if (D == 2) {
            const double a = 2 * pi * (j + 0.5) / m;
            s[2 * j] = c[0] + R * std::cos(a);
            s[2 * j + 1] = c[1] + R * std::sin(a);
        } else {
            const double z = 1.0 - 2.0 * (j + 0.5) / m, r = std::sqrt(std::max(0.0, 1.0 - z * z));
            const double ph = pi * (1.0 + std::sqrt(5.0)) * (j + 0.5);
            s[3 * j] = c[0] + R * r * std::cos(ph);
            s[3 * j + 1] = c[1] + R * r * std::sin(ph);
            s[3 * j + 2] = c[2] + R * z;
        }
    }
    return s;
}

// In-place Cholesky solve of the SPD system A x = b (n x n, row-major); false if A is not
// positive definite.
bool cholesky_solve(std::vector<double> &A, std::vector<double> &b, int n) {
    for (int j = 0; j < n; ++j) {
        double d = A[(size_t)j * n + j];
        for (int k = 0; k < j; ++k) d -= A[(size_t)j * n + k] * A[(size_t)j * n + k];
        if (!(d > 0)) return false;
        const double l = std::sqrt(d);
        A[(size_t)j * n + j] = l;
        for (int i = j + 1; i < n; ++i) {
            double v = A[(size_t)i * n + j];
            for (int k = 0; k < j; ++k) v -= A[(size_t)i * n + k] * A[(size_t)j * n + k];
            A[(size_t)i * n + j] = v / l;
        }
    }
    for (int i = 0; i < n; ++i) {
        double v = b[i];
        for (int k = 0; k < i; ++k) v -= A[(size_t)i * n + k] * b[k];
        b[i] = v / A[(size_t)i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double v = b[i];
        for (int k = i + 1; k < n; ++k) v -= A[(size_t)k * n + i] * b[k];
        b[i] = v / A[(size_t)i * n + i];
    }
    return true;
}

wob_status bind_sources(wob_scene *scene, const std::vector<float> &rec, uint32_t m, float c0, int side) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)m * sizeof(float4));
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc harmonic control variate: ") + hipGetErrorString(e));
    e = hipMemcpy(p, rec.data(), (size_t)m * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return wob_hip_fail(e, "hipMemcpy harmonic control variate");
    }
    return wob_scene_bind_hcv(scene, (float4 *)p, m, c0, side);
}

}  // namespace

extern "C" {

wob_status wob_harmonic_cv_targets(const wob_scene *scene, const wob_hcv_desc *d, float *d_records, void *stream) {
    wob_status st = check_desc(scene, d);
    if (st != WOB_OK) return st;
    if (!d_records) return fail(WOB_INVALID_ARG, "null device buffer");
    const wobhost::HostScene &h = wob_scene_host(scene);
    const size_t lds = wobl::walk_lds_bytes(wobdev::FL_UNIFORM, true, 256);
    const float eps = d->volume_samples > 0 ? d->eps : 1.f;
    const bool interior = !(d->flags & WOB_HCV_EXTERIOR);   // the volume term's domain
    hipError_t e = h.dim == 2 ? wobl::launch_hcv_targets_2(wob_scene_dev(scene), d->n_train, d->seed, d->stream_id,
                                                           d->volume_samples, eps, interior, d_records, lds,
                                                           (hipStream_t)stream)
                              : wobl::launch_hcv_targets_3(wob_scene_dev(scene), d->n_train, d->seed, d->stream_id,
                                                           d->volume_samples, eps, interior, d_records, lds,
                                                           (hipStream_t)stream);
    WOB_HIP(e, "harmonic control-variate targets launch");
    return WOB_OK;
}

wob_status wob_scene_fit_harmonic_cv(wob_scene *scene, const wob_hcv_desc *d, float *rms_residual, void *stream) {
    wob_status st = check_desc(scene, d);
    if (st != WOB_OK) return st;
    const wobhost::HostScene &h = wob_scene_host(scene);
    const int D = (int)h.dim, W = 2 * D + 4;
    const uint32_t n = d->n_train, m = d->n_sources, P = m + 1, Q = m + 2;
    const hipStream_t hs = (hipStream_t)stream;
    const bool exterior = (d->flags & WOB_HCV_EXTERIOR) != 0;
    const std::vector<double> src = source_positions(h, m, d->radius_scale, exterior);
    // device buffers: the target records, the source positions, the Gram matrix, residuals
    const size_t rec_bytes = ((size_t)n * W * sizeof(float) + 255) & ~(size_t)255,
                 src_bytes = (src.size() * sizeof(double) + 255) & ~(size_t)255,
                 g_bytes = ((size_t)Q * Q * sizeof(double) + 255) & ~(size_t)255,
                 res_bytes = ((size_t)n * sizeof(float) + 255) & ~(size_t)255;
    char *buf = nullptr;
    if ((st = wob_scene_hcv_workspace(scene, rec_bytes + src_bytes + g_bytes + res_bytes + wobl::hcv_gram_scratch(n, m),
                                      &buf)) != WOB_OK)
        return st;
    float *d_rec = (float *)buf;
    double *d_src = (double *)(buf + rec_bytes), *d_G = (double *)(buf + rec_bytes + src_bytes);
    float *d_res = (float *)(buf + rec_bytes + src_bytes + g_bytes);
    double *d_part = (double *)(buf + rec_bytes + src_bytes + g_bytes + res_bytes);
    std::vector<double> G((size_t)Q * Q);
    hipError_t e = hipMemcpyAsync(d_src, src.data(), src.size() * sizeof(double), hipMemcpyHostToDevice, hs);
    if (e == hipSuccess && (st = wob_harmonic_cv_targets(scene, d, d_rec, stream)) == WOB_OK) {
        e = wobl::launch_hcv_gram(D, d_rec, n, d_src, m, d_part, d_G, hs);
        if (e == hipSuccess) e = hipMemcpyAsync(G.data(), d_G, G.size() * sizeof(double), hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
    }
    if (st != WOB_OK) { return st; }
    if (e != hipSuccess) { return wob_hip_fail(e, "harmonic control-variate Gram matrix"); }

    // normal equations with the columns scaled to unit RMS, and a ridge relative to the
    // mean diagonal.  An exterior fit leaves out column 0 (c0 = 0); a zero-charge fit solves for
    // w_0 .. w_{m-2} with w_{m-1} = -(w_0 + .. + w_{m-2}), i.e. on the columns b_j - b_{m-1}:
    // Gram entries G_ij - G_i,L - G_L,j + G_LL and targets G_iT - G_LT (L the last source).
    const bool zero = (d->flags & WOB_HCV_ZERO_CHARGE) != 0;
    const uint32_t a0 = exterior ? 1u : 0u, P1 = P - a0 - (zero ? 1u : 0u), L = P - 1;
    auto gram = [&](uint32_t i, uint32_t j) {   // unknowns i, j (Gram rows a0 + i, a0 + j)
        const uint32_t gi = a0 + i, gj = a0 + j;
        double v = G[(size_t)gi * Q + gj];
        if (zero) v += -G[(size_t)gi * Q + L] - G[(size_t)L * Q + gj] + G[(size_t)L * Q + L];
        return v;
    };
    auto rhs = [&](uint32_t i) {
        double v = G[(size_t)(a0 + i) * Q + P];
        if (zero) v -= G[(size_t)L * Q + P];
        return v;
    };
    std::vector<double> A((size_t)P1 * P1), rr(P1), scale(P1), r(P, 0.0);
    for (uint32_t a = 0; a < P1; ++a) {
        const double g = gram(a, a);
        scale[a] = g > 0 ? 1.0 / std::sqrt(g / n) : 1.0;
    }
    double tr = 0;
    for (uint32_t a = 0; a < P1; ++a) {
        for (uint32_t b = 0; b < P1; ++b) A[(size_t)a * P1 + b] = gram(a, b) * scale[a] * scale[b];
        rr[a] = rhs(a) * scale[a];
        tr += A[(size_t)a * P1 + a];
    }
    const double lam = (double)d->ridge * tr / P1;
    for (uint32_t a = 0; a < P1; ++a) A[(size_t)a * P1 + a] += lam;
    if (!cholesky_solve(A, rr, (int)P1)) {
        return fail(WOB_INVALID_ARG, "harmonic control-variate fit: singular normal equations (raise ridge, or the "
                                     "samples hold no Dirichlet or Robin datum to fix the constant)");
    }
    for (uint32_t a = 0; a < P1; ++a) r[a + a0] = rr[a] * scale[a];
    if (zero) {
        double sum = 0;
        for (uint32_t a = 0; a < P1; ++a) sum += r[a + a0];
        r[L] = -sum;
    }

    // device records (FP32), bound to the scene
    std::vector<float> out((size_t)m * 4, 0.f);
    for (uint32_t j = 0; j < m; ++j) {
        for (int k = 0; k < D; ++k) out[(size_t)j * 4 + k] = (float)src[(size_t)j * D + k];
        out[(size_t)j * 4 + D] = (float)r[1 + j];
    }
    st = bind_sources(scene, out, m, (float)r[0],
                      !exterior ? wob_hcv_interior : zero ? wob_hcv_exterior_zero_charge : wob_hcv_exterior);
    if (st != WOB_OK) { return st; }
    if (rms_residual) {   // the residual datum of every sample as the walk computes it
        std::vector<float> res(n);
        e = wobl::launch_hcv_residual(D, wob_scene_dev(scene), d_rec, n, d_res, hs);
        if (e == hipSuccess) e = hipMemcpyAsync(res.data(), d_res, n * sizeof(float), hipMemcpyDeviceToHost, hs);
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
        if (e != hipSuccess) { return wob_hip_fail(e, "harmonic control-variate residual"); }
        double ss = 0;
        for (uint32_t i = 0; i < n; ++i) ss += (double)res[i] * res[i];
        *rms_residual = (float)std::sqrt(ss / n);
    }
    return WOB_OK;
}

wob_status wob_scene_set_harmonic_cv(wob_scene *scene, uint32_t n_sources, const float *sources, float c0) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n_sources == 0) return wob_scene_bind_hcv(scene, nullptr, 0, 0.f, wob_hcv_any);
    if (n_sources > kMaxSources) return fail(WOB_INVALID_ARG, "n_sources must be <= 1024");
    if (!sources) return fail(WOB_INVALID_ARG, "null sources");
    const int D = (int)wob_scene_host(scene).dim;
    std::vector<float> rec((size_t)n_sources * 4, 0.f);
    for (uint32_t j = 0; j < n_sources; ++j)
        for (int k = 0; k <= D; ++k) rec[(size_t)j * 4 + k] = sources[(size_t)j * (D + 1) + k];
    return bind_sources(scene, rec, n_sources, c0, wob_hcv_any);
}

wob_status wob_scene_get_harmonic_cv(const wob_scene *scene, uint32_t *n_sources, float *sources, float *c0) {
    if (!scene || !n_sources) return fail(WOB_INVALID_ARG, "null scene or output");
    const wobdev::DevScene &S = wob_scene_dev(scene);
    *n_sources = S.hcv_m;
    if (c0) *c0 = S.hcv_c0;
    if (sources && S.hcv_m) {
        const int D = (int)wob_scene_host(scene).dim;
        std::vector<float> rec((size_t)S.hcv_m * 4);
        WOB_HIP(hipMemcpy(rec.data(), S.hcv_src, rec.size() * sizeof(float), hipMemcpyDeviceToHost),
                "hipMemcpy harmonic control variate");
        for (uint32_t j = 0; j < S.hcv_m; ++j)
            for (int k = 0; k <= D; ++k) sources[(size_t)j * (D + 1) + k] = rec[(size_t)j * 4 + k];
    }
    return WOB_OK;
}

wob_status wob_eval_harmonic_cv(const wob_scene *scene, const float *d_points, uint32_t n, float *d_out, void *stream) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n == 0) return WOB_OK;
    if (!d_points || !d_out) return fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobl::launch_hcv_eval((int)wob_scene_host(scene).dim, wob_scene_dev(scene), n, d_points, d_out,
                                  (hipStream_t)stream),
            "harmonic control-variate evaluation launch");
    return WOB_OK;
}

}  // extern "C"
