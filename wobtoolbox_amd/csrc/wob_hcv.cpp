// C ABI of the learned harmonic control variate (include/wob_c.h): a combination of
// fundamental solutions u_h(x) = c0 + sum_j w_j phi(|x - s_j|^2) (phi(z) = z^(-1/2) in 3D,
// log2 z in 2D) with its sources on a sphere (circle) outside the scene, fitted by FP64
// least squares to the walk's own known-term data at uniform boundary samples (the datum
// plus the volume term as the estimator samples it, bie_models.cuh:301-316, 389-399), each
// sample through the datum of its type (u_h, du_h/dn or du_h/dn + alpha u_h), and bound to
// the scene for WOB_EST_HARMONIC_CV walks.
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
    if (!(d->radius_scale > 1.f)) return fail(WOB_INVALID_ARG, "radius_scale must be > 1");
    if (d->n_train < 2 * (d->n_sources + 1)) return fail(WOB_INVALID_ARG, "n_train must be >= 2 (n_sources + 1)");
    if (!(d->ridge >= 0.f)) return fail(WOB_INVALID_ARG, "ridge must be >= 0");
    if (d->volume_samples > 0 && wob_scene_dev(scene).vol.kind == wobdev::kVolNone)
        return fail(WOB_INVALID_ARG, "volume_samples > 0 needs a volume source on the scene");
    if (d->volume_samples > 0 && !(d->eps > 0.f)) return fail(WOB_INVALID_ARG, "eps must be > 0");
    return WOB_OK;
}

// Sources on a sphere (Fibonacci lattice) or a circle (equal angles) of radius
// radius_scale * R about the centre of the scene's vertex box, R the largest vertex distance
// from that centre; every source lies outside the scene.
std::vector<double> source_positions(const wobhost::HostScene &h, uint32_t m, double scale) {
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
    double c[3] = {0, 0, 0}, R2 = 0;
    for (int k = 0; k < D; ++k) c[k] = 0.5 * (lo[k] + hi[k]);
    for (uint32_t e = 0; e < h.n; ++e)
        for (int v = 0; v < D; ++v) {
            double r2 = 0;
            for (int k = 0; k < D; ++k) {
                const double d = h.elem_attr[(size_t)e * aw + v * D + k] - c[k];
                r2 += d * d;
            }
            R2 = std::max(R2, r2);
        }
    const double R = scale * std::sqrt(R2);
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

inline double phi(int D, const double *x, const double *s) {
    double z = 0;
    for (int k = 0; k < D; ++k) {
        const double d = x[k] - s[k];
        z += d * d;
    }
    return D == 3 ? 1.0 / std::sqrt(z) : std::log2(z);
}

// d phi / dn at x: 3D -(x - s).n |x - s|^-3, 2D (2 / ln 2) (x - s).n / |x - s|^2
inline double dphi_dn(int D, const double *x, const double *s, const double *n) {
    double z = 0, dn = 0;
    for (int k = 0; k < D; ++k) {
        const double d = x[k] - s[k];
        z += d * d;
        dn += d * n[k];
    }
    return D == 3 ? -dn / (z * std::sqrt(z)) : 2.0 / std::log(2.0) * dn / z;
}

// The fit's row of one sample record [p, n, ia, target, type, alpha]: the datum of its type
// of each basis function (1, phi_1 .. phi_m).
void basis_row(int D, const float *rec, const std::vector<double> &src, uint32_t m, std::vector<double> &row) {
    double x[3], n[3];
    for (int k = 0; k < D; ++k) { x[k] = rec[k]; n[k] = rec[D + k]; }
    const int type = (int)rec[2 * D + 2];
    const double alpha = rec[2 * D + 3];
    row[0] = type == WOB_DIRICHLET ? 1.0 : type == WOB_NEUMANN ? 0.0 : alpha;
    for (uint32_t j = 0; j < m; ++j) {
        const double *s = &src[(size_t)j * D];
        if (type == WOB_DIRICHLET) row[1 + j] = phi(D, x, s);
        else if (type == WOB_NEUMANN) row[1 + j] = dphi_dn(D, x, s, n);
        else row[1 + j] = dphi_dn(D, x, s, n) + alpha * phi(D, x, s);
    }
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

wob_status bind_sources(wob_scene *scene, const std::vector<float> &rec, uint32_t m, float c0) {
    void *p = nullptr;
    hipError_t e = hipMalloc(&p, (size_t)m * sizeof(float4));
    if (e != hipSuccess) return fail(WOB_OOM, std::string("hipMalloc harmonic control variate: ") + hipGetErrorString(e));
    e = hipMemcpy(p, rec.data(), (size_t)m * sizeof(float4), hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(p);
        return wob_hip_fail(e, "hipMemcpy harmonic control variate");
    }
    return wob_scene_bind_hcv(scene, (float4 *)p, m, c0);
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
    hipError_t e = h.dim == 2 ? wobl::launch_hcv_targets_2(wob_scene_dev(scene), d->n_train, d->seed, d->stream_id,
                                                           d->volume_samples, eps, true, d_records, lds,
                                                           (hipStream_t)stream)
                              : wobl::launch_hcv_targets_3(wob_scene_dev(scene), d->n_train, d->seed, d->stream_id,
                                                           d->volume_samples, eps, true, d_records, lds,
                                                           (hipStream_t)stream);
    WOB_HIP(e, "harmonic control-variate targets launch");
    return WOB_OK;
}

wob_status wob_scene_fit_harmonic_cv(wob_scene *scene, const wob_hcv_desc *d, float *rms_residual, void *stream) {
    wob_status st = check_desc(scene, d);
    if (st != WOB_OK) return st;
    const wobhost::HostScene &h = wob_scene_host(scene);
    const int D = (int)h.dim, W = 2 * D + 4;
    const uint32_t n = d->n_train, m = d->n_sources, P = m + 1;
    float *buf = nullptr;
    WOB_HIP(hipMalloc((void **)&buf, (size_t)n * W * sizeof(float)), "hipMalloc harmonic control-variate samples");
    std::vector<float> rec((size_t)n * W);
    st = wob_harmonic_cv_targets(scene, d, buf, stream);
    hipError_t e = hipSuccess;
    if (st == WOB_OK) e = hipStreamSynchronize((hipStream_t)stream);
    if (st == WOB_OK && e == hipSuccess) e = hipMemcpy(rec.data(), buf, rec.size() * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(buf);
    if (st != WOB_OK) return st;
    WOB_HIP(e, "harmonic control-variate samples");

    const std::vector<double> src = source_positions(h, m, d->radius_scale);
    // normal equations of the rows, columns scaled to unit RMS
    std::vector<double> row(P), G((size_t)P * P, 0.0), r(P, 0.0), scale(P, 0.0);
    for (uint32_t i = 0; i < n; ++i) {
        basis_row(D, &rec[(size_t)i * W], src, m, row);
        for (uint32_t a = 0; a < P; ++a) scale[a] += row[a] * row[a];
    }
    for (uint32_t a = 0; a < P; ++a) scale[a] = scale[a] > 0 ? 1.0 / std::sqrt(scale[a] / n) : 1.0;
    for (uint32_t i = 0; i < n; ++i) {
        const float *ri = &rec[(size_t)i * W];
        basis_row(D, ri, src, m, row);
        const double t = ri[2 * D + 1];
        for (uint32_t a = 0; a < P; ++a) row[a] *= scale[a];
        for (uint32_t a = 0; a < P; ++a) {
            r[a] += row[a] * t;
            const double ra = row[a];
            double *g = &G[(size_t)a * P];
            for (uint32_t b = 0; b <= a; ++b) g[b] += ra * row[b];
        }
    }
    double tr = 0;
    for (uint32_t a = 0; a < P; ++a) tr += G[(size_t)a * P + a];
    const double lam = (double)d->ridge * tr / P;
    for (uint32_t a = 0; a < P; ++a) {
        G[(size_t)a * P + a] += lam;
        for (uint32_t b = a + 1; b < P; ++b) G[(size_t)a * P + b] = G[(size_t)b * P + a];
    }
    if (!cholesky_solve(G, r, (int)P))
        return fail(WOB_INVALID_ARG, "harmonic control-variate fit: singular normal equations (raise ridge, or the "
                                     "samples hold no Dirichlet or Robin datum to fix the constant)");
    for (uint32_t a = 0; a < P; ++a) r[a] *= scale[a];

    // device records (FP32) and the fit's residual with the coefficients as bound
    std::vector<float> out((size_t)m * 4, 0.f);
    std::vector<double> src32((size_t)m * D), coef(P);
    coef[0] = (float)r[0];
    for (uint32_t j = 0; j < m; ++j) {
        for (int k = 0; k < D; ++k) {
            out[(size_t)j * 4 + k] = (float)src[(size_t)j * D + k];
            src32[(size_t)j * D + k] = out[(size_t)j * 4 + k];
        }
        out[(size_t)j * 4 + D] = (float)r[1 + j];
        coef[1 + j] = out[(size_t)j * 4 + D];
    }
    if (rms_residual) {
        double ss = 0;
        for (uint32_t i = 0; i < n; ++i) {
            const float *ri = &rec[(size_t)i * W];
            basis_row(D, ri, src32, m, row);
            double v = 0;
            for (uint32_t a = 0; a < P; ++a) v += row[a] * coef[a];
            ss += (ri[2 * D + 1] - v) * (ri[2 * D + 1] - v);
        }
        *rms_residual = (float)std::sqrt(ss / n);
    }
    return bind_sources(scene, out, m, (float)r[0]);
}

wob_status wob_scene_set_harmonic_cv(wob_scene *scene, uint32_t n_sources, const float *sources, float c0) {
    if (!scene) return fail(WOB_INVALID_ARG, "null scene");
    if (n_sources == 0) return wob_scene_bind_hcv(scene, nullptr, 0, 0.f);
    if (n_sources > kMaxSources) return fail(WOB_INVALID_ARG, "n_sources must be <= 1024");
    if (!sources) return fail(WOB_INVALID_ARG, "null sources");
    const int D = (int)wob_scene_host(scene).dim;
    std::vector<float> rec((size_t)n_sources * 4, 0.f);
    for (uint32_t j = 0; j < n_sources; ++j)
        for (int k = 0; k <= D; ++k) rec[(size_t)j * 4 + k] = sources[(size_t)j * (D + 1) + k];
    return bind_sources(scene, rec, n_sources, c0);
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
