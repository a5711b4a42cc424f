// C ABI of the learned control variate (include/wob_c.h): the Fourier-feature MLP (training
// and inference on the BF16 MFMA kernels of mlp_kernels.hip) and its binding to a scene as a
// per-element-vertex value table used by the control-variate walk.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <random>
#include <string>
#include <vector>

#include "../../include/wob_c.h"
#include "mlp_launch.hpp"
#include "scene_build.hpp"
#include "wob_capi_internal.hpp"

namespace {

constexpr uint32_t kWidth = 128, kNF = 64, kNO = 16, kLayers = 5;

uint16_t host_f2bf(float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}

}  // namespace

struct wob_cv {
    uint32_t dim = 0, n_out = 0;
    MlpParams P{};
    std::vector<float> h_master, h_freq;
    float *d_master = nullptr, *d_m = nullptr, *d_v = nullptr, *d_freq = nullptr;
    uint16_t *d_w = nullptr, *d_wt = nullptr;
    uint16_t *d_partial = nullptr;   // BF16 gradient partials, one block per 64-sample tile
    size_t partial_cap = 0;
    float *d_loss = nullptr;
    uint64_t step = 0;
};

namespace {

void cv_free(wob_cv *c) {
    for (void *p : {(void *)c->d_master, (void *)c->d_m, (void *)c->d_v, (void *)c->d_freq, (void *)c->d_w,
                    (void *)c->d_wt, (void *)c->d_partial, (void *)c->d_loss})
        if (p) (void)hipFree(p);
}

// BF16 copies (row-major and transposed) from the FP32 masters, on the host
wob_status cv_upload(wob_cv *c) {
    std::vector<uint16_t> w(c->P.n_params, 0), wt(c->P.n_params, 0);
    for (uint32_t l = 0; l < kLayers; ++l) {
        const uint32_t nout = l == kLayers - 1 ? kNO : kWidth;
        for (uint32_t o = 0; o < nout; ++o)
            for (uint32_t i = 0; i < kWidth; ++i) {
                const uint16_t h = host_f2bf(c->h_master[c->P.w_off[l] + o * kWidth + i]);
                w[c->P.w_off[l] + o * kWidth + i] = h;
                wt[c->P.w_off[l] + i * nout + o] = h;
            }
    }
    WOB_HIP(hipMemcpy(c->d_master, c->h_master.data(), c->P.n_params * 4, hipMemcpyHostToDevice), "hipMemcpy");
    WOB_HIP(hipMemcpy(c->d_w, w.data(), c->P.n_params * 2, hipMemcpyHostToDevice), "hipMemcpy");
    WOB_HIP(hipMemcpy(c->d_wt, wt.data(), c->P.n_params * 2, hipMemcpyHostToDevice), "hipMemcpy");
    WOB_HIP(hipMemcpy(c->d_freq, c->h_freq.data(), kNF * 3 * 4, hipMemcpyHostToDevice), "hipMemcpy");
    return WOB_OK;
}

wob_status ensure_partial(wob_cv *c, uint32_t n) {
    const size_t parts = (n + 63) / 64;
    const size_t bytes = parts * c->P.n_params_padded * sizeof(uint16_t);
    if (bytes <= c->partial_cap) return WOB_OK;
    if (c->d_partial) WOB_HIP(hipFree(c->d_partial), "hipFree");
    c->d_partial = nullptr;
    c->partial_cap = 0;
    if (hipMalloc(&c->d_partial, bytes) != hipSuccess) return wob_fail(WOB_OOM, "hipMalloc MLP gradient partials");
    c->partial_cap = bytes;
    return WOB_OK;
}

}  // namespace

extern "C" {

wob_status wob_cv_create(uint32_t dim, uint32_t n_outputs, float sigma, uint64_t seed, wob_cv **out) {
    if (!out) return wob_fail(WOB_INVALID_ARG, "null output");
    *out = nullptr;
    if (dim != 2 && dim != 3) return wob_fail(WOB_INVALID_ARG, "dim must be 2 or 3");
    if (n_outputs == 0 || n_outputs > kNO) return wob_fail(WOB_INVALID_ARG, "n_outputs must be 1..16");
    if (!(sigma > 0.f)) return wob_fail(WOB_INVALID_ARG, "sigma must be > 0");
    auto *c = new (std::nothrow) wob_cv();
    if (!c) return wob_fail(WOB_OOM, "host allocation failed");
    c->dim = dim;
    c->n_out = n_outputs;
    uint32_t off = 0;
    for (uint32_t l = 0; l < kLayers; ++l) {
        const uint32_t nout = l == kLayers - 1 ? kNO : kWidth;
        c->P.w_off[l] = off;
        off += nout * kWidth;
        c->P.b_off[l] = off;
        c->P.b_goff[l] = off;
        off += nout;
    }
    c->P.dim = dim;
    c->P.n_params = off;
    c->P.n_params_padded = (off + 1 + 63) / 64 * 64;
    // He-uniform weights, zero biases; Gaussian frequencies of scale sigma (all from seed)
    std::mt19937_64 g(seed);
    auto unif = [&]() { return (double)(g() >> 11) * (1.0 / 9007199254740992.0); };
    c->h_master.assign(off, 0.f);
    for (uint32_t l = 0; l < kLayers; ++l) {
        const uint32_t nout = l == kLayers - 1 ? kNO : kWidth;
        const double bound = std::sqrt(6.0 / kWidth) * (l == kLayers - 1 ? 0.1 : 1.0);
        for (uint32_t i = 0; i < nout * kWidth; ++i) c->h_master[c->P.w_off[l] + i] = (float)((2 * unif() - 1) * bound);
    }
    c->h_freq.assign(kNF * 3, 0.f);
    for (uint32_t j = 0; j < kNF; ++j)
        for (uint32_t k = 0; k < 3; ++k) {
            const double u1 = std::max(unif(), 1e-300), u2 = unif();
            c->h_freq[3 * j + k] = k < dim ? (float)(sigma * std::sqrt(-2 * std::log(u1)) * std::cos(2 * M_PI * u2)) : 0.f;
        }
    const size_t np = off;
    if (hipMalloc(&c->d_master, np * 4) != hipSuccess || hipMalloc(&c->d_m, np * 4) != hipSuccess ||
        hipMalloc(&c->d_v, np * 4) != hipSuccess || hipMalloc(&c->d_freq, kNF * 3 * 4) != hipSuccess ||
        hipMalloc(&c->d_w, np * 2) != hipSuccess || hipMalloc(&c->d_wt, np * 2) != hipSuccess ||
        hipMalloc(&c->d_loss, 64) != hipSuccess) {
        cv_free(c);
        delete c;
        return wob_fail(WOB_OOM, "hipMalloc MLP parameters");
    }
    if (hipMemset(c->d_m, 0, np * 4) != hipSuccess || hipMemset(c->d_v, 0, np * 4) != hipSuccess) {
        cv_free(c);
        delete c;
        return wob_fail(WOB_DEVICE, "hipMemset");
    }
    c->P.freq = c->d_freq;
    c->P.w_bf16 = c->d_w;
    c->P.wt_bf16 = c->d_wt;
    c->P.b_f32 = c->d_master;
    wob_status st = cv_upload(c);
    if (st == WOB_OK && wobml::set_smem_limits() != hipSuccess) st = wob_fail(WOB_DEVICE, "hipFuncSetAttribute");
    if (st != WOB_OK) {
        cv_free(c);
        delete c;
        return st;
    }
    *out = c;
    return WOB_OK;
}

wob_status wob_cv_destroy(wob_cv *c) {
    if (!c) return WOB_OK;
    cv_free(c);
    delete c;
    return WOB_OK;
}

wob_status wob_cv_get_info(const wob_cv *c, uint32_t *n_params, uint32_t *n_outputs) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (n_params) *n_params = c->P.n_params;
    if (n_outputs) *n_outputs = c->n_out;
    return WOB_OK;
}

wob_status wob_cv_get_params(const wob_cv *c, float *params, float *freq) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (params) WOB_HIP(hipMemcpy(params, c->d_master, c->P.n_params * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (freq) std::memcpy(freq, c->h_freq.data(), kNF * 3 * 4);
    return WOB_OK;
}

wob_status wob_cv_set_params(wob_cv *c, const float *params, const float *freq) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (params) std::memcpy(c->h_master.data(), params, c->P.n_params * 4);
    else WOB_HIP(hipMemcpy(c->h_master.data(), c->d_master, c->P.n_params * 4, hipMemcpyDeviceToHost), "hipMemcpy");
    if (freq) std::memcpy(c->h_freq.data(), freq, kNF * 3 * 4);
    WOB_HIP(hipMemset(c->d_m, 0, c->P.n_params * 4), "hipMemset");
    WOB_HIP(hipMemset(c->d_v, 0, c->P.n_params * 4), "hipMemset");
    c->step = 0;
    return cv_upload(c);
}

wob_status wob_cv_train(wob_cv *c, const float *d_pos, const int32_t *d_idx, const float *d_target, uint32_t n,
                        uint32_t batch, uint32_t steps, float lr, float *losses, void *stream) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (!d_pos || !d_idx || !d_target || n == 0 || batch == 0) return wob_fail(WOB_INVALID_ARG, "empty training set");
    batch = std::min(batch, n);
    wob_status st = ensure_partial(c, batch);
    if (st != WOB_OK) return st;
    hipStream_t hs = (hipStream_t)stream;
    const uint32_t parts = (batch + 63) / 64;
    for (uint32_t s = 0; s < steps; ++s) {
        uint32_t start = (uint32_t)(((uint64_t)s * batch) % n);
        const uint32_t cnt = std::min(batch, n - start);
        const float frac = steps > 1 ? (float)s / (float)(steps - 1) : 0.f;
        const float lr_s = lr * (0.05f + 0.95f * 0.5f * (1.f + std::cos((float)M_PI * frac)));
        c->step += 1;
        const float bc1 = 1.f - std::pow(0.9f, (float)c->step), bc2 = 1.f - std::pow(0.999f, (float)c->step);
        WOB_HIP(wobml::launch_train_step(c->P, d_pos + (size_t)start * c->dim, d_idx + start, d_target + start, cnt,
                                         c->d_partial, hs),
                "mlp train launch");
        WOB_HIP(wobml::launch_adam(c->P, c->d_partial, (cnt + 63) / 64, lr_s, 0.9f, 0.999f, 1e-8f, bc1, bc2,
                                   c->d_master, c->d_m, c->d_v, hs),
                "mlp adam launch");
        if (losses) {
            WOB_HIP(wobml::launch_loss_sum(c->d_partial, (cnt + 63) / 64, c->P.n_params_padded, c->P.n_params, c->d_loss,
                                           hs),
                    "loss launch");
            float l = 0.f;
            WOB_HIP(hipMemcpyAsync(&l, c->d_loss, 4, hipMemcpyDeviceToHost, hs), "hipMemcpyAsync");
            WOB_HIP(hipStreamSynchronize(hs), "hipStreamSynchronize");
            losses[s] = l / (float)cnt;
        }
        (void)parts;
    }
    return WOB_OK;
}

wob_status wob_cv_gradient(wob_cv *c, const float *d_pos, const int32_t *d_idx, const float *d_target, uint32_t n,
                           float *grad, float *loss, void *stream) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (!d_pos || !d_idx || !d_target || n == 0) return wob_fail(WOB_INVALID_ARG, "empty batch");
    wob_status st = ensure_partial(c, n);
    if (st != WOB_OK) return st;
    hipStream_t hs = (hipStream_t)stream;
    WOB_HIP(wobml::launch_train_step(c->P, d_pos, d_idx, d_target, n, c->d_partial, hs), "mlp train launch");
    WOB_HIP(hipStreamSynchronize(hs), "hipStreamSynchronize");
    const uint32_t parts = (n + 63) / 64;
    std::vector<uint16_t> h((size_t)parts * c->P.n_params_padded);
    WOB_HIP(hipMemcpy(h.data(), c->d_partial, h.size() * 2, hipMemcpyDeviceToHost), "hipMemcpy");
    auto bf2f = [](uint16_t b) {
        const uint32_t u = (uint32_t)b << 16;
        float f;
        std::memcpy(&f, &u, 4);
        return f;
    };
    if (grad)
        for (uint32_t q = 0; q < c->P.n_params; ++q) {
            float g = 0.f;
            for (uint32_t k = 0; k < parts; ++k) g += bf2f(h[(size_t)k * c->P.n_params_padded + q]);
            grad[partial_param(c->P, q)] = g;
        }
    if (loss) {
        float l = 0.f;
        for (uint32_t k = 0; k < parts; ++k) {
            float t;
            std::memcpy(&t, &h[(size_t)k * c->P.n_params_padded + c->P.n_params], 4);
            l += t;
        }
        *loss = l / (float)n;
    }
    return WOB_OK;
}

wob_status wob_cv_infer(wob_cv *c, const float *d_pos, uint32_t n, float *d_out, void *stream) {
    if (!c) return wob_fail(WOB_INVALID_ARG, "null network");
    if (n == 0) return WOB_OK;
    if (!d_pos || !d_out) return wob_fail(WOB_INVALID_ARG, "null device buffer");
    WOB_HIP(wobml::launch_infer(c->P, d_pos, n, d_out, c->n_out, (hipStream_t)stream), "mlp infer launch");
    return WOB_OK;
}

wob_status wob_scene_set_control_variate(wob_scene *scene, wob_cv *c, const float *d_table, uint32_t n_values,
                                         void *stream) {
    if (!scene) return wob_fail(WOB_INVALID_ARG, "null scene");
    const wobhost::HostScene &h = wob_scene_host(scene);
    const uint32_t D = h.dim, n = h.n;
    if (!c && !d_table) return wob_scene_bind_cv_table(scene, nullptr, 0);
    if (n_values == 0 || n_values > kNO) return wob_fail(WOB_INVALID_ARG, "n_values must be 1..16");
    if (c && (c->dim != D || n_values > c->n_out))
        return wob_fail(WOB_INVALID_ARG, "network dimension / output count does not match");
    const size_t tab_floats = (size_t)n_values * n * D;
    float *tab = nullptr;
    if (hipMalloc(&tab, tab_floats * 4) != hipSuccess) return wob_fail(WOB_OOM, "hipMalloc control-variate table");
    hipStream_t hs = (hipStream_t)stream;
    if (!c) {
        hipError_t e = hipMemcpyAsync(tab, d_table, tab_floats * 4, hipMemcpyDeviceToDevice, hs);
        if (e == hipSuccess) e = hipStreamSynchronize(hs);
        if (e != hipSuccess) { (void)hipFree(tab); return wob_hip_fail(e, "hipMemcpy table"); }
        return wob_scene_bind_cv_table(scene, tab, n_values);
    }
    // element vertices (element order), evaluated in one inference launch
    const int AW = D == 2 ? wobhost::kAttrWords2 : wobhost::kAttrWords3;
    std::vector<float> pos((size_t)n * D * D);
    for (uint32_t e = 0; e < n; ++e)
        for (uint32_t j = 0; j < D; ++j)
            for (uint32_t k = 0; k < D; ++k) pos[((size_t)e * D + j) * D + k] = h.elem_attr[(size_t)e * AW + j * D + k];
    float *d_pos = nullptr, *d_out = nullptr;
    if (hipMalloc(&d_pos, pos.size() * 4) != hipSuccess ||
        hipMalloc(&d_out, (size_t)n * D * c->n_out * 4) != hipSuccess) {
        (void)hipFree(tab); (void)hipFree(d_pos);
        return wob_fail(WOB_OOM, "hipMalloc control-variate evaluation buffers");
    }
    std::vector<float> out((size_t)n * D * c->n_out), t(tab_floats);
    hipError_t e = hipMemcpy(d_pos, pos.data(), pos.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess) e = wobml::launch_infer(c->P, d_pos, n * D, d_out, c->n_out, hs);
    if (e == hipSuccess) e = hipStreamSynchronize(hs);
    if (e == hipSuccess) e = hipMemcpy(out.data(), d_out, out.size() * 4, hipMemcpyDeviceToHost);
    if (e == hipSuccess) {
        for (uint32_t r = 0; r < n_values; ++r)
            for (size_t v = 0; v < (size_t)n * D; ++v) t[(size_t)r * n * D + v] = out[v * c->n_out + r];
        e = hipMemcpy(tab, t.data(), tab_floats * 4, hipMemcpyHostToDevice);
    }
    (void)hipFree(d_pos);
    (void)hipFree(d_out);
    if (e != hipSuccess) { (void)hipFree(tab); return wob_hip_fail(e, "control-variate table"); }
    return wob_scene_bind_cv_table(scene, tab, n_values);
}

}  // extern "C"
