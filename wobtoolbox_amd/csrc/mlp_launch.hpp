// Launch interface of the Fourier-feature MLP kernels (mlp_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

// Flat FP32 parameter space (masters, Adam moments, gradients): layer l holds W_l
// (out x in, row-major) at w_off[l] followed by b_l (out) at b_off[l].  The BF16 copies
// w_bf16 / wt_bf16 use the same offsets for W_l / W_l^T (in x out); biases are read from
// the FP32 masters.
struct MlpParams {
    uint32_t dim;               // 2 or 3
    uint32_t n_params;          // flat parameter count
    uint32_t n_params_padded;   // per-workgroup partial stride (n_params + loss, padded)
    uint32_t w_off[5];
    uint32_t b_off[5];
    uint32_t b_goff[5];         // == b_off (gradient partials use the same layout)
    const float *freq;          // 64 x 3 Fourier frequencies
    const uint16_t *w_bf16;
    const uint16_t *wt_bf16;
    const float *b_f32;         // FP32 masters (biases read from here)
};

namespace wobml {
size_t smem_bytes();
hipError_t launch_infer(const MlpParams &P, const float *pos, uint32_t n, float *out, uint32_t n_out, hipStream_t st);
hipError_t launch_train_step(const MlpParams &P, const float *pos, const int32_t *ridx, const float *target, uint32_t n,
                             float *partial, hipStream_t st);
hipError_t launch_adam(const MlpParams &P, const float *partial, uint32_t n_parts, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, float *master, float *m, float *v, hipStream_t st);
hipError_t launch_loss_sum(const float *partial, uint32_t n_parts, uint32_t stride, uint32_t off, float *loss,
                           hipStream_t st);
hipError_t set_smem_limits();
}  // namespace wobml
