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

// Gradient partials (one block of n_params_padded BF16 values per 64-sample tile; the tile's
// loss sum is an FP32 value in the two slots at n_params) use the flat layout above except
// inside each W_l: there dW_l is stored by 16x16 tile (tile t = row block
// t / 8, column block t % 8, 256 floats each) and, inside a tile, in the MFMA accumulator order
// (lane L's 4 values at 4L .. 4L+3 are rows 4 (L / 16) + j, column L % 16), so that each wave
// writes a tile as one contiguous 1 KB store.  partial_param maps a partial slot to its
// parameter index; biases, the loss slot and padding map to themselves.
__host__ __device__ inline uint32_t partial_param(const MlpParams &P, uint32_t q) {
    for (int l = 0; l < 5; ++l) {
        const uint32_t nout = l == 4 ? 16u : 128u;
        if (q >= P.w_off[l] && q < P.w_off[l] + nout * 128u) {
            const uint32_t r = q - P.w_off[l], t = r >> 8, L = (r & 255u) >> 2, j = r & 3u;
            const uint32_t m = (t >> 3) * 16u + 4u * (L >> 4) + j, n = (t & 7u) * 16u + (L & 15u);
            return P.w_off[l] + m * 128u + n;
        }
    }
    return q;
}

namespace wobml {
size_t smem_bytes();
hipError_t launch_infer(const MlpParams &P, const float *pos, uint32_t n, float *out, uint32_t n_out, hipStream_t st);
hipError_t launch_train_step(const MlpParams &P, const float *pos, const int32_t *ridx, const float *target, uint32_t n,
                             uint16_t *partial, hipStream_t st);
hipError_t launch_adam(const MlpParams &P, const uint16_t *partial, uint32_t n_parts, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, float *master, float *m, float *v, hipStream_t st);
hipError_t launch_loss_sum(const uint16_t *partial, uint32_t n_parts, uint32_t stride, uint32_t off, float *loss,
                           hipStream_t st);
hipError_t set_smem_limits();
}  // namespace wobml
