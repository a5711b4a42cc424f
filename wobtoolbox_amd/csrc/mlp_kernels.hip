// Fourier-feature MLP for the learned control variate: BF16 MFMA training and inference
// kernels with FP32 accumulation (gfx950, wave64).
//
// Network: gamma(y) = [sin(2 pi B y), cos(2 pi B y)] (64 fixed Gaussian frequencies B, 128
// features) -> 4 x (Linear 128->128, ReLU) -> Linear 128->NO (NO <= 16 outputs, one per
// remaining path length).  Layer l computes Z = H W_l^T + b_l with W_l stored [out][in].
//
// Tiles of 64 samples per workgroup (16 waves: 16 rows and a quarter of the columns each).  Matrix products use
// v_mfma_f32_16x16x16bf16_1k: lane l holds A[l%16][4(l/16)+j], B[4(l/16)+j][l%16] and
// D[4(l/16)+j][l%16], j = 0..3.  Activations live in LDS as BF16; the backward pass keeps
// dZ in FP32 in LDS.  Weight gradients of a tile go to a per-workgroup partial buffer that
// mlp_adam_kernel sums in a fixed order (deterministic) before the Adam update of the FP32
// master weights, which also refreshes the BF16 row-major and transposed copies.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_launch.hpp"

#pragma clang fp contract(fast)

namespace wobm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int T = 64;          // samples per tile
constexpr int W = 128;         // width
constexpr int NF = 64;         // Fourier frequencies
constexpr int NO = 16;         // padded outputs
constexpr int LAYERS = 5;
constexpr int kThreads = 1024;  // 16 waves: wave w owns rows 16 (w % 4) .. +16 and column group w / 4
constexpr int kColGroups = kThreads / 256;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    u += 0x7FFFu + ((u >> 16) & 1u);   // round to nearest even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 4 consecutive bf16 (8 bytes, 8-byte aligned)
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) {
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFF); r[1] = (short)(v.x >> 16); r[2] = (short)(v.y & 0xFFFF); r[3] = (short)(v.y >> 16);
    return r;
}

// Row strides are padded (BF16 rows by 4 elements, FP32 rows by 1) so the 16 rows a
// 16x16x16 MFMA operand spans fall in distinct LDS banks.
constexpr int HS = W + 4;         // BF16 activation / weight row stride
constexpr int ZS = W + 1;         // FP32 dz row stride
constexpr int TS = T + 4;         // row stride of the transposed (feature-major) BF16 tiles
struct Smem {
    uint16_t h[LAYERS][T][HS];    // layer inputs (h[0] = Fourier features), BF16
    float dz[T][ZS];              // backward: dL/dZ of the current layer, FP32
    float out[T][NO];             // network outputs, FP32
    // the current layer's weights (W_l or W_l^T, BF16, row stride cols + 4); during the
    // weight-gradient product the same space holds H_l^T and dZ^T (BF16, row stride TS)
    union {
        uint16_t wbuf[W * HS];
        struct {
            uint16_t hT[W][TS];
            uint16_t dzT[W][TS];
        } tr;
    };
};

// Copy a rows x cols BF16 matrix (row-major, cols a multiple of 4) into LDS with row stride
// cols + 4, 8 bytes per access.
__device__ __forceinline__ void stage(const uint16_t *__restrict__ src, int rows, int cols, uint16_t *dst) {
    const int chunks = rows * cols / 4;
    for (int c = threadIdx.x; c < chunks; c += blockDim.x) {
        const int e = 4 * c, r = e / cols, k = e - r * cols;
        *reinterpret_cast<uint2 *>(&dst[r * (cols + 4) + k]) = *reinterpret_cast<const uint2 *>(&src[e]);
    }
}

__device__ __forceinline__ void features(const MlpParams &P, const float *pos, uint32_t n, uint32_t row0, Smem &s) {
    // thread t: row t / (NF / PER), frequencies (t % (NF / PER)) * PER .. +PER
    constexpr int PER = NF * T / kThreads;
    const int t = threadIdx.x;
    const int r = t / (NF / PER);
    const uint32_t gi = row0 + r;
    float y[3] = {0.f, 0.f, 0.f};
    if (gi < n)
        for (int k = 0; k < (int)P.dim; ++k) y[k] = pos[(size_t)gi * P.dim + k];
    const int j0 = (t % (NF / PER)) * PER;
    for (int j = j0; j < j0 + PER; ++j) {
        const float a = 2.0f * (P.freq[3 * j] * y[0] + P.freq[3 * j + 1] * y[1] + P.freq[3 * j + 2] * y[2]);
        float sn, cs;
        sincospif(a, &sn, &cs);
        s.h[0][r][j] = f2bf(sn);
        s.h[0][r][NF + j] = f2bf(cs);
    }
}

// Z = H_l W_l^T + b_l for this wave's 16 rows; hidden layers store ReLU(Z) as BF16 into
// h[l+1]; the output layer stores Z into out.  W_l is staged in LDS first (all threads take
// part; the caller synchronises after the layer).
__device__ __forceinline__ void forward_layer(const MlpParams &P, int l, Smem &s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int nout = l == LAYERS - 1 ? NO : W;
    stage(P.w_bf16 + P.w_off[l], nout, W, s.wbuf);
    __syncthreads();
    const float *bl = P.b_f32 + P.b_off[l];
    const int r0 = (wave & 3) * 16;
    // column group (wave >> 2); the 16 outputs of the last layer are one tile, on waves 0-3
    const int nhalf = nout > 16 ? nout / kColGroups : 16;
    const int nbeg = nout > 16 ? (wave >> 2) * nhalf : ((wave >> 2) ? nout : 0);
    for (int n0 = nbeg; n0 < nbeg + nhalf && n0 < nout; n0 += 16) {
        f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int k0 = 0; k0 < W; k0 += 16) {
            const bf16x4 a = ld4(&s.h[l][r0 + (lane & 15)][k0 + 4 * (lane >> 4)]);
            const bf16x4 b = ld4(&s.wbuf[(n0 + (lane & 15)) * HS + k0 + 4 * (lane >> 4)]);
            acc = mfma(a, b, acc);
        }
        const int col = n0 + (lane & 15);
        const float bias = bl[col];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int row = r0 + 4 * (lane >> 4) + j;
            const float z = acc[j] + bias;
            if (l == LAYERS - 1) s.out[row][col] = z;
            else s.h[l + 1][row][col] = f2bf(fmaxf(z, 0.f));
        }
    }
}

__global__ __launch_bounds__(kThreads) void mlp_infer_kernel(MlpParams P, const float *pos, uint32_t n, float *out,
                                                        uint32_t n_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem &s = *reinterpret_cast<Smem *>(smem_raw);
    const uint32_t row0 = blockIdx.x * T;
    features(P, pos, n, row0, s);
    __syncthreads();
    for (int l = 0; l < LAYERS; ++l) {
        forward_layer(P, l, s);
        __syncthreads();
    }
    for (int i = threadIdx.x; i < T * (int)n_out; i += blockDim.x) {
        const int r = i / n_out, c = i - r * n_out;
        if (row0 + r < n) out[(size_t)(row0 + r) * n_out + c] = s.out[r][c];
    }
}

// One tile: forward, masked MSE (sample i supervises output r_idx[i]), backward.
// partial layout per workgroup: for each layer l, dW_l (out x in) then db_l (out); then loss.
__global__ __launch_bounds__(kThreads) void mlp_train_kernel(MlpParams P, const float *pos, const int32_t *ridx,
                                                        const float *target, uint32_t n, float inv_batch,
                                                        float *partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem &s = *reinterpret_cast<Smem *>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row0 = blockIdx.x * T;
    features(P, pos, n, row0, s);
    __syncthreads();
    for (int l = 0; l < LAYERS; ++l) {
        forward_layer(P, l, s);
        __syncthreads();
    }
    float *part = partial + (size_t)blockIdx.x * P.n_params_padded;
    // dL/dout (masked MSE, mean over the batch) into dz[:, :NO]; loss partial
    __shared__ float loss_acc[T];
    if (threadIdx.x < T) {
        const int r = threadIdx.x;
        float g = 0.f, lsum = 0.f;
        const uint32_t gi = row0 + r;
        int c = -1;
        if (gi < n) {
            c = min(max(ridx[gi], 0), NO - 1);
            const float e = s.out[r][c] - target[gi];
            g = 2.f * e * inv_batch;
            lsum = e * e;
        }
        for (int j = 0; j < NO; ++j) s.dz[r][j] = j == c ? g : 0.f;
        loss_acc[r] = lsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int r = 0; r < T; ++r) t += loss_acc[r];
        part[P.n_params] = t;
    }
    for (int l = LAYERS - 1; l >= 0; --l) {
        const int nout = l == LAYERS - 1 ? NO : W;
        float *gW = part + P.w_off[l];
        float *gb = part + P.b_goff[l];
        // dW_l = dZ^T H_l  (nout x W), reduction over the 64 rows of the tile.  Its MFMA operands
        // run along the samples, so H_l and dZ (rounded to BF16) are first transposed into
        // feature-major tiles; each operand is then one 8-byte LDS read.
        for (int it = threadIdx.x; it < W * (T / 4); it += blockDim.x) {
            const int f = it % W, r4 = (it / W) * 4;   // feature f, samples r4 .. r4 + 3
            uint2 hv, zv;
            hv.x = (uint32_t)s.h[l][r4][f] | ((uint32_t)s.h[l][r4 + 1][f] << 16);
            hv.y = (uint32_t)s.h[l][r4 + 2][f] | ((uint32_t)s.h[l][r4 + 3][f] << 16);
            *reinterpret_cast<uint2 *>(&s.tr.hT[f][r4]) = hv;
            if (f < nout) {
                zv.x = (uint32_t)f2bf(s.dz[r4][f]) | ((uint32_t)f2bf(s.dz[r4 + 1][f]) << 16);
                zv.y = (uint32_t)f2bf(s.dz[r4 + 2][f]) | ((uint32_t)f2bf(s.dz[r4 + 3][f]) << 16);
                *reinterpret_cast<uint2 *>(&s.tr.dzT[f][r4]) = zv;
            }
        }
        __syncthreads();
        const int tiles = (nout / 16) * (W / 16);
        for (int tix = wave; tix < tiles; tix += kThreads / 64) {
            const int m0 = (tix / (W / 16)) * 16, n0 = (tix % (W / 16)) * 16;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k0 = 0; k0 < T; k0 += 16) {
                const bf16x4 a = ld4(&s.tr.dzT[m0 + (lane & 15)][k0 + 4 * (lane >> 4)]);
                const bf16x4 b = ld4(&s.tr.hT[n0 + (lane & 15)][k0 + 4 * (lane >> 4)]);
                acc = mfma(a, b, acc);
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) gW[(size_t)(m0 + 4 * (lane >> 4) + j) * W + n0 + (lane & 15)] = acc[j];
        }
        // db_l = column sums of dZ
        for (int c = threadIdx.x; c < nout; c += blockDim.x) {
            float t = 0.f;
            for (int r = 0; r < T; ++r) t += s.dz[r][c];
            gb[c] = t;
        }
        if (l == 0) break;
        // dH_l = dZ W_l (T x W), masked by ReLU'(H_l > 0) -> new dz (overwrites after use);
        // W_l^T ([in][out]) is staged in LDS with row stride nout + 4, over the transposed tiles
        __syncthreads();
        stage(P.wt_bf16 + P.w_off[l], W, nout, s.wbuf);
        __syncthreads();
        const uint16_t *WT = s.wbuf;
        constexpr int NT = W / 16 / kColGroups;   // column tiles per wave
        f32x4 accs[NT];
        const int r0 = (wave & 3) * 16, nt0 = (wave >> 2) * NT;
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int n0 = (nt0 + q) * 16;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            for (int k0 = 0; k0 < nout; k0 += 16) {
                bf16x4 a;
#pragma unroll
                for (int j = 0; j < 4; ++j) a[j] = (short)f2bf(s.dz[r0 + (lane & 15)][k0 + 4 * (lane >> 4) + j]);
                const bf16x4 b = ld4(&WT[(n0 + (lane & 15)) * (nout + 4) + k0 + 4 * (lane >> 4)]);
                acc = mfma(a, b, acc);
            }
            accs[q] = acc;
        }
        __syncthreads();   // all reads of dz done
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int col = (nt0 + q) * 16 + (lane & 15);
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int row = r0 + 4 * (lane >> 4) + j;
                s.dz[row][col] = bf2f(s.h[l][row][col]) > 0.f ? accs[q][j] : 0.f;
            }
        }
        __syncthreads();
    }
}

// Sum the per-workgroup partials, Adam-update the FP32 masters, refresh the BF16 copies.
// A workgroup of kAdamSlices waves covers 256 parameters, 4 consecutive ones per lane (16-byte
// loads): wave s sums partials [s*n/S, (s+1)*n/S), and the slice sums are combined in slice
// order, so the result does not depend on scheduling.
constexpr int kAdamSlices = 8;
__device__ __forceinline__ void adam_one(const MlpParams &P, uint32_t p, float g, float lr, float beta1, float beta2,
                                         float eps, float bc1, float bc2, float *master, float *m, float *v) {
    const float mm = beta1 * m[p] + (1.f - beta1) * g;
    const float vv = beta2 * v[p] + (1.f - beta2) * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + eps);
    master[p] = w;
    // locate the parameter: weights of layer l occupy [w_off[l], w_off[l] + out*in), biases
    // of layer l occupy [b_goff[l], b_goff[l] + out) in the same flat index space
    for (int l = 0; l < LAYERS; ++l) {
        const uint32_t nout = l == LAYERS - 1 ? NO : W;
        if (p >= P.w_off[l] && p < P.w_off[l] + nout * W) {
            const uint32_t q = p - P.w_off[l], o = q / W, i = q - o * W;
            const uint16_t h = f2bf(w);
            const_cast<uint16_t *>(P.w_bf16)[P.w_off[l] + q] = h;
            const_cast<uint16_t *>(P.wt_bf16)[P.w_off[l] + (size_t)i * nout + o] = h;
            return;
        }
        if (p >= P.b_goff[l] && p < P.b_goff[l] + nout) {
            const_cast<float *>(P.b_f32)[P.b_off[l] + (p - P.b_goff[l])] = w;
            return;
        }
    }
}

__global__ __launch_bounds__(64 * kAdamSlices) void mlp_adam_kernel(MlpParams P, const float *partial,
                                                                    uint32_t n_parts, float lr, float beta1,
                                                                    float beta2, float eps, float bc1, float bc2,
                                                                    float *master, float *m, float *v) {
    __shared__ float4 red[kAdamSlices][64];
    const uint32_t lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const uint32_t p4 = blockIdx.x * 256 + 4 * lane;   // n_params_padded is a multiple of 64
    const uint32_t k0 = slice * n_parts / kAdamSlices, k1 = (slice + 1) * n_parts / kAdamSlices;
    float4 g = make_float4(0.f, 0.f, 0.f, 0.f);
    if (p4 < P.n_params_padded)
        for (uint32_t k = k0; k < k1; ++k) {
            const float4 t = *reinterpret_cast<const float4 *>(partial + (size_t)k * P.n_params_padded + p4);
            g.x += t.x; g.y += t.y; g.z += t.z; g.w += t.w;
        }
    red[slice][lane] = g;
    __syncthreads();
    if (slice != 0) return;
    float4 tot = red[0][lane];
    for (int q = 1; q < kAdamSlices; ++q) {
        const float4 t = red[q][lane];
        tot.x += t.x; tot.y += t.y; tot.z += t.z; tot.w += t.w;
    }
    const float gs[4] = {tot.x, tot.y, tot.z, tot.w};
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (p4 + j < P.n_params) adam_one(P, p4 + j, gs[j], lr, beta1, beta2, eps, bc1, bc2, master, m, v);
}

__global__ void mlp_loss_sum_kernel(const float *partial, uint32_t n_parts, uint32_t stride, uint32_t off,
                                    float *loss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float t = 0.f;
        for (uint32_t k = 0; k < n_parts; ++k) t += partial[(size_t)k * stride + off];
        *loss = t;
    }
}

}  // namespace wobm

namespace wobml {

size_t smem_bytes() { return sizeof(wobm::Smem); }

hipError_t set_smem_limits() {
    const int b = (int)smem_bytes() + 1024;
    hipError_t e = hipFuncSetAttribute((const void *)wobm::mlp_infer_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, b);
    if (e != hipSuccess) return e;
    return hipFuncSetAttribute((const void *)wobm::mlp_train_kernel, hipFuncAttributeMaxDynamicSharedMemorySize, b);
}

hipError_t launch_infer(const MlpParams &P, const float *pos, uint32_t n, float *out, uint32_t n_out,
                        hipStream_t st) {
    const uint32_t grid = (n + wobm::T - 1) / wobm::T;
    hipLaunchKernelGGL(wobm::mlp_infer_kernel, dim3(grid), dim3(wobm::kThreads), smem_bytes(), st, P, pos, n, out, n_out);
    return hipGetLastError();
}

hipError_t launch_train_step(const MlpParams &P, const float *pos, const int32_t *ridx, const float *target, uint32_t n,
                             float *partial, hipStream_t st) {
    const uint32_t grid = (n + wobm::T - 1) / wobm::T;
    hipLaunchKernelGGL(wobm::mlp_train_kernel, dim3(grid), dim3(wobm::kThreads), smem_bytes(), st, P, pos, ridx, target, n,
                       1.0f / (float)n, partial);
    return hipGetLastError();
}

hipError_t launch_adam(const MlpParams &P, const float *partial, uint32_t n_parts, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, float *master, float *m, float *v, hipStream_t st) {
    const uint32_t block = 64 * wobm::kAdamSlices, grid = (P.n_params + 255) / 256;
    hipLaunchKernelGGL(wobm::mlp_adam_kernel, dim3(grid), dim3(block), 0, st, P, partial, n_parts, lr, beta1, beta2,
                       eps, bc1, bc2, master, m, v);
    return hipGetLastError();
}

hipError_t launch_loss_sum(const float *partial, uint32_t n_parts, uint32_t stride, uint32_t off, float *loss,
                           hipStream_t st) {
    hipLaunchKernelGGL(wobm::mlp_loss_sum_kernel, dim3(1), dim3(64), 0, st, partial, n_parts, stride, off, loss);
    return hipGetLastError();
}

}  // namespace wobml
