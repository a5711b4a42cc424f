// Fourier-feature MLP for the learned control variate: BF16 MFMA training and inference
// kernels with FP32 accumulation (gfx950, wave64).
//
// Network: gamma(y) = [sin(2 pi B y), cos(2 pi B y)] (64 fixed Gaussian frequencies B, 128
// features) -> 4 x (Linear 128->128, ReLU) -> Linear 128->NO (NO <= 16 outputs, one per
// remaining path length).  Layer l computes Z = H W_l^T + b_l with W_l stored [out][in].
//
// Tiles of 64 samples per workgroup (16 waves: 16 rows and a quarter of the columns each).
// Matrix products use gfx950's v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)+j],
// B[8(l/16)+j][l%16] (j = 0..7) and D[4(l/16)+i][l%16] (i = 0..3), so every operand is one
// 16-byte LDS read; the one product whose reduction is 16 long (the input gradient through
// the 16-output layer) uses v_mfma_f32_16x16x16bf16_1k.  Activations live in LDS as BF16, the
// backward pass keeps dZ in FP32 in LDS; FP32 -> BF16 is the hardware's round-to-nearest-even
// conversion.  Weight gradients of a tile go to a per-workgroup partial buffer that
// mlp_adam_kernel sums in a fixed order (deterministic) before the Adam update of the FP32
// master weights, which also refreshes the BF16 row-major and transposed copies.  The partials
// are BF16 (rounded to nearest even from the tile's FP32 sums) and summed in FP32: they are the
// largest memory stream of a training step (one 138 KB partial per 64 samples).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "mlp_launch.hpp"

#pragma clang fp contract(fast)

namespace wobm {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

constexpr int T = 64;          // samples per tile
constexpr int W = 128;         // width
constexpr int NF = 64;         // Fourier frequencies
constexpr int NO = 16;         // padded outputs
constexpr int LAYERS = 5;
constexpr int kThreads = 1024;  // 16 waves: wave w owns rows 16 (w % 4) .. +16 and column group w / 4
constexpr int kColGroups = kThreads / 256;

__device__ __forceinline__ uint16_t f2bf(float f) { return __builtin_bit_cast(uint16_t, (__bf16)f); }
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

__device__ __forceinline__ f32x4 mfma32(bf16x8 a, bf16x8 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(bf16x4 a, bf16x4 b, f32x4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// 8 consecutive bf16 (16 bytes, 16-byte aligned)
__device__ __forceinline__ bf16x8 ld8(const uint16_t *p) {
    return __builtin_bit_cast(bf16x8, *reinterpret_cast<const uint4 *>(p));
}
// 8 consecutive FP32 (two 16-byte reads) rounded to bf16
__device__ __forceinline__ bf16x8 ld8f(const float *p) {
    const float4 x = *reinterpret_cast<const float4 *>(p), y = *reinterpret_cast<const float4 *>(p + 4);
    bf16x8 r;
    r[0] = (__bf16)x.x; r[1] = (__bf16)x.y; r[2] = (__bf16)x.z; r[3] = (__bf16)x.w;
    r[4] = (__bf16)y.x; r[5] = (__bf16)y.y; r[6] = (__bf16)y.z; r[7] = (__bf16)y.w;
    return r;
}
__device__ __forceinline__ bf16x4 ld4f(const float *p) {
    const float4 x = *reinterpret_cast<const float4 *>(p);
    bf16x4 r;
    r[0] = (short)f2bf(x.x); r[1] = (short)f2bf(x.y); r[2] = (short)f2bf(x.z); r[3] = (short)f2bf(x.w);
    return r;
}
__device__ __forceinline__ bf16x4 ld4(const uint16_t *p) {
    const uint2 v = *reinterpret_cast<const uint2 *>(p);
    bf16x4 r;
    r[0] = (short)(v.x & 0xFFFF); r[1] = (short)(v.x >> 16); r[2] = (short)(v.y & 0xFFFF); r[3] = (short)(v.y >> 16);
    return r;
}

// Workgroup barrier that orders LDS only.  __syncthreads() also waits for every outstanding
// global load, which would stall on the weights fetched ahead for the next phase; the phases
// here exchange data through LDS alone (global loads are consumed by the thread that issued
// them, global stores are not read back in the kernel).
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// Row strides are padded so that rows start 16 bytes apart in the bank cycle and every
// 16-byte operand read is aligned: BF16 rows by 8 elements, FP32 rows by 4.
constexpr int HS = W + 8;         // BF16 activation / weight row stride
constexpr int ZS = W + 4;         // FP32 dz row stride
constexpr int TS = T + 8;         // row stride of the transposed (feature-major) BF16 tiles
struct Smem {
    uint16_t h[LAYERS][T][HS];    // layer inputs (h[0] = Fourier features), BF16
    float dz[T][ZS];              // backward: dL/dZ of the current layer, FP32
    float out[T][NO];             // network outputs, FP32
    // the current layer's weights (W_l or W_l^T, BF16, row stride cols + 8); during the
    // weight-gradient product the same space holds H_l^T and dZ^T (BF16, row stride TS)
    union {
        uint16_t wbuf[W * HS];
        struct {
            uint16_t hT[W][TS];
            uint16_t dzT[W][TS];
        } tr;
    };
};

// Copy a ROWS x COLS BF16 matrix (row-major) into LDS with row stride COLS + 8, 16 bytes per
// access, in two halves: fetch() issues the global loads into registers, commit() waits for them
// and writes LDS.  Issuing the next matrix's loads right after a commit hides their latency
// behind the work of the phase in between.
constexpr int kFetchChunks = W * W / 8 / kThreads;   // 16-byte chunks per thread, largest matrix
struct Fetch {
    u32x4 v[kFetchChunks];
};
template <int ROWS, int COLS>
__device__ __forceinline__ void fetch(const uint16_t *__restrict__ src, Fetch &f) {
    constexpr int chunks = ROWS * COLS / 8;
#pragma unroll
    for (int i = 0; i < kFetchChunks; ++i) {
        const int c = threadIdx.x + i * kThreads;
        if (c < chunks) f.v[i] = *reinterpret_cast<const u32x4 *>(&src[8 * c]);
    }
}
template <int ROWS, int COLS>
__device__ __forceinline__ void commit(const Fetch &f, uint16_t *dst) {
    constexpr int chunks = ROWS * COLS / 8;
#pragma unroll
    for (int i = 0; i < kFetchChunks; ++i) {
        const int c = threadIdx.x + i * kThreads;
        if (c < chunks) {
            const int e = 8 * c, r = e / COLS, k = e - r * COLS;
            *reinterpret_cast<u32x4 *>(&dst[r * (COLS + 8) + k]) = f.v[i];
        }
    }
}
// The weight matrices in the order the training step uses them: W_0 .. W_4 (forward), then
// W_4^T .. W_1^T (input gradients).  Stage s < LAYERS is W_s, stage s >= LAYERS is
// W_{2 LAYERS - 1 - s}^T.
__device__ __forceinline__ void fetch_stage(const MlpParams &P, int s, Fetch &f) {
    if (s < LAYERS - 1) fetch<W, W>(P.w_bf16 + P.w_off[s], f);
    else if (s == LAYERS - 1) fetch<NO, W>(P.w_bf16 + P.w_off[s], f);
    else if (s == LAYERS) fetch<W, NO>(P.wt_bf16 + P.w_off[LAYERS - 1], f);
    else if (s < 2 * LAYERS - 1) fetch<W, W>(P.wt_bf16 + P.w_off[2 * LAYERS - 1 - s], f);
}
__device__ __forceinline__ void commit_stage(int s, const Fetch &f, uint16_t *dst) {
    if (s < LAYERS - 1) commit<W, W>(f, dst);
    else if (s == LAYERS - 1) commit<NO, W>(f, dst);
    else if (s == LAYERS) commit<W, NO>(f, dst);
    else commit<W, W>(f, dst);
}

// Per-thread inputs of a tile, loaded before the first weight fetch.  Loads complete in issue
// order, so anything the kernel waits for early must be issued before the 32 KB weight fetches.
constexpr int kPer = NF * T / kThreads;   // Fourier frequencies per thread
constexpr int kNQ = W / 16 / kColGroups;  // 16-column tiles per wave in a hidden layer
struct Inputs {
    float y[3];             // position of sample row threadIdx.x / (NF / kPer)
    float fr[3 * kPer];     // frequencies (threadIdx.x % (NF / kPer)) * kPer .. + kPer
    f32x4 bias[LAYERS][kNQ];   // biases of the 4 output features this lane produces per tile
};
// The columns a wave owns in forward layer l: hidden layers, column group (wave >> 2) takes
// kNQ tiles of 16 columns; the 16 outputs of the last layer are one tile, on waves 0-3.
__device__ __forceinline__ int fwd_tiles(int l, int wave) {
    return l == LAYERS - 1 ? ((wave >> 2) == 0 ? 1 : 0) : kNQ;
}
__device__ __forceinline__ int fwd_col0(int l, int wave) { return l == LAYERS - 1 ? 0 : (wave >> 2) * kNQ * 16; }

__device__ __forceinline__ void load_inputs(const MlpParams &P, const float *pos, uint32_t n, uint32_t row0,
                                            Inputs &in) {
    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const uint32_t gi = row0 + t / (NF / kPer);
    in.y[0] = in.y[1] = in.y[2] = 0.f;
    if (gi < n) {
        const float *p = pos + (size_t)gi * P.dim;
        in.y[0] = p[0];
        in.y[1] = p[1];
        if (P.dim == 3) in.y[2] = p[2];
    }
    const int j0 = (t % (NF / kPer)) * kPer;
#pragma unroll
    for (int k = 0; k < 3 * kPer; ++k) in.fr[k] = P.freq[3 * j0 + k];
#pragma unroll
    for (int l = 0; l < LAYERS; ++l) {
        const float *bl = P.b_f32 + P.b_off[l];
        const int nt = fwd_tiles(l, wave), c0 = fwd_col0(l, wave);
#pragma unroll
        for (int q = 0; q < kNQ; ++q)
            in.bias[l][q] = q < nt ? *reinterpret_cast<const f32x4 *>(&bl[c0 + 16 * q + 4 * (lane >> 4)])
                                   : f32x4{0.f, 0.f, 0.f, 0.f};
    }
}

// h[0] = [sin(2 pi B y), cos(2 pi B y)] in BF16
__device__ __forceinline__ void features(const Inputs &in, Smem &s) {
    const int t = threadIdx.x;
    const int r = t / (NF / kPer), j0 = (t % (NF / kPer)) * kPer;
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
        const float a = 2.0f * (in.fr[3 * q] * in.y[0] + in.fr[3 * q + 1] * in.y[1] + in.fr[3 * q + 2] * in.y[2]);
        float sn, cs;
        sincospif(a, &sn, &cs);
        s.h[0][r][j0 + q] = f2bf(sn);
        s.h[0][r][NF + j0 + q] = f2bf(cs);
    }
}

// Z = H_l W_l^T + b_l for this wave's 16 rows; hidden layers store ReLU(Z) as BF16 into
// h[l+1]; the output layer stores Z into out.  The product is formed as Z^T = W_l H_l^T (the
// same operand reads with the MFMA operands swapped), so a lane ends with 4 consecutive output
// features of one sample and writes them with one 8-byte (16-byte for out) LDS store.  W_l arrives in `f` (fetched one phase earlier)
// and is written to LDS first; the loads of stage `next` (if >= 0) are issued right after.  The
// caller synchronises after the layer.
__device__ __forceinline__ void forward_layer(const MlpParams &P, int l, const Inputs &in, Fetch &f, int next,
                                              Smem &s) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const bool last = l == LAYERS - 1;
    commit_stage(l, f, s.wbuf);
    if (next >= 0) fetch_stage(P, next, f);
    const int r0 = (wave & 3) * 16;
    const int ntiles = fwd_tiles(l, wave), nbeg = fwd_col0(l, wave);
    lds_barrier();
    if (ntiles == 0) return;
    f32x4 acc[kNQ];
#pragma unroll
    for (int q = 0; q < kNQ; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int k0 = 0; k0 < W; k0 += 32) {
        const bf16x8 a = ld8(&s.h[l][r0 + (lane & 15)][k0 + 8 * (lane >> 4)]);
#pragma unroll
        for (int q = 0; q < kNQ; ++q)
            if (q < ntiles) {
                const bf16x8 b = ld8(&s.wbuf[(nbeg + 16 * q + (lane & 15)) * HS + k0 + 8 * (lane >> 4)]);
                acc[q] = mfma32(b, a, acc[q]);
            }
    }
#pragma unroll
    for (int q = 0; q < kNQ; ++q)
        if (q < ntiles) {
            // lane: sample r0 + lane % 16, output features col .. col + 3
            const int row = r0 + (lane & 15), col = nbeg + 16 * q + 4 * (lane >> 4);
            const f32x4 z = acc[q] + in.bias[l][q];
            if (last) {
                *reinterpret_cast<f32x4 *>(&s.out[row][col]) = z;
            } else {
                uint2 pk;
                pk.x = (uint32_t)f2bf(fmaxf(z[0], 0.f)) | ((uint32_t)f2bf(fmaxf(z[1], 0.f)) << 16);
                pk.y = (uint32_t)f2bf(fmaxf(z[2], 0.f)) | ((uint32_t)f2bf(fmaxf(z[3], 0.f)) << 16);
                *reinterpret_cast<uint2 *>(&s.h[l + 1][row][col]) = pk;
            }
        }
}

__global__ __launch_bounds__(kThreads) void mlp_infer_kernel(MlpParams P, const float *pos, uint32_t n, float *out,
                                                        uint32_t n_out) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem &s = *reinterpret_cast<Smem *>(smem_raw);
    const uint32_t row0 = blockIdx.x * T;
    Inputs in;
    load_inputs(P, pos, n, row0, in);
    Fetch f;
    fetch_stage(P, 0, f);
    features(in, s);
    lds_barrier();
#pragma unroll
    for (int l = 0; l < LAYERS; ++l) {
        forward_layer(P, l, in, f, l + 1 < LAYERS ? l + 1 : -1, s);
        lds_barrier();
    }
    for (int i = threadIdx.x; i < T * (int)n_out; i += blockDim.x) {
        const int r = i / n_out, c = i - r * n_out;
        if (row0 + r < n) out[(size_t)(row0 + r) * n_out + c] = s.out[r][c];
    }
}

// One tile: forward, masked MSE (sample i supervises output r_idx[i]), backward.
// partial layout per workgroup: for each layer l, dW_l (tile order, see partial_param) then db_l
// (out); then the loss.
__global__ __launch_bounds__(kThreads) void mlp_train_kernel(MlpParams P, const float *pos, const int32_t *ridx,
                                                        const float *target, uint32_t n, float inv_batch,
                                                        uint16_t *partial) {
    extern __shared__ __align__(16) unsigned char smem_raw[];
    Smem &s = *reinterpret_cast<Smem *>(smem_raw);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t row0 = blockIdx.x * T;
    Inputs in;
    load_inputs(P, pos, n, row0, in);
    // supervision of sample row threadIdx.x (threads < T)
    int sup_c = -1;
    float sup_t = 0.f;
    if (threadIdx.x < T && row0 + threadIdx.x < n) {
        sup_c = min(max(ridx[row0 + threadIdx.x], 0), NO - 1);
        sup_t = target[row0 + threadIdx.x];
    }
    Fetch f;
    fetch_stage(P, 0, f);
    features(in, s);
    lds_barrier();
#pragma unroll
    for (int l = 0; l < LAYERS; ++l) {
        forward_layer(P, l, in, f, l + 1, s);
        lds_barrier();
    }
    uint16_t *part = partial + (size_t)blockIdx.x * P.n_params_padded;
    // dL/dout (masked MSE, mean over the batch) into dz[:, :NO]; loss partial
    __shared__ float loss_acc[T];
    if (threadIdx.x < T) {
        const int r = threadIdx.x;
        float g = 0.f, lsum = 0.f;
        const int c = sup_c;
        if (c >= 0) {
            const float e = s.out[r][c] - sup_t;
            g = 2.f * e * inv_batch;
            lsum = e * e;
        }
        for (int j = 0; j < NO; ++j) s.dz[r][j] = j == c ? g : 0.f;
        loss_acc[r] = lsum;
    }
    lds_barrier();
    if (threadIdx.x == 0) {
        float t = 0.f;
        for (int r = 0; r < T; ++r) t += loss_acc[r];
        *reinterpret_cast<float *>(part + P.n_params) = t;
    }
#pragma unroll
    for (int l = LAYERS - 1; l >= 0; --l) {
        const int nout = l == LAYERS - 1 ? NO : W;
        uint16_t *gW = part + P.w_off[l];
        uint16_t *gb = part + P.b_goff[l];
        // dW_l = dZ^T H_l  (nout x W), reduction over the 64 rows of the tile.  Its MFMA operands
        // run along the samples, so H_l and dZ (rounded to BF16) are first transposed into
        // feature-major tiles; each operand is then one 16-byte LDS read.  The same pass sums
        // dZ over groups of 8 samples in FP32 (into the output area, free during the backward
        // pass) for db_l.
        float(*dbp)[W] = reinterpret_cast<float(*)[W]>(&s.out[0][0]);   // [T / 8][W]
        for (int it = threadIdx.x; it < W * (T / 8); it += blockDim.x) {
            const int ft = it % W, g8 = it / W, r8 = g8 * 8;   // feature ft, samples r8 .. r8 + 7
            uint4 hv;
            hv.x = (uint32_t)s.h[l][r8][ft] | ((uint32_t)s.h[l][r8 + 1][ft] << 16);
            hv.y = (uint32_t)s.h[l][r8 + 2][ft] | ((uint32_t)s.h[l][r8 + 3][ft] << 16);
            hv.z = (uint32_t)s.h[l][r8 + 4][ft] | ((uint32_t)s.h[l][r8 + 5][ft] << 16);
            hv.w = (uint32_t)s.h[l][r8 + 6][ft] | ((uint32_t)s.h[l][r8 + 7][ft] << 16);
            *reinterpret_cast<uint4 *>(&s.tr.hT[ft][r8]) = hv;
            if (ft < nout) {
                float z[8];
#pragma unroll
                for (int j = 0; j < 8; ++j) z[j] = s.dz[r8 + j][ft];
                uint4 zv;
                zv.x = (uint32_t)f2bf(z[0]) | ((uint32_t)f2bf(z[1]) << 16);
                zv.y = (uint32_t)f2bf(z[2]) | ((uint32_t)f2bf(z[3]) << 16);
                zv.z = (uint32_t)f2bf(z[4]) | ((uint32_t)f2bf(z[5]) << 16);
                zv.w = (uint32_t)f2bf(z[6]) | ((uint32_t)f2bf(z[7]) << 16);
                *reinterpret_cast<uint4 *>(&s.tr.dzT[ft][r8]) = zv;
                dbp[g8][ft] = ((z[0] + z[1]) + (z[2] + z[3])) + ((z[4] + z[5]) + (z[6] + z[7]));
            }
        }
        lds_barrier();
        {
            // 4 (hidden layers) or 1 (output layer) 16x16 tiles of dW per wave, interleaved
            constexpr int NQ = W * W / 256 / (kThreads / 64);
            const int ntile = (nout / 16) * (W / 16);
            f32x4 acc[NQ];
#pragma unroll
            for (int q = 0; q < NQ; ++q) acc[q] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int k0 = 0; k0 < T; k0 += 32) {
#pragma unroll
                for (int q = 0; q < NQ; ++q) {
                    const int tix = wave + q * (kThreads / 64);
                    if (tix < ntile) {
                        const int m0 = (tix / (W / 16)) * 16, n0 = (tix % (W / 16)) * 16;
                        const bf16x8 a = ld8(&s.tr.dzT[m0 + (lane & 15)][k0 + 8 * (lane >> 4)]);
                        const bf16x8 b = ld8(&s.tr.hT[n0 + (lane & 15)][k0 + 8 * (lane >> 4)]);
                        acc[q] = mfma32(a, b, acc[q]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < NQ; ++q) {
                const int tix = wave + q * (kThreads / 64);
                // tile-order partial layout (partial_param): one contiguous 512-byte store per wave
                if (tix < ntile) {
                    uint2 pk;
                    pk.x = (uint32_t)f2bf(acc[q][0]) | ((uint32_t)f2bf(acc[q][1]) << 16);
                    pk.y = (uint32_t)f2bf(acc[q][2]) | ((uint32_t)f2bf(acc[q][3]) << 16);
                    *reinterpret_cast<uint2 *>(gW + 256 * tix + 4 * lane) = pk;
                }
            }
        }
        // db_l: the 8 group sums of each column, in group order
        for (int c = threadIdx.x; c < nout; c += blockDim.x) {
            float t = 0.f;
#pragma unroll
            for (int g = 0; g < T / 8; ++g) t += dbp[g][c];
            gb[c] = f2bf(t);
        }
        if (l == 0) break;
        // dH_l = dZ W_l (T x W), masked by ReLU'(H_l > 0) -> new dz (overwrites after use);
        // W_l^T ([in][out]) goes to LDS with row stride nout + 8, over the transposed tiles.  As
        // in the forward pass the MFMA forms the transpose, dH_l^T = W_l^T dZ^T, so a lane holds
        // 4 consecutive input features of one sample.
        lds_barrier();
        const int st = 2 * LAYERS - 1 - l;
        commit_stage(st, f, s.wbuf);
        if (st + 1 < 2 * LAYERS - 1) fetch_stage(P, st + 1, f);
        lds_barrier();
        const uint16_t *WT = s.wbuf;
        constexpr int NT = W / 16 / kColGroups;   // column tiles per wave
        f32x4 accs[NT];
        const int r0 = (wave & 3) * 16, nt0 = (wave >> 2) * NT;
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int n0 = (nt0 + q) * 16;
            f32x4 acc = {0.f, 0.f, 0.f, 0.f};
            if (nout == NO) {   // a 16-long reduction: the 16x16x16 instruction
                const bf16x4 a = ld4f(&s.dz[r0 + (lane & 15)][4 * (lane >> 4)]);
                const bf16x4 b = ld4(&WT[(n0 + (lane & 15)) * (NO + 8) + 4 * (lane >> 4)]);
                acc = mfma16(b, a, acc);
            } else {
#pragma unroll
                for (int k0 = 0; k0 < W; k0 += 32) {
                    const bf16x8 a = ld8f(&s.dz[r0 + (lane & 15)][k0 + 8 * (lane >> 4)]);
                    const bf16x8 b = ld8(&WT[(n0 + (lane & 15)) * (W + 8) + k0 + 8 * (lane >> 4)]);
                    acc = mfma32(b, a, acc);
                }
            }
            accs[q] = acc;
        }
        lds_barrier();   // all reads of dz done
#pragma unroll
        for (int q = 0; q < NT; ++q) {
            const int row = r0 + (lane & 15), col = (nt0 + q) * 16 + 4 * (lane >> 4);
            const uint2 hv = *reinterpret_cast<const uint2 *>(&s.h[l][row][col]);
            f32x4 d = accs[q];
            d[0] = __uint_as_float(hv.x << 16) > 0.f ? d[0] : 0.f;
            d[1] = __uint_as_float(hv.x & 0xFFFF0000u) > 0.f ? d[1] : 0.f;
            d[2] = __uint_as_float(hv.y << 16) > 0.f ? d[2] : 0.f;
            d[3] = __uint_as_float(hv.y & 0xFFFF0000u) > 0.f ? d[3] : 0.f;
            *reinterpret_cast<f32x4 *>(&s.dz[row][col]) = d;
        }
        lds_barrier();
    }
}

// Sum the per-workgroup partials, Adam-update the FP32 masters, refresh the BF16 copies.
// A workgroup of kAdamSlices waves covers 256 partial slots, 4 consecutive ones per lane (16-byte
// loads): wave s sums partials [s*n/S, (s+1)*n/S), the slice sums are combined in slice order
// (so the result does not depend on scheduling), and 256 threads then update one parameter each.
constexpr int kAdamSlices = 16;
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

// At most 64 VGPRs (8 waves per SIMD): two workgroups fit on a CU, so the ~270 workgroups of
// the network run in a single round.
__global__ __launch_bounds__(64 * kAdamSlices) __attribute__((amdgpu_waves_per_eu(8))) void mlp_adam_kernel(
    MlpParams P, const uint16_t *partial, uint32_t n_parts, float lr, float beta1, float beta2, float eps, float bc1,
    float bc2, float *master, float *m, float *v) {
    __shared__ float red[kAdamSlices][256];
    const uint32_t lane = threadIdx.x & 63, slice = threadIdx.x >> 6;
    const uint32_t p4 = blockIdx.x * 256 + 4 * lane;   // n_params_padded is a multiple of 64
    const uint32_t k0 = slice * n_parts / kAdamSlices, k1 = (slice + 1) * n_parts / kAdamSlices;
    f32x4 g = {0.f, 0.f, 0.f, 0.f};
    if (p4 < P.n_params) {   // the FP32 loss slots at n_params are not gradients
        const uint16_t *src = partial + (size_t)k0 * P.n_params_padded + p4;
#pragma unroll 8
        for (uint32_t k = k0; k < k1; ++k, src += P.n_params_padded) {
            const uint2 v = *reinterpret_cast<const uint2 *>(src);
            g += f32x4{__uint_as_float(v.x << 16), __uint_as_float(v.x & 0xFFFF0000u), __uint_as_float(v.y << 16),
                       __uint_as_float(v.y & 0xFFFF0000u)};
        }
    }
    *reinterpret_cast<f32x4 *>(&red[slice][4 * lane]) = g;
    __syncthreads();
    if (threadIdx.x >= 256) return;
    const uint32_t q = blockIdx.x * 256 + threadIdx.x;
    if (q >= P.n_params) return;
    float tot = red[0][threadIdx.x];
    for (int k = 1; k < kAdamSlices; ++k) tot += red[k][threadIdx.x];
    adam_one(P, partial_param(P, q), tot, lr, beta1, beta2, eps, bc1, bc2, master, m, v);
}

__global__ void mlp_loss_sum_kernel(const uint16_t *partial, uint32_t n_parts, uint32_t stride, uint32_t off,
                                    float *loss) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        float t = 0.f;
        for (uint32_t k = 0; k < n_parts; ++k) t += *reinterpret_cast<const float *>(partial + (size_t)k * stride + off);
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
                             uint16_t *partial, hipStream_t st) {
    const uint32_t grid = (n + wobm::T - 1) / wobm::T;
    hipLaunchKernelGGL(wobm::mlp_train_kernel, dim3(grid), dim3(wobm::kThreads), smem_bytes(), st, P, pos, ridx, target, n,
                       1.0f / (float)n, partial);
    return hipGetLastError();
}

hipError_t launch_adam(const MlpParams &P, const uint16_t *partial, uint32_t n_parts, float lr, float beta1, float beta2,
                       float eps, float bc1, float bc2, float *master, float *m, float *v, hipStream_t st) {
    const uint32_t block = 64 * wobm::kAdamSlices, grid = (P.n_params + 255) / 256;
    hipLaunchKernelGGL(wobm::mlp_adam_kernel, dim3(grid), dim3(block), 0, st, P, partial, n_parts, lr, beta1, beta2,
                       eps, bc1, bc2, master, m, v);
    return hipGetLastError();
}

hipError_t launch_loss_sum(const uint16_t *partial, uint32_t n_parts, uint32_t stride, uint32_t off, float *loss,
                           hipStream_t st) {
    hipLaunchKernelGGL(wobm::mlp_loss_sum_kernel, dim3(1), dim3(64), 0, st, partial, n_parts, stride, off, loss);
    return hipGetLastError();
}

}  // namespace wobml
