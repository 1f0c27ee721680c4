// erfx_mlp.hip -- the learned stand-in for Kessler's cell tendencies: an
// 8 -> 128 -> 128 -> 4 multilayer perceptron with GELU activations, BF16
// weights and activations and FP32 accumulation, trained on the GPU by the
// kernels below (forward, backward, Adam).  It has no counterpart in ERF; the
// plain Kessler scheme (Source/Microphysics/Kessler/ERF_Kessler.cpp:11-250)
// produces its training data and is its fallback and the reference for its error.
//
// Inputs (T, p [mbar], rho, qv, qc, qr, w, dt), normalised per feature;
// outputs the local process tendencies of one step (dqv, dqc, dqr, dtheta,
// sedimentation excluded), scaled per output.
//
// Matrix work runs on the CDNA4 matrix cores: v_mfma_f32_32x32x16_bf16 (one
// wave computes a 32x32 FP32 tile from 32x16 and 16x32 BF16 operands).
// Operand layouts of that instruction for lane l of a 64-wide wave:
//   A (32 x 16): row l % 32, columns 8 (l / 32) .. 8 (l / 32) + 7
//   B (16 x 32): column l % 32, rows 8 (l / 32) .. 8 (l / 32) + 7
//   D (32 x 32): column l % 32, rows 8 (i / 4) + 4 (l / 32) + i % 4, i = 0..15
// The forward pass runs all three layers on the matrix cores (the input and
// output layers zero-padded to the instruction's K = 16 and N = 32); the
// backward pass runs the hidden layer's two products there and the 8-wide and
// 4-wide layers on the vector ALUs.
#include "erfx_common.h"
#include "erfx_internal.h"
#include "erfx_mlp.h"

namespace erfx {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}
__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }

// LDS row pitch of a 32 x 128 BF16 tile: 8 elements of padding keep the
// 16-byte operand reads of 32 rows on distinct banks
constexpr int TP = NH + 8;

// One wave: D (32 x 128, four 32 x 32 tiles) = A (32 x 128 from LDS, row
// pitch TP) times B (128 x 128) whose column n is stored contiguously in k
// (bt[n * NH + k]).
__device__ __forceinline__ void wave_gemm_32x128x128(const __bf16* a_lds, const __bf16* bt, floatx16 acc[4]) {
    const int l = threadIdx.x & 63;
    const int r = l & 31, kh = (l >> 5) * 8;
#pragma unroll
    for (int t = 0; t < 4; ++t)
        for (int i = 0; i < 16; ++i) acc[t][i] = 0.0f;
#pragma unroll 2
    for (int c = 0; c < NH / 16; ++c) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(a_lds + r * TP + 16 * c + kh);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(bt + (32 * t + r) * NH + 16 * c + kh);
            acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[t], 0, 0, 0);
        }
    }
}
// row of accumulator element i of lane l in a 32 x 32 tile
__device__ __forceinline__ int drow(int l, int i) { return 8 * (i >> 2) + 4 * (l >> 5) + (i & 3); }

// Forward pass of 32 samples per wave, WPB waves per block.  x: normalised
// inputs [n][NI]; y: scaled outputs [n][NO].  When `train` the activations the
// backward pass needs are stored: h1 (bf16), z1 and z2 pre-activations (fp32),
// h2 (bf16).  All three layers run on the matrix cores: the input layer as one
// 32x32x16 product per 32 units (the 8 inputs in the first half of K, zeros in
// the second), the output layer as eight K-steps into one 32 x 32 tile of
// which the first 4 columns are the outputs; the small weights and the biases
// are staged in LDS once per block.  The GELUs stay on the vector ALUs (two
// per unit and sample: they, not the products, bound the kernel).
template <bool train, int WPB>
__global__ void __launch_bounds__(64 * WPB) fwd_kernel(const float* __restrict__ x, int n, const Weights W,
                                                        float* __restrict__ y, TrainBuf T) {
    __shared__ __bf16 tile[WPB][32 * TP];
    __shared__ __bf16 sw1[NI * NH];   // [k][u]
    __shared__ __bf16 sw3[NH * NO];   // [u][o]
    __shared__ float sb1[NH], sb2[NH], sb3[NO];
    for (int e = threadIdx.x; e < NI * NH; e += 64 * WPB) sw1[e] = W.w1[e];
    for (int e = threadIdx.x; e < NH * NO; e += 64 * WPB) sw3[e] = W.w3[e];
    for (int e = threadIdx.x; e < NH; e += 64 * WPB) { sb1[e] = W.b1[e]; sb2[e] = W.b2[e]; }
    if (threadIdx.x < NO) sb3[threadIdx.x] = W.b3[threadIdx.x];
    __syncthreads();
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int s0 = (blockIdx.x * WPB + wv) * 32;   // waves past n run on dead lanes (no early exit: block barriers)
    __bf16* h = tile[wv];
    const int cell = l & 31, half = l >> 5;
    const int s = s0 + cell;
    const bool live = s < n;
    const __bf16 zero = f2bf(0.0f);
    // ---- input layer: A = x (sample `cell`, inputs 0..7 in lanes 0..31), B = W1
    bf16x8 xa;
#pragma unroll
    for (int q = 0; q < 8; ++q) xa[q] = (half == 0 && live) ? f2bf(x[(long)s * NI + q]) : zero;
    floatx16 acc[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        bf16x8 wb;
#pragma unroll
        for (int q = 0; q < 8; ++q) wb[q] = half == 0 ? sw1[q * NH + 32 * t + cell] : zero;
        floatx16 z0;
#pragma unroll
        for (int i = 0; i < 16; ++i) z0[i] = 0.0f;
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(xa, wb, z0, 0, 0, 0);
    }
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int u = 32 * t + cell;
        const float b = sb1[u];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = drow(l, i);
            const float z = acc[t][i] + b;
            const __bf16 hv = f2bf(gelu(z));
            h[row * TP + u] = hv;
            const int sr = s0 + row;
            if (train && sr < n) {
                T.z1[(long)sr * NH + u] = z;
                T.h1[(long)sr * NH + u] = hv;
            }
        }
    }
    __syncthreads();
    // ---- hidden layer
    wave_gemm_32x128x128(h, W.w2t, acc);
    __syncthreads();
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int u = 32 * t + cell;
        const float b = sb2[u];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int row = drow(l, i);
            const float z = acc[t][i] + b;
            const __bf16 hv = f2bf(gelu(z));
            h[row * TP + u] = hv;
            const int sr = s0 + row;
            if (train && sr < n) {
                T.z2[(long)sr * NH + u] = z;
                T.h2[(long)sr * NH + u] = hv;
            }
        }
    }
    __syncthreads();
    // ---- output layer: column o = cell of the tile is output o for cell < NO
    floatx16 ao;
#pragma unroll
    for (int i = 0; i < 16; ++i) ao[i] = 0.0f;
    const int kh = half * 8;
    const int oc = cell < NO ? cell : NO - 1;   // lanes of the padding columns read a valid entry, then use 0
#pragma unroll
    for (int c = 0; c < NH / 16; ++c) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(h + cell * TP + 16 * c + kh);
        bf16x8 wb;
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            const __bf16 w = sw3[(16 * c + kh + q) * NO + oc];
            wb[q] = cell < NO ? w : zero;
        }
        ao = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, wb, ao, 0, 0, 0);
    }
    if (cell < NO) {
        const float b = sb3[cell];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int sr = s0 + drow(l, i);
            if (sr < n) y[(long)sr * NO + cell] = ao[i] + b;
        }
    }
}

// Backward through the output and hidden layers for 32 samples per wave:
// dy = d(mean squared error)/dy; dz2 = (dy W3^T) * gelu'(z2) (stored bf16);
// dz1 = (dz2 W2^T) * gelu'(z1) (matrix cores; stored fp32).  loss_part gets
// the per-sample squared error.
__global__ void __launch_bounds__(256) bwd_kernel(const float* __restrict__ y, const float* __restrict__ tgt, int n,
                                                  const Weights W, TrainBuf T, float scale) {
    __shared__ __bf16 tile[4][32 * TP];
    const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int s0 = (blockIdx.x * 4 + wv) * 32;   // no early exit: block barrier below
    __bf16* dz = tile[wv];
    const int cell = l & 31, half = l >> 5;
    const int s = s0 + cell;
    const bool live = s < n;
    float dy[NO];
    float se = 0.0f;
#pragma unroll
    for (int o = 0; o < NO; ++o) {
        float d = live ? y[(long)s * NO + o] - tgt[(long)s * NO + o] : 0.0f;
        se += d * d;
        dy[o] = scale * 2.0f * d;
    }
    if (live && half == 0) T.loss[s] = se;
    if (live)
        for (int o = 2 * half; o < 2 * half + 2; ++o) T.dy[(long)s * NO + o] = dy[o];
    for (int j = 0; j < 64; ++j) {
        const int u = half * 64 + j;
        float dh = 0.0f;
#pragma unroll
        for (int o = 0; o < NO; ++o) dh += dy[o] * bf2f(W.w3[u * NO + o]);
        const float z = live ? T.z2[(long)s * NH + u] : 0.0f;
        const __bf16 v = f2bf(dh * gelu_grad(z));
        dz[cell * TP + u] = v;
        if (live) T.dz2[(long)s * NH + u] = v;
    }
    __syncthreads();
    // dh1[s][k] = sum_u dz2[s][u] W2[k][u]: B column k is row k of W2, contiguous in u
    floatx16 acc[4];
    wave_gemm_32x128x128(dz, W.w2, acc);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int u = 32 * t + cell;
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const int sr = s0 + drow(l, i);
            if (sr < n) T.dz1[(long)sr * NH + u] = acc[t][i] * gelu_grad(T.z1[(long)sr * NH + u]);
        }
    }
}

// Weight gradients over a chunk of WG_S samples per block, as partial sums
// per block (summed in a fixed order by the Adam kernel):
//   dW2[k][u] = sum_s h1[s][k] dz2[s][u]    (matrix cores, K = samples)
//   dW1[k][u] = sum_s x[s][k] dz1[s][u], dW3[u][o] = sum_s h2[s][u] dy[s][o],
//   db1, db2, db3 column sums                (vector ALUs)
constexpr int WG_S = 64;
__global__ void __launch_bounds__(256) wgrad_kernel(const float* __restrict__ x, int n, TrainBuf T,
                                                    float* __restrict__ partial) {
    __shared__ __bf16 h1t[NH * (WG_S + 8)];   // [k][s]
    __shared__ __bf16 dzt[NH * (WG_S + 8)];   // [u][s]
    constexpr int SP = WG_S + 8;
    const int s0 = blockIdx.x * WG_S;
    const int tid = threadIdx.x;
    for (int e = tid; e < WG_S * NH; e += 256) {
        const int s = e / NH, k = e % NH;
        const bool live = s0 + s < n;
        h1t[k * SP + s] = live ? T.h1[(long)(s0 + s) * NH + k] : f2bf(0.0f);
        dzt[k * SP + s] = live ? T.dz2[(long)(s0 + s) * NH + k] : f2bf(0.0f);
    }
    __syncthreads();
    float* out = partial + (long)blockIdx.x * NPARAM;
    // 16 tiles of dW2 (4 x 4), four per wave
    const int wv = tid >> 6, l = tid & 63;
    const int r = l & 31, kh = (l >> 5) * 8;
    for (int tt = 0; tt < 4; ++tt) {
        const int tile_id = wv * 4 + tt;
        const int tm = tile_id >> 2, tn = tile_id & 3;
        floatx16 acc;
        for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
        for (int c = 0; c < WG_S / 16; ++c) {
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(h1t + (32 * tm + r) * SP + 16 * c + kh);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(dzt + (32 * tn + r) * SP + 16 * c + kh);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        for (int i = 0; i < 16; ++i) {
            const int k = 32 * tm + drow(l, i), u = 32 * tn + r;
            out[OFF_W2 + k * NH + u] = acc[i];
        }
    }
    // dW1, db1, db2: one thread per unit u
    if (tid < NH) {
        const int u = tid;
        float g[NI] = {0, 0, 0, 0, 0, 0, 0, 0};
        float gb = 0.0f, gb2 = 0.0f;
        for (int s = 0; s < WG_S && s0 + s < n; ++s) {
            const float d = T.dz1[(long)(s0 + s) * NH + u];
#pragma unroll
            for (int k = 0; k < NI; ++k) g[k] += bf2f(f2bf(x[(long)(s0 + s) * NI + k])) * d;
            gb += d;
            gb2 += bf2f(dzt[u * SP + s]);
        }
#pragma unroll
        for (int k = 0; k < NI; ++k) out[OFF_W1 + k * NH + u] = g[k];
        out[OFF_B1 + u] = gb;
        out[OFF_B2 + u] = gb2;
    } else if (tid < 2 * NH) {
        const int u = tid - NH;
        float g[NO] = {0, 0, 0, 0};
        for (int s = 0; s < WG_S && s0 + s < n; ++s) {
            const float hv = bf2f(T.h2[(long)(s0 + s) * NH + u]);
#pragma unroll
            for (int o = 0; o < NO; ++o) g[o] += hv * T.dy[(long)(s0 + s) * NO + o];
        }
#pragma unroll
        for (int o = 0; o < NO; ++o) out[OFF_W3 + u * NO + o] = g[o];
        if (u < NO) {
            float gb = 0.0f;
            for (int s = 0; s < WG_S && s0 + s < n; ++s) gb += T.dy[(long)(s0 + s) * NO + u];
            out[OFF_B3 + u] = gb;
        }
    }
}

// Sum of the block partials per parameter in block order (deterministic), then
// either the gradient is written out or Adam updates the FP32 master weights and
// refreshes the BF16 copies used by the kernels.
__global__ void __launch_bounds__(256) adam_kernel(const float* __restrict__ partial, int nblk, Params_ pm,
                                                   float* __restrict__ grad_out, bool update, float lr, float b1c,
                                                   float b2c) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= NPARAM) return;
    float g = 0.0f;
    for (int b = 0; b < nblk; ++b) g += partial[(long)b * NPARAM + p];
    if (grad_out) grad_out[p] = g;
    if (!update) return;
    const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
    float m = beta1 * pm.m[p] + (1.0f - beta1) * g;
    float v = beta2 * pm.v[p] + (1.0f - beta2) * g * g;
    pm.m[p] = m;
    pm.v[p] = v;
    const float mh = m / b1c, vh = v / b2c;
    const float w = pm.w[p] - lr * mh / (sqrtf(vh) + eps);
    pm.w[p] = w;
    write_bf16_copy(pm, p, w);
}

__global__ void refresh_kernel(Params_ pm) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < NPARAM) write_bf16_copy(pm, p, pm.w[p]);
}

__global__ void loss_kernel(const float* __restrict__ loss, int n, float* __restrict__ out) {
    // one block: fixed-order sum
    __shared__ float sh[256];
    float a = 0.0f;
    for (int s = threadIdx.x; s < n; s += 256) a += loss[s];
    sh[threadIdx.x] = a;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = sh[0];
}

}  // namespace mlp

// ---- host side ------------------------------------------------------------------
using namespace mlp;

// Forward launches: 4 waves per block (the weights staged once for 128
// samples) when that still gives every CU several blocks, else one wave per
// block so that a small batch spreads over the 256 CUs.
template <bool train>
static void launch_fwd(const Net& N, const float* x, float* y, int n, const TrainBuf& T, hipStream_t s) {
    if ((n + 127) / 128 >= 1024) {
        hipLaunchKernelGGL((fwd_kernel<train, 4>), dim3((n + 127) / 128), dim3(256), 0, s, x, n, N.W, y, T);
    } else {
        hipLaunchKernelGGL((fwd_kernel<train, 1>), dim3((n + 31) / 32), dim3(64), 0, s, x, n, N.W, y, T);
    }
}

void mlp_forward(const Net& N, const float* x, float* y, int n, hipStream_t s) {
    if (n <= 0) return;
    launch_fwd<false>(N, x, y, n, TrainBuf{}, s);
}

void mlp_refresh(const Net& N, hipStream_t s) {
    hipLaunchKernelGGL(refresh_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, N.P);
}

// forward + backward + gradient of the mean squared error over n samples
// (n <= N.cap); with `update`, one Adam step at learning rate lr (step count
// N.t incremented by the caller).  The loss sum lands in N.loss_dev[0].
void mlp_train_pass(Net& N, const float* x, const float* tgt, int n, bool update, float lr, float* grad_out,
                    hipStream_t s) {
    const int blocks = (n + 127) / 128;
    launch_fwd<true>(N, x, N.ybuf, n, N.T, s);
    const float scale = 1.0f / (float)(n * NO);
    hipLaunchKernelGGL(bwd_kernel, dim3(blocks), dim3(256), 0, s, (const float*)N.ybuf, tgt, n, N.W, N.T, scale);
    const int nblk = (n + WG_S - 1) / WG_S;
    hipLaunchKernelGGL(wgrad_kernel, dim3(nblk), dim3(256), 0, s, x, n, N.T, N.partial);
    float b1c = 1.0f, b2c = 1.0f;
    if (update) {
        b1c = 1.0f - powf(0.9f, (float)N.t);
        b2c = 1.0f - powf(0.999f, (float)N.t);
    }
    hipLaunchKernelGGL(adam_kernel, dim3((NPARAM + 255) / 256), dim3(256), 0, s, (const float*)N.partial, nblk, N.P,
                       grad_out, update, lr, b1c, b2c);
    hipLaunchKernelGGL(loss_kernel, dim3(1), dim3(256), 0, s, (const float*)N.T.loss, n, N.loss_dev);
}

}  // namespace erfx
