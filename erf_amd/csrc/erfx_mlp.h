// erfx_mlp.h -- shapes and buffers of the learned Kessler stand-in
// (erfx_mlp.hip, erfx_api.hip).
#pragma once

#include "erfx_common.h"

namespace erfx {
namespace mlp {

constexpr int NI = 8, NH = 128, NO = 4;
// parameter vector (FP32 master copy): W1 [k][u], b1, W2 [k][u], b2, W3 [u][o], b3
constexpr int OFF_W1 = 0;
constexpr int OFF_B1 = OFF_W1 + NI * NH;
constexpr int OFF_W2 = OFF_B1 + NH;
constexpr int OFF_B2 = OFF_W2 + NH * NH;
constexpr int OFF_W3 = OFF_B2 + NH;
constexpr int OFF_B3 = OFF_W3 + NH * NO;
constexpr int NPARAM = OFF_B3 + NO;   // 18180
static_assert(NPARAM == 18180, "8-128-128-4 parameter count");

// what the kernels read: BF16 weights (W2 also transposed), FP32 biases
struct Weights {
    const __bf16* w1;    // [k][u]
    const float* b1;
    const __bf16* w2;    // [k][u]
    const __bf16* w2t;   // [u][k]
    const float* b2;
    const __bf16* w3;    // [u][o]
    const float* b3;
};

// FP32 master weights, Adam moments and the BF16 copies the update refreshes
struct Params_ {
    float* w;
    float* m;
    float* v;
    __bf16* w1b;
    __bf16* w2b;
    __bf16* w2tb;
    __bf16* w3b;
};

__device__ __forceinline__ void write_bf16_copy(const Params_& pm, int p, float w) {
    const __bf16 b = (__bf16)w;
    if (p < OFF_B1) {
        pm.w1b[p - OFF_W1] = b;
    } else if (p >= OFF_W2 && p < OFF_B2) {
        const int e = p - OFF_W2, k = e / NH, u = e % NH;
        pm.w2b[e] = b;
        pm.w2tb[u * NH + k] = b;
    } else if (p >= OFF_W3 && p < OFF_B3) {
        pm.w3b[p - OFF_W3] = b;
    }
}

// activations the backward pass reads, per sample
struct TrainBuf {
    float* z1;      // [s][NH] pre-activation of layer 1
    __bf16* h1;     // [s][NH]
    float* z2;      // [s][NH]
    __bf16* h2;     // [s][NH]
    __bf16* dz2;    // [s][NH]
    float* dz1;     // [s][NH]
    float* dy;      // [s][NO]
    float* loss;    // [s] squared error
};

}  // namespace mlp

// a network with its training workspace (device memory owned by the API object)
struct Net {
    mlp::Weights W;
    mlp::Params_ P;
    mlp::TrainBuf T;
    float* ybuf = nullptr;      // [cap][NO]
    float* partial = nullptr;   // [(cap + 63) / 64 + 1][NPARAM]
    float* loss_dev = nullptr;  // [1]
    int cap = 0;                // samples per training pass
    long t = 0;                 // Adam step count
};

// feature scaling of the network and the raw range it was trained on
struct MlpScaling {
    double x_mean[mlp::NI], x_scale[mlp::NI], x_lo[mlp::NI], x_hi[mlp::NI], y_scale[mlp::NO];
};

void mlp_forward(const Net& N, const float* x, float* y, int n, hipStream_t s);
void mlp_refresh(const Net& N, hipStream_t s);
void mlp_train_pass(Net& N, const float* x, const float* tgt, int n, bool update, float lr, float* grad_out,
                    hipStream_t s);

// erfx_kessler.hip: Kessler's local tendencies as training rows, and the
// microphysics step with the network (see there)
void kessler_samples(const Params& P, const double* cons, const double* w, double dt, float* out, long max_rows,
                     unsigned long long* counter, int* bad, hipStream_t s);
void kessler_mlp_advance(const Params& P, const Net& N, const MlpScaling& S, double* cons, const double* w,
                         double* fz, double* rain_accum, double dt, int* slot, float* x, float* y, long* rows_cell,
                         unsigned long long* counter, unsigned long long* host_counts, int* bad, hipStream_t s);

}  // namespace erfx
