/* Surrogate MLP: kernel argument blocks and launchers (k_mlp.hip). */
#pragma once

#include <cstdint>

#include "sgpu_internal.h"

namespace sgpu {

constexpr int kFeatures = 53;      // used inputs
constexpr int kInPad = 64;         // padded input width (multiple of the MFMA K = 32)
constexpr int kHidden = 256;
constexpr int kLayers = 5;         // 4 hidden + output
constexpr int kOutputs = 2;

enum { kEpiPartial = 0, kEpiForward = 1, kEpiBackward = 2 };

/* C[M x N] = A[M x K] . Bt[N x K]^T over k in [k_begin + z*k_len, +k_len) for grid.z = z. */
struct GemmArgs {
    const uint16_t* A;
    int lda;
    const uint16_t* Bt;
    int ldb;
    int M, N, K;
    int k_begin, k_len;
    const float* bias;      // forward: N
    float* Cf;              // partial: [splits][M][N] fp32
    uint16_t* Z;            // forward: pre-activation [M][N] (optional)
    uint16_t* H;            // forward: GELU output [M][N]
    uint16_t* Ht;           // forward: GELU output transposed [N][ldt] (optional)
    const uint16_t* Zin;    // backward: pre-activation [M][N]
    uint16_t* dZ;           // backward: [M][N]
    uint16_t* dZt;          // backward: [N][ldt]
    int ldt;
};
void launch_gemm(const GemmArgs& g, int epi, int splits, hipStream_t s);

struct OutArgs {
    const uint16_t* H4;     // [B][256]
    const uint16_t* Z4;     // [B][256] (training)
    const uint16_t* W5;     // [2][256] bf16
    const float* b5;        // [2]
    int64_t B;
    float* Y;               // inference output [B][2]
    const float* T;         // training targets [B][2]
    float wt[2];            // per-output loss weights
    float scale;            // 1 / B
    uint16_t* dZ4;          // [B][256]
    uint16_t* dZ4t;         // [256][ldt]
    int ldt;
    float* gW5;             // [2][256] (accumulated)
    float* gb5;             // [2]
    double* loss;           // summed weighted squared error
};
void launch_out(const OutArgs& o, bool train, hipStream_t s);
void launch_bias_grad(const uint16_t* dZt, int ldt, int B, int n, float* gb, hipStream_t s);

/* Fused inference of B rows: X [B][64] bf16 -> Y [B][2] fp32 (k_mlp_infer, `rows` voxels per
   workgroup: 32 or 64). */
struct InferArgs {
    const uint16_t* X;
    const uint16_t* W[4];   // hidden layer l: [256][in_l] bf16
    const float* b[4];      // [256]
    const uint16_t* W5;     // [2][256] bf16
    const float* b5;        // [2]
    int64_t B;
    float* Y;
};
void launch_infer(const InferArgs& f, int rows, hipStream_t s);

struct AdamArgs {
    float* w;               // fp32 master
    float* m;
    float* v;
    const float* grad;      // [splits][n]
    int splits;
    float grad_scale;
    int64_t n;
    int rows, cols;         // w viewed as [rows][cols] for the transposed copy
    uint16_t* wb;           // bf16 copy (optional)
    uint16_t* wt;           // bf16 transposed copy (optional)
    float lr, beta1, beta2, eps, bc1, bc2;
};
void launch_adam(const AdamArgs& a, hipStream_t s);

}  // namespace sgpu
