/* Per-voxel surrogate MLP on gfx950: BF16 weights and activations, FP32
 * accumulation, hand-written MFMA GEMMs for forward, backward and the weight
 * gradients, and an Adam kernel.
 *
 * Network: 64 (53 features + zero padding) -> 256 -> 256 -> 256 -> 256 -> 2,
 * exact-erf GELU on the hidden layers.  Layout: every matrix is row-major with
 * the reduction (K) dimension contiguous, so an MFMA operand fragment is one
 * 16-byte load.  Weights are kept as W[out][in] (forward) and Wt[in][out]
 * (backward-data); activations as H[batch][feat] (next layer) and
 * Ht[feat][batch] (weight gradient, K = batch).
 *
 * v_mfma_f32_16x16x32_bf16: lane l holds A[l%16][8(l/16)..+8],
 * B[8(l/16)..+8][l%16] and D[4(l/16)..+4][l%16]. */
#include <cmath>
#include <cstring>

#include "sgpu_internal.h"
#include "sgpu_mlp.h"

namespace sgpu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }
__device__ __forceinline__ uint16_t f2bf(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);   // round to nearest even
}

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float z) {
    return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ __forceinline__ bf16x8 load8(const uint16_t* p) {
    const uint4 q = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, q);
}

constexpr int kTile = 64;   // workgroup tile (M and N); 4 waves of 32 x 32

template <int EPI>
__global__ __launch_bounds__(256) void k_gemm(GemmArgs g) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int m0 = blockIdx.x * kTile + (wave >> 1) * 32;
    const int n0 = blockIdx.y * kTile + (wave & 1) * 32;
    const int r = lane & 15, q = lane >> 4;
    const int kb = g.k_begin + blockIdx.z * g.k_len;
    const int ke = min(kb + g.k_len, g.k_begin + g.K);
    f32x4 acc[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
    const bf16x8 zero = __builtin_bit_cast(bf16x8, make_uint4(0, 0, 0, 0));
    for (int k = kb; k < ke; k += 32) {
        bf16x8 a[2], b[2];
#pragma unroll
        for (int i = 0; i < 2; ++i) {
            const int m = m0 + 16 * i + r;
            a[i] = (m < g.M) ? load8(g.A + (int64_t)m * g.lda + k + 8 * q) : zero;
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + 16 * j + r;
            b[j] = (n < g.N) ? load8(g.Bt + (int64_t)n * g.ldb + k + 8 * q) : zero;
        }
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const int n = n0 + 16 * j + r;
            const int mb = m0 + 16 * i + 4 * q;
            if (n >= g.N) continue;
            if constexpr (EPI == kEpiPartial) {
                float* dst = g.Cf + (int64_t)blockIdx.z * g.M * g.N;
#pragma unroll
                for (int e = 0; e < 4; ++e)
                    if (mb + e < g.M) dst[(int64_t)(mb + e) * g.N + n] = acc[i][j][e];
            } else if constexpr (EPI == kEpiForward) {
                uint16_t hv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = mb + e;
                    const float z = acc[i][j][e] + g.bias[n];
                    const float h = gelu(z);
                    hv[e] = f2bf(h);
                    if (m < g.M) {
                        if (g.Z) g.Z[(int64_t)m * g.N + n] = f2bf(z);
                        g.H[(int64_t)m * g.N + n] = hv[e];
                    }
                }
                if (g.Ht && mb + 3 < g.M)
                    *reinterpret_cast<uint2*>(g.Ht + (int64_t)n * g.ldt + mb) =
                        make_uint2(hv[0] | ((uint32_t)hv[1] << 16), hv[2] | ((uint32_t)hv[3] << 16));
            } else {   // kEpiBackward: dZ = dH * gelu'(Z)
                uint16_t dv[4];
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int m = mb + e;
                    float dz = 0.0f;
                    if (m < g.M) dz = acc[i][j][e] * gelu_grad(bf2f(g.Zin[(int64_t)m * g.N + n]));
                    dv[e] = f2bf(dz);
                    if (m < g.M) g.dZ[(int64_t)m * g.N + n] = dv[e];
                }
                if (mb + 3 < g.M)
                    *reinterpret_cast<uint2*>(g.dZt + (int64_t)n * g.ldt + mb) =
                        make_uint2(dv[0] | ((uint32_t)dv[1] << 16), dv[2] | ((uint32_t)dv[3] << 16));
            }
        }
    }
}

/* Output layer 256 -> 2 (lane = sample).  Training: squared-error loss gradient,
 * dZ4 = (dY W5) * gelu'(Z4) in both layouts, dW5/db5 and the loss via LDS + atomics. */
template <bool TRAIN>
__global__ __launch_bounds__(256) void k_out(OutArgs o) {
    __shared__ float w5[2][kHidden];
    __shared__ float red[3][kHidden];
    const int tid = threadIdx.x;
    for (int i = tid; i < 2 * kHidden; i += 256) w5[i / kHidden][i % kHidden] = bf2f(o.W5[i]);
    if (TRAIN)
        for (int i = tid; i < 3 * kHidden; i += 256) red[i / kHidden][i % kHidden] = 0.0f;
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.x * 256 + tid;
    const bool live = b < o.B;
    const uint16_t* h = o.H4 + b * kHidden;
    float y0 = o.b5[0], y1 = o.b5[1];
    if (live) {
        for (int i = 0; i < kHidden; i += 8) {
            const uint4 q = *reinterpret_cast<const uint4*>(h + i);
            const uint32_t wv[4] = {q.x, q.y, q.z, q.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float lo = __uint_as_float(wv[e] << 16), hi = __uint_as_float(wv[e] & 0xFFFF0000u);
                y0 += lo * w5[0][i + 2 * e] + hi * w5[0][i + 2 * e + 1];
                y1 += lo * w5[1][i + 2 * e] + hi * w5[1][i + 2 * e + 1];
            }
        }
    }
    if (!TRAIN) {
        if (live) {
            o.Y[2 * b] = y0;
            o.Y[2 * b + 1] = y1;
        }
        return;
    }
    float g0 = 0.0f, g1 = 0.0f;
    if (live) {
        const float e0 = y0 - o.T[2 * b], e1 = y1 - o.T[2 * b + 1];
        g0 = 2.0f * e0 * o.scale * o.wt[0];
        g1 = 2.0f * e1 * o.scale * o.wt[1];
        atomicAdd(o.loss, (double)(o.wt[0] * e0 * e0 + o.wt[1] * e1 * e1));
        atomicAdd(&red[2][0], g0);
        atomicAdd(&red[2][1], g1);
    }
    for (int i = 0; i < kHidden; ++i) {
        float hv = 0.0f, dz = 0.0f;
        if (live) {
            hv = bf2f(h[i]);
            const float dh = g0 * w5[0][i] + g1 * w5[1][i];
            dz = dh * gelu_grad(bf2f(o.Z4[b * kHidden + i]));
            o.dZ4[b * kHidden + i] = f2bf(dz);
            o.dZ4t[(int64_t)i * o.ldt + b] = f2bf(dz);
        }
        /* wave-reduce g*h for dW5 */
        float p0 = g0 * hv, p1 = g1 * hv;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            p0 += __shfl_xor(p0, m, 64);
            p1 += __shfl_xor(p1, m, 64);
        }
        if ((tid & 63) == 0) {
            atomicAdd(&red[0][i], p0);
            atomicAdd(&red[1][i], p1);
        }
    }
    __syncthreads();
    for (int i = tid; i < kHidden; i += 256) {
        atomicAdd(&o.gW5[i], red[0][i]);
        atomicAdd(&o.gW5[kHidden + i], red[1][i]);
    }
    if (tid == 0) {
        atomicAdd(&o.gb5[0], red[2][0]);
        atomicAdd(&o.gb5[1], red[2][1]);
    }
}

/* Fused inference: one workgroup takes ROWS voxels through all five layers.  The activations
 * stay in LDS (two ping-pong tiles of ROWS x 256 bf16, rows padded by 16 bytes so the 16 rows of
 * an MFMA A fragment fall in different banks); the weights stream from L2.  Wave w computes
 * output columns 64w..64w+63 of every layer with the MFMA steps, bias, GELU and bf16 rounding of
 * k_gemm<kEpiForward>, and the 256 -> 2 output layer is k_out's per-row loop, so the outputs equal
 * the layer-by-layer path bit for bit. */
constexpr int kActPitch = kHidden + 8;

__device__ __forceinline__ bf16x8 lds8(const uint16_t* p) {
    const uint4 q = *reinterpret_cast<const uint4*>(p);
    return __builtin_bit_cast(bf16x8, q);
}

template <int ROWS>
__global__ __launch_bounds__(256) void k_mlp_infer(InferArgs f) {
    static_assert(ROWS == 32 || ROWS == 64, "tile rows");
    constexpr int MI = ROWS / 16;   // 16-row MFMA blocks per wave
    __shared__ __align__(16) uint16_t act[2][ROWS * kActPitch];
    __shared__ float w5[2][kHidden];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int r = lane & 15, q = lane >> 4;
    const int64_t row0 = (int64_t)blockIdx.x * ROWS;
    for (int i = threadIdx.x; i < 2 * kHidden; i += 256) w5[i / kHidden][i % kHidden] = bf2f(f.W5[i]);
    for (int i = threadIdx.x; i < ROWS * (kInPad / 8); i += 256) {
        const int rr = i / (kInPad / 8), cc = (i % (kInPad / 8)) * 8;
        uint4 v = make_uint4(0u, 0u, 0u, 0u);
        if (row0 + rr < f.B) v = *reinterpret_cast<const uint4*>(f.X + (row0 + rr) * kInPad + cc);
        *reinterpret_cast<uint4*>(act[0] + rr * kActPitch + cc) = v;
    }
    __syncthreads();
    const int n0 = wave * 64;
    int cur = 0;
    for (int l = 0; l < 4; ++l) {
        const int K = (l == 0) ? kInPad : kHidden;
        const uint16_t* W = f.W[l];
        f32x4 acc[MI][4];
#pragma unroll
        for (int i = 0; i < MI; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
        for (int k = 0; k < K; k += 32) {
            bf16x8 a[MI], b[4];
#pragma unroll
            for (int j = 0; j < 4; ++j) b[j] = load8(W + (int64_t)(n0 + 16 * j + r) * K + k + 8 * q);
#pragma unroll
            for (int i = 0; i < MI; ++i) a[i] = lds8(act[cur] + (16 * i + r) * kActPitch + k + 8 * q);
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
        }
        uint16_t* out = act[cur ^ 1];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int n = n0 + 16 * j + r;
            const float bias = f.b[l][n];
#pragma unroll
            for (int i = 0; i < MI; ++i)
#pragma unroll
                for (int e = 0; e < 4; ++e) out[(16 * i + 4 * q + e) * kActPitch + n] = f2bf(gelu(acc[i][j][e] + bias));
        }
        __syncthreads();
        cur ^= 1;
    }
    /* output layer: one lane per row, k_out's arithmetic */
    for (int rr = threadIdx.x; rr < ROWS; rr += 256) {
        const int64_t b = row0 + rr;
        if (b >= f.B) continue;
        const uint16_t* h = act[cur] + rr * kActPitch;
        float y0 = f.b5[0], y1 = f.b5[1];
        for (int i = 0; i < kHidden; i += 8) {
            const uint4 qv = *reinterpret_cast<const uint4*>(h + i);
            const uint32_t wv[4] = {qv.x, qv.y, qv.z, qv.w};
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const float lo = __uint_as_float(wv[e] << 16), hi = __uint_as_float(wv[e] & 0xFFFF0000u);
                y0 += lo * w5[0][i + 2 * e] + hi * w5[0][i + 2 * e + 1];
                y1 += lo * w5[1][i + 2 * e] + hi * w5[1][i + 2 * e + 1];
            }
        }
        f.Y[2 * b] = y0;
        f.Y[2 * b + 1] = y1;
    }
}

/* db[n] = sum over the batch of dZt[n][.] (one wave per n). */
__global__ __launch_bounds__(64) void k_bias_grad(const uint16_t* dZt, int ldt, int B, float* gb) {
    const int n = blockIdx.x, lane = threadIdx.x;
    float s = 0.0f;
    for (int b = lane; b < B; b += 64) s += bf2f(dZt[(int64_t)n * ldt + b]);
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) s += __shfl_xor(s, m, 64);
    if (lane == 0) gb[n] = s;
}

/* Adam over one tensor; the gradient is the sum of `splits` partial buffers. */
__global__ __launch_bounds__(256) void k_adam(AdamArgs a) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= a.n) return;
    float g = 0.0f;
    for (int s = 0; s < a.splits; ++s) g += a.grad[(int64_t)s * a.n + p];
    g *= a.grad_scale;
    const float m = a.beta1 * a.m[p] + (1.0f - a.beta1) * g;
    const float v = a.beta2 * a.v[p] + (1.0f - a.beta2) * g * g;
    a.m[p] = m;
    a.v[p] = v;
    const float mh = m * a.bc1, vh = v * a.bc2;
    const float w = a.w[p] - a.lr * mh / (sqrtf(vh) + a.eps);
    a.w[p] = w;
    if (a.wb) a.wb[p] = f2bf(w);
    if (a.wt) {   // transposed bf16 copy: p = row * cols + col -> col * rows + row
        const int64_t row = p / a.cols, col = p % a.cols;
        a.wt[col * a.rows + row] = f2bf(w);
    }
}

}  // namespace

void launch_gemm(const GemmArgs& g, int epi, int splits, hipStream_t s) {
    if (g.K % 32 || g.k_len % 32 || g.lda % 8 || g.ldb % 8) fail(SGPU_ERR_ARG, "gemm: K and strides must be multiples of 32/8");
    dim3 grid((unsigned)((g.M + kTile - 1) / kTile), (unsigned)((g.N + kTile - 1) / kTile), (unsigned)splits);
    switch (epi) {
        case kEpiPartial: hipLaunchKernelGGL(k_gemm<kEpiPartial>, grid, dim3(256), 0, s, g); break;
        case kEpiForward: hipLaunchKernelGGL(k_gemm<kEpiForward>, grid, dim3(256), 0, s, g); break;
        case kEpiBackward: hipLaunchKernelGGL(k_gemm<kEpiBackward>, grid, dim3(256), 0, s, g); break;
        default: fail(SGPU_ERR_ARG, "gemm: bad epilogue");
    }
    SGPU_HIP(hipGetLastError());
}

void launch_out(const OutArgs& o, bool train, hipStream_t s) {
    const unsigned g = (unsigned)((o.B + 255) / 256);
    if (train) hipLaunchKernelGGL(k_out<true>, dim3(g), dim3(256), 0, s, o);
    else hipLaunchKernelGGL(k_out<false>, dim3(g), dim3(256), 0, s, o);
    SGPU_HIP(hipGetLastError());
}

void launch_infer(const InferArgs& f, int rows, hipStream_t s) {
    if (f.B <= 0) return;
    if (rows == 32)
        hipLaunchKernelGGL(k_mlp_infer<32>, dim3((unsigned)((f.B + 31) / 32)), dim3(256), 0, s, f);
    else if (rows == 64)
        hipLaunchKernelGGL(k_mlp_infer<64>, dim3((unsigned)((f.B + 63) / 64)), dim3(256), 0, s, f);
    else
        fail(SGPU_ERR_ARG, "inference tiles of 32 or 64 rows");
    SGPU_HIP(hipGetLastError());
}

void launch_bias_grad(const uint16_t* dZt, int ldt, int B, int n, float* gb, hipStream_t s) {
    hipLaunchKernelGGL(k_bias_grad, dim3(n), dim3(64), 0, s, dZt, ldt, B, gb);
    SGPU_HIP(hipGetLastError());
}

void launch_adam(const AdamArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_adam, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
