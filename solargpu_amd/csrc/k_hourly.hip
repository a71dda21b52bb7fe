/* Stage 2 on gfx950: the sunlit-hour loop.
 *
 * A pass covers K consecutive sunlit hours (K = 1 is the reference's one
 * hour per launch).  Per pass:
 *   k_solar_top  one lane per top cell: direct (SVF-gated) + Perez diffuse
 *                for the K hours -> hourly_top[cell][K] (cell-major, so a
 *                neighbour gather is one contiguous K-float load);
 *   k_reflect    one lane per voxel: its own direct + diffuse for the K hours,
 *                one walk of its neighbour list (CSR ids or in-kernel ba-RLE
 *                decode) with the hour-invariant Lambertian geometry computed
 *                once per neighbour and applied to the K gathered hours, then
 *                the FP32 Kahan-compensated annual accumulation, hour by hour
 *                in hour order (total += reflected + hourly).
 * Formulas: solar_kernel.cu:54-77 (+ Perez :8-50, hour-only terms hoisted
 * into HourConst), reflection_kernels.cu:18-65 (geometry), :70-89 (CSR walk),
 * :95-183 (ba-RLE decode), accumulate solar_kernel.cu:79-85. */
#include "sgpu_internal.h"

namespace sgpu {
namespace {

constexpr int kBlock = 256;

__device__ __forceinline__ float hour_irradiance(const HourConst& hc, float cs, float ss, float aspect,
                                                 bool shaded) {
    const float cos_aoi = hc.cos_z * cs + hc.sin_z * ss * cosf(hc.azk - aspect);
    float I = 0.0f;
    if (cos_aoi > 0.0f && hc.dense >= 0 && !shaded) I += fmaxf(0.0f, hc.dni_eff * cos_aoi);
    const float a = fmaxf(0.0f, cos_aoi);
    const float dif = hc.dhi * ((1 - hc.f1) * ((1 + cs) / 2) + hc.f1 * (a / hc.b) + hc.f2 * ss);
    I += fmaxf(0.0f, dif);
    return I;
}

__device__ __forceinline__ bool svf_blocked(const uint8_t* svf, int64_t v, int dense, int64_t n) {
    if (dense < 0) return false;
    return (svf[(int64_t)(dense >> 3) * n + v] >> (dense & 7)) & 1;   // voxel_types.h:36-44
}

template <int K>
__global__ __launch_bounds__(kBlock) void k_solar_top(HourArgs a) {
    const int64_t c = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (c >= a.n_top) return;
    const float slope = a.slopes[c], aspect = a.aspects[c];
    const float cs = cosf(slope), ss = sinf(slope);
    float out[K];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        out[h] = 0.0f;
        if (h < a.nh) {
            const HourConst hc = a.hours[h];
            out[h] = hour_irradiance(hc, cs, ss, aspect, svf_blocked(a.svf, c, hc.dense, a.n_total));
        }
    }
    float* dst = a.hourly_top + c * K;
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int h = 0; h < K; h += 4)
            *reinterpret_cast<float4*>(dst + h) = make_float4(out[h], out[h + 1], out[h + 2], out[h + 3]);
    } else {
#pragma unroll
        for (int h = 0; h < K; ++h) dst[h] = out[h];
    }
}

struct Receiver {
    int gx, gy;
    float z, nx, ny, nz;
};

/* One neighbour's contribution for the K hours (reflection_kernels.cu:18-38).
 * K == 1 keeps the reference's product order; K > 1 folds the hour-invariant
 * factor once and applies it to each gathered hour. */
template <int K>
__device__ __forceinline__ void add_neighbour(int ngid, const Receiver& r, const HourArgs& a, float* R) {
    if (ngid < 0 || ngid >= a.W * a.H) return;
    const int ngx = ngid % a.W, ngy = ngid / a.W;
    const float dx = (float)(ngx - r.gx), dy = (float)(ngy - r.gy), dz = a.grid[ngid] - r.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e-4f) return;
    const float d = sqrtf(d2);
    const float cr = (r.nx * dx + r.ny * dy + r.nz * dz) / d;
    if (cr <= 0.0f) return;
    const float cs = fmaxf(fabsf(-dz) / d, 0.1f);
    const float* hs = a.hourly_top + (int64_t)(ngy * a.W + ngx) * K;
    if constexpr (K == 1) {
        R[0] += 0.2f * hs[0] * cr * cs / (3.14159265f * d2);
    } else {
        const float wgt = 0.2f * cr * cs / (3.14159265f * d2);
        float hv[K];
        if constexpr (K % 4 == 0) {
#pragma unroll
            for (int h = 0; h < K; h += 4) {
                const float4 q = *reinterpret_cast<const float4*>(hs + h);
                hv[h] = q.x; hv[h + 1] = q.y; hv[h + 2] = q.z; hv[h + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int h = 0; h < K; ++h) hv[h] = hs[h];
        }
#pragma unroll
        for (int h = 0; h < K; ++h) R[h] += wgt * hv[h];
    }
}

template <int K, int FMT, bool DBG>
__global__ __launch_bounds__(kBlock) void k_reflect(HourArgs a) {
    const int64_t v = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (v >= a.n_total) return;
    Receiver r;
    const int top = a.pix_parent[v];
    r.gx = top % a.W;
    r.gy = top / a.W;
    r.z = (v < a.n_top) ? a.grid[r.gx + r.gy * a.W] : a.wall_z[v - a.n_top];
    const float slope = a.slopes[v], aspect = a.aspects[v];
    const float ss = sinf(slope), cs = cosf(slope);
    r.nx = ss * sinf(aspect);
    r.ny = -ss * cosf(aspect);
    r.nz = cs;

    float own[K], R[K];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        R[h] = 0.0f;
        own[h] = 0.0f;
        if (h < a.nh) {
            const HourConst hc = a.hours[h];
            own[h] = hour_irradiance(hc, cs, ss, aspect, svf_blocked(a.svf, v, hc.dense, a.n_total));
        }
    }

    const int64_t start = a.nb_off[v], end = a.nb_off[v + 1];
    if constexpr (FMT == SGPU_FMT_CSR) {
        for (int64_t k = start; k < end; ++k) add_neighbour<K>(a.nb_ids[k] - 1, r, a, R);
    } else {
        /* ba-RLE block stream (reflection_kernels.cu:132-180). */
        const uint8_t* p = a.nb_bytes + start;
        const int size = (int)(end - start);
        int i = 0;
        unsigned prev = 0;
        bool first = true;
        while (i < size) {
            const unsigned bits = ((unsigned)p[i] << 16) | ((unsigned)p[i + 1] << 8) | (unsigned)p[i + 2];
            i += 3;
            int nc = 0;
            unsigned widths = 0;   // 3 bits per run
            for (int j = 0; j < 8; ++j) {
                const unsigned c = (bits >> (21 - 3 * j)) & 7u;
                if (c == 7u) break;
                widths |= c << (3 * j);
                ++nc;
            }
            const int len_at = i;
            i += nc;
            for (int j = 0; j < nc; ++j) {
                const unsigned w = (widths >> (3 * j)) & 7u;
                const int run = p[len_at + j];
                if (w == 0) {
                    for (int k = 0; k < run;) {
                        const unsigned byte = p[i++];
                        unsigned val = first ? ((byte >> 4) & 15u) : prev - ((byte >> 4) & 15u);
                        first = false;
                        prev = val;
                        if (val > 0) add_neighbour<K>((int)val - 1, r, a, R);
                        ++k;
                        if (k < run) {
                            val = prev - (byte & 15u);
                            prev = val;
                            if (val > 0) add_neighbour<K>((int)val - 1, r, a, R);
                            ++k;
                        }
                    }
                } else {
                    for (int k = 0; k < run; ++k) {
                        unsigned dlt = 0;
                        for (unsigned q = 0; q < w; ++q) dlt = (dlt << 8) | p[i++];
                        const unsigned val = first ? dlt : prev - dlt;
                        first = false;
                        prev = val;
                        if (val > 0) add_neighbour<K>((int)val - 1, r, a, R);
                    }
                }
            }
        }
    }

    if constexpr (DBG) {
        a.dbg_hourly[v] = own[0];
        a.dbg_reflected[v] = R[0];
        return;
    }
    float s = a.sum[v], c = a.comp[v];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        if (h < a.nh) {
            const float x = R[h] + own[h];
            const float y = x - c;
            const float t = s + y;
            c = (t - s) - y;
            s = t;
        }
    }
    a.sum[v] = s;
    a.comp[v] = c;
}

template <int K>
void solar_top_k(const HourArgs& a, hipStream_t s) {
    const unsigned g = (unsigned)((a.n_top + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_solar_top<K>, dim3(g), dim3(kBlock), 0, s, a);
}

template <int K>
void reflect_k(const HourArgs& a, unsigned fmt, hipStream_t s) {
    const unsigned g = (unsigned)((a.n_total + kBlock - 1) / kBlock);
    const bool dbg = a.dbg_hourly != nullptr;
    if (fmt == SGPU_FMT_CSR) {
        if (dbg) hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_CSR, true>), dim3(g), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_CSR, false>), dim3(g), dim3(kBlock), 0, s, a);
    } else {
        if (dbg) hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_BARLE, true>), dim3(g), dim3(kBlock), 0, s, a);
        else hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_BARLE, false>), dim3(g), dim3(kBlock), 0, s, a);
    }
}

}  // namespace

bool supported_hours_per_pass(int K) { return K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32; }

void launch_solar_top(const HourArgs& a, int K, hipStream_t s) {
    if (a.n_top <= 0) return;
    switch (K) {
        case 1: solar_top_k<1>(a, s); break;
        case 2: solar_top_k<2>(a, s); break;
        case 4: solar_top_k<4>(a, s); break;
        case 8: solar_top_k<8>(a, s); break;
        case 16: solar_top_k<16>(a, s); break;
        case 32: solar_top_k<32>(a, s); break;
        default: fail(SGPU_ERR_ARG, "unsupported hours per pass %d", K);
    }
    SGPU_HIP(hipGetLastError());
}

void launch_reflect_accum(const HourArgs& a, int K, unsigned fmt, hipStream_t s) {
    if (a.n_total <= 0) return;
    if (fmt != SGPU_FMT_CSR && fmt != SGPU_FMT_BARLE) fail(SGPU_ERR_ARG, "format must be CSR or BARLE");
    switch (K) {
        case 1: reflect_k<1>(a, fmt, s); break;
        case 2: reflect_k<2>(a, fmt, s); break;
        case 4: reflect_k<4>(a, fmt, s); break;
        case 8: reflect_k<8>(a, fmt, s); break;
        case 16: reflect_k<16>(a, fmt, s); break;
        case 32: reflect_k<32>(a, fmt, s); break;
        default: fail(SGPU_ERR_ARG, "unsupported hours per pass %d", K);
    }
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
