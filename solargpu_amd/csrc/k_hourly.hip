/* Stage 2 on gfx950: the sunlit-hour loop.
 *
 * A pass covers K consecutive sunlit hours (K = 1 is the reference's one hour
 * per launch).  Per pass:
 *   k_solar_top  one lane per top cell: direct (SVF-gated) + Perez diffuse for
 *                the K hours -> hourly_top[cell][K] (cell-major: a neighbour
 *                gather is one contiguous K-float load).
 *   k_reflect    one wave per 64 consecutive voxels.
 *                  phase 1, lane = voxel: receiver geometry and the voxel's own
 *                           direct + diffuse for the K hours (registers);
 *                  phase 2, lane = neighbour: for each of the 64 voxels in turn
 *                           the whole wave walks its list -- CSR ids read
 *                           coalesced, or the ba-RLE stream decoded wave-parallel
 *                           (run lookup per element, delta prefix sum across
 *                           lanes) -- computes the hour-invariant Lambertian
 *                           weight once per neighbour, applies it to the K
 *                           gathered hours, and reduces across lanes with a
 *                           butterfly transpose (K + log2(64/K) shuffles);
 *                  phase 3, lane = voxel: FP32 Kahan-compensated annual
 *                           accumulation, hour by hour in hour order
 *                           (total += reflected + hourly).
 * Workgroups are remapped so each XCD walks one contiguous eighth of the
 * voxels, keeping its neighbour gathers in its own L2.
 *
 * Formulas: solar_kernel.cu:54-77 (+ Perez :8-50; hour-only terms hoisted into
 * HourConst), reflection_kernels.cu:18-65 (geometry), :70-89 (CSR walk),
 * :95-183 (ba-RLE decode), accumulate solar_kernel.cu:79-85.  Transcendentals
 * and reciprocals use the device's fast forms; the parity bar is 1e-4 on the
 * annual sums against the FP64 oracle. */
#include "sgpu_internal.h"

namespace sgpu {
namespace {

constexpr int kSolarBlock = 256;
constexpr int kXcds = 8;

__device__ __forceinline__ float hour_irradiance(const HourConst& hc, float cs, float ss, float ca, float sa,
                                                 bool shaded) {
    /* cos(az - aspect) by angle addition: cos az cos a + sin az sin a */
    const float cdiff = hc.cos_az * ca + hc.sin_az * sa;
    const float cos_aoi = hc.cos_z * cs + hc.sin_z * ss * cdiff;
    float I = 0.0f;
    if (cos_aoi > 0.0f && hc.dense >= 0 && !shaded) I += fmaxf(0.0f, hc.dni_eff * cos_aoi);
    const float a = fmaxf(0.0f, cos_aoi);
    const float dif = hc.dhi * ((1 - hc.f1) * ((1 + cs) / 2) + hc.f1 * (a * hc.inv_b) + hc.f2 * ss);
    I += fmaxf(0.0f, dif);
    return I;
}

__device__ __forceinline__ bool svf_blocked(const uint8_t* svf, int64_t v, int dense, int64_t n) {
    if (dense < 0) return false;
    return (svf[(int64_t)(dense >> 3) * n + v] >> (dense & 7)) & 1;   // voxel_types.h:36-44
}

template <int K>
__global__ __launch_bounds__(kSolarBlock) void k_solar_top(HourArgs a) {
    const int64_t c = (int64_t)blockIdx.x * kSolarBlock + threadIdx.x;
    if (c >= a.n_top) return;
    const float slope = a.slopes[c], aspect = a.aspects[c];
    float ss, cs, sa, ca;
    sincosf(slope, &ss, &cs);
    sincosf(aspect, &sa, &ca);
    float out[K];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        out[h] = 0.0f;
        if (h < a.nh) {
            const HourConst& hc = a.hours[h];
            out[h] = hour_irradiance(hc, cs, ss, ca, sa, svf_blocked(a.svf, c, hc.dense, a.n_total));
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

struct Recv {
    float gx, gy, z, nx, ny, nz;
};

/* Division of a cell id by the grid width via a float reciprocal and one fix-up
 * (exact for ids < 2^26 and widths <= 8192). */
__device__ __forceinline__ void cell_xy(int id, int W, float inv_w, int& x, int& y) {
    int q = (int)((float)id * inv_w);
    int r = id - q * W;
    if (r < 0) { --q; r += W; }
    else if (r >= W) { ++q; r -= W; }
    x = r;
    y = q;
}

/* Contribution of neighbour cell `id` (0-based) to the K hourly sums of receiver r
 * (reflection_kernels.cu:18-38). */
template <int K>
__device__ __forceinline__ void add_neighbour(int id, const Recv& r, const HourArgs& a, float inv_w, float* R) {
    if (id < 0 || id >= a.W * a.H) return;
    int nx, ny;
    cell_xy(id, a.W, inv_w, nx, ny);
    const float dx = (float)nx - r.gx, dy = (float)ny - r.gy, dz = a.grid[id] - r.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e-4f) return;
    const float inv_d = rsqrtf(d2);
    const float cr = (r.nx * dx + r.ny * dy + r.nz * dz) * inv_d;
    if (cr <= 0.0f) return;
    const float cs = fmaxf(fabsf(dz) * inv_d, 0.1f);
    const float wgt = (0.2f / 3.14159265f) * cr * cs * (inv_d * inv_d);
    const float* hs = a.hourly_top + (int64_t)id * K;
    if constexpr (K % 4 == 0) {
#pragma unroll
        for (int h = 0; h < K; h += 4) {
            const float4 q = *reinterpret_cast<const float4*>(hs + h);
            R[h] += wgt * q.x;
            R[h + 1] += wgt * q.y;
            R[h + 2] += wgt * q.z;
            R[h + 3] += wgt * q.w;
        }
    } else {
#pragma unroll
        for (int h = 0; h < K; ++h) R[h] += wgt * hs[h];
    }
}

/* Butterfly transpose-reduce of R[K] over the 64 lanes.  Afterwards lane L holds the
 * full sum of hour hour_of_lane(L) in R[0]. */
template <int K>
__device__ __forceinline__ void wave_transpose_reduce(float* R, int lane) {
    int width = K;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        if (width > 1) {
            const int half = width >> 1;
            const bool upper = lane & m;
#pragma unroll
            for (int h = 0; h < K / 2; ++h) {
                if (h < half) {
                    const float keep = upper ? R[h + half] : R[h];
                    const float give = upper ? R[h] : R[h + half];
                    R[h] = keep + __shfl_xor(give, m, 64);
                }
            }
            width = half;
        } else {
            R[0] += __shfl_xor(R[0], m, 64);
        }
    }
}

/* Which hour lane L holds after wave_transpose_reduce<K>. */
template <int K>
__device__ __forceinline__ int hour_of_lane(int lane) {
    int h = 0, width = K;
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        if (width > 1) {
            width >>= 1;
            if (lane & m) h += width;
        }
    }
    return h;
}

/* ba-RLE stream of one voxel, decoded by the whole wave; calls f(id) per non-zero value.
 * Block layout: 3 bytes of 3-bit width codes (7 = unused), one length byte per run,
 * payload (4-bit deltas two per byte, high nibble first; wider ones big-endian). */
template <typename F>
__device__ __forceinline__ void wave_decode(const uint8_t* __restrict__ p, int size, int lane, F&& f) {
    int i = 0;
    unsigned prev = 0;
    bool first = true;
    while (i < size) {
        const unsigned bits = ((unsigned)p[i] << 16) | ((unsigned)p[i + 1] << 8) | (unsigned)p[i + 2];
        i += 3;
        int nc = 0;
        unsigned wid[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const unsigned cd = (bits >> (21 - 3 * j)) & 7u;
            wid[j] = cd;
            if (cd != 7u && nc == j) nc = j + 1;
        }
        /* run lengths, element prefix and payload offsets (wave-uniform) */
        int len[8], estart[9], pstart[8];
        estart[0] = 0;
        int pay = i + nc;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            len[j] = (j < nc) ? (int)p[i + j] : 0;
            estart[j + 1] = estart[j] + len[j];
            pstart[j] = pay;
            pay += (wid[j] == 0u) ? (len[j] + 1) >> 1 : len[j] * (int)wid[j];
        }
        const int total = estart[8];
        for (int e0 = 0; e0 < total; e0 += 64) {
            const int e = e0 + lane;
            unsigned delta = 0;
            if (e < total) {
                int r = 0;
#pragma unroll
                for (int j = 1; j < 8; ++j) r += (e >= estart[j]) ? 1 : 0;
                const int k = e - estart[r];
                const unsigned w = wid[r];
                if (w == 0u) {
                    const unsigned byte = p[pstart[r] + (k >> 1)];
                    delta = (k & 1) ? (byte & 15u) : (byte >> 4);
                } else {
                    const uint8_t* q = p + pstart[r] + k * (int)w;
                    for (unsigned b = 0; b < w; ++b) delta = (delta << 8) | q[b];
                }
                if (first && e == 0) delta = 0u - delta;   // first value is absolute
            }
            /* inclusive prefix sum of deltas (mod 2^32) across the wave */
            unsigned s = delta;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const unsigned y = __shfl_up(s, o, 64);
                if (lane >= o) s += y;
            }
            const unsigned val = prev - s;
            if (e < total && val > 0u) f((int)val - 1);
            prev = __shfl(val, 63, 64);   // lanes past the end add zero deltas
            first = false;
        }
        i = pay;
    }
}

/* WAVES waves per workgroup (2 or 4): the LDS footprint / occupancy knob the autotuner picks. */
template <int K, int FMT, bool DBG, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_reflect(HourArgs a) {
    __shared__ float rbuf[WAVES][64][K + 1];
    __shared__ Recv rec[WAVES][64];
    __shared__ int64_t beg[WAVES][64 + 1];

    /* XCD-aware block order: block b runs on XCD b % 8; give each XCD a contiguous range. */
    const unsigned nb = gridDim.x;
    const unsigned per = (nb + kXcds - 1) / kXcds;
    const unsigned xcd = blockIdx.x % kXcds, slot = blockIdx.x / kXcds;
    const unsigned lb = xcd * per + slot;   // the grid is a multiple of 8, so lb < gridDim.x
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    /* positions p index the neighbour offsets and the accumulators; p is the voxel id
       itself unless a subset of voxel ids is given */
    const int64_t n = a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total);
    const int64_t p0 = ((int64_t)lb * WAVES + wave) * 64;
    const int nv = (p0 < n) ? (int)min<int64_t>(64, n - p0) : 0;
    const int64_t p = p0 + lane;
    const bool live = lane < nv;
    const int64_t v = live ? (a.subset ? (int64_t)a.subset[p] : a.v_base + p) : 0;
    const float inv_w = 1.0f / (float)a.W;

    /* phase 1: lane = voxel */
    float own[K];
    {
        Recv r{};
        float ss = 0, cs = 1, sa = 0, ca = 1;
        if (live) {
            const int top = a.pix_parent[v];
            int gx, gy;
            cell_xy(top, a.W, inv_w, gx, gy);
            r.gx = (float)gx;
            r.gy = (float)gy;
            r.z = (v < a.n_top) ? a.grid[top] : a.wall_z[v - a.n_top];
            sincosf(a.slopes[v], &ss, &cs);
            sincosf(a.aspects[v], &sa, &ca);
            r.nx = ss * sa;
            r.ny = -ss * ca;
            r.nz = cs;
            beg[wave][lane] = a.nb_off[p];
            if (lane == nv - 1) beg[wave][nv] = a.nb_off[p + 1];
        }
        rec[wave][lane] = r;
#pragma unroll
        for (int h = 0; h < K; ++h) {
            own[h] = 0.0f;
            if (live && h < a.nh) {
                const HourConst& hc = a.hours[h];
                own[h] = hour_irradiance(hc, cs, ss, ca, sa, svf_blocked(a.svf, v, hc.dense, a.n_total));
            }
        }
    }
    __syncthreads();

    /* phase 2: lane = neighbour, one voxel of the wave at a time */
    const int my_hour = hour_of_lane<K>(lane);
    const bool writer = (lane & ((64 / K) - 1)) == 0;   // one of the 64/K replicas of each hour
    for (int j = 0; j < nv; ++j) {
        const Recv r = rec[wave][j];
        const int64_t s0 = beg[wave][j], s1 = beg[wave][j + 1];
        float R[K];
#pragma unroll
        for (int h = 0; h < K; ++h) R[h] = 0.0f;
        if constexpr (FMT == SGPU_FMT_CSR) {
            for (int64_t k = s0 + lane; k < s1; k += 64) add_neighbour<K>(a.nb_ids[k] - 1, r, a, inv_w, R);
        } else {
            wave_decode(a.nb_bytes + s0, (int)(s1 - s0), lane,
                        [&](int id) { add_neighbour<K>(id, r, a, inv_w, R); });
        }
        wave_transpose_reduce<K>(R, lane);
        if (writer) rbuf[wave][j][my_hour] = R[0];
    }
    __syncthreads();

    /* phase 3: lane = voxel */
    if (!live) return;
    if constexpr (DBG) {
        a.dbg_hourly[p] = own[0];
        a.dbg_reflected[p] = rbuf[wave][lane][0];
        return;
    }
    float s = a.sum[p], c = a.comp[p];
#pragma unroll
    for (int h = 0; h < K; ++h) {
        if (h < a.nh) {
            const float x = rbuf[wave][lane][h] + own[h];
            const float y = x - c;
            const float t = s + y;
            c = (t - s) - y;
            s = t;
        }
    }
    a.sum[p] = s;
    a.comp[p] = c;
}

template <int K>
void solar_top_k(const HourArgs& a, hipStream_t s) {
    const unsigned g = (unsigned)((a.n_top + kSolarBlock - 1) / kSolarBlock);
    hipLaunchKernelGGL(k_solar_top<K>, dim3(g), dim3(kSolarBlock), 0, s, a);
}

template <int K, int WAVES>
void reflect_kw(const HourArgs& a, unsigned fmt, hipStream_t s) {
    const int64_t n = a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total);
    const int64_t blocks = (n + 64 * WAVES - 1) / (64 * WAVES);
    /* round up to a multiple of the XCD count so the remap covers every block */
    const unsigned g = (unsigned)(((blocks + kXcds - 1) / kXcds) * kXcds);
    const bool dbg = a.dbg_hourly != nullptr;
    const dim3 blk(64 * WAVES);
    if (fmt == SGPU_FMT_CSR) {
        if (dbg) hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_CSR, true, WAVES>), dim3(g), blk, 0, s, a);
        else hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_CSR, false, WAVES>), dim3(g), blk, 0, s, a);
    } else {
        if (dbg) hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_BARLE, true, WAVES>), dim3(g), blk, 0, s, a);
        else hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_BARLE, false, WAVES>), dim3(g), blk, 0, s, a);
    }
}

template <int K>
void reflect_k(const HourArgs& a, unsigned fmt, hipStream_t s) {
    if (a.waves == 2) reflect_kw<K, 2>(a, fmt, s);
    else reflect_kw<K, 4>(a, fmt, s);
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
    if ((a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total)) <= 0) return;
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
