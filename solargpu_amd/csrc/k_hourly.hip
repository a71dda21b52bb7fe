/* Stage 2 on gfx950: the sunlit-hour loop.
 *
 * A pass covers K consecutive sunlit hours (K = 1 is the reference's one hour
 * per launch; K <= 256).  Per pass:
 *   k_solar_top  one lane per top cell: direct (SVF-gated) + Perez diffuse for
 *                the K hours -> hourly_top[cell][K] (cell-major: one cell's K
 *                hours are one contiguous row; written through LDS as whole
 *                128-byte pieces of rows when K is a multiple of 32).
 *   k_reflect    one wave per 64 consecutive voxels (receivers).
 *                  phase 1, lane = voxel: receiver geometry and the voxel's own
 *                           direct + diffuse summed over the K hours;
 *                  phase 2, one receiver of the wave at a time:
 *                           lane = neighbour: CSR ids read coalesced, or the
 *                           ba-RLE stream decoded wave-parallel out of an
 *                           LDS-staged window (run lookup per element, delta
 *                           prefix sum across lanes); the hour-invariant
 *                           Lambertian weight; neighbours with a non-zero weight
 *                           are compacted into a per-wave LDS list (ballot +
 *                           prefix count);
 *                           lane = (neighbour slot, hour): the list is applied to
 *                           the gathered hourly_top rows, so one load instruction
 *                           reads 64/K whole rows, into one FP32 accumulator per
 *                           lane; a butterfly over the wave gives the receiver's
 *                           reflected irradiance summed over the pass;
 *                  phase 3, lane = voxel: the pass total (reflected + own)
 *                           enters the FP32 Kahan-compensated annual sum.
 * Workgroups are remapped XCD-aware: each XCD walks one contiguous eighth of
 * the voxels, or the eight XCDs walk adjacent bands side by side (xcd_block).
 *
 * Formulas: solar_kernel.cu:54-77 (+ Perez :8-50; hour-only terms hoisted into
 * HourConst), reflection_kernels.cu:18-65 (geometry), :70-89 (CSR walk),
 * :95-183 (ba-RLE decode), accumulate solar_kernel.cu:79-85 (total +=
 * reflected + hourly, here once per pass of K hours with Kahan compensation).
 * Transcendentals and reciprocals use the device's fast forms; the parity bar
 * is 1e-4 on the annual sums against the FP64 oracle. */
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

/* The row of cell c in the hourly buffer: c itself, or (swz = S > 0) c with its low 10 bits
 * XOR-ed with bits S.. (the cell's grid row when S = log2 of the width), a bijection within each
 * aligned block of 1024 rows.  Cells of one grid column sit W rows apart in cell order; the swizzle
 * gives them different low address bits, so a patch of neighbour rows spreads over the L2's sets. */
__device__ __forceinline__ unsigned row_slot(unsigned c, int swz) {
    return swz > 0 ? c ^ ((c >> swz) & 1023u) : c;
}

template <int K>
__device__ __forceinline__ void solar_top_cell(const HourArgs& a, int64_t c) {
    if (c >= a.n_top) return;
    const float slope = a.slopes[c], aspect = a.aspects[c];
    float ss, cs, sa, ca;
    sincosf(slope, &ss, &cs);
    sincosf(aspect, &sa, &ca);
    float* dst = a.hourly_top + (int64_t)row_slot((unsigned)c, a.row_swz) * K;
    constexpr int V = (K % 4 == 0) ? 4 : 1;   // hours per store
    for (int h0 = 0; h0 < K; h0 += V) {
        float out[V];
#pragma unroll
        for (int u = 0; u < V; ++u) {
            const int h = h0 + u;
            out[u] = 0.0f;
            if (h < a.nh) {
                const HourConst& hc = a.hours[h];
                out[u] = hour_irradiance(hc, cs, ss, ca, sa, svf_blocked(a.svf, c, hc.dense, a.n_total));
            }
        }
        if constexpr (V == 4) *reinterpret_cast<float4*>(dst + h0) = make_float4(out[0], out[1], out[2], out[3]);
        else dst[h0] = out[0];
    }
}

/* Hours per LDS round of the wave form below: 32 hours of a cell are one 128-byte piece of its row. */
constexpr int kSolarHours = 32;
struct SolarLds {
    float t[kSolarBlock / 64][64][kSolarHours + 4];   // per wave: 64 cells x 32 hours (rows padded by 16 B)
};

/* solar_top_cell for the 64 cells c0 .. c0 + 63 of a wave (K a multiple of kSolarHours): each lane
   computes its cell's hours 32 at a time into LDS, then the wave stores them as whole 128-byte
   pieces of the rows, 8 cells per store instruction (a lane's own 16-byte pieces of its row land in
   64 rows 1 KB apart, which the memory system writes back as twice the bytes).  Every lane of the
   wave takes part. */
template <int K>
__device__ __forceinline__ void solar_top_wave(const HourArgs& a, int64_t c0, float (*tile)[kSolarHours + 4]) {
    static_assert(K % kSolarHours == 0, "whole 32-hour rounds");
    const int lane = threadIdx.x & 63;
    const int64_t c = c0 + lane;
    const bool valid = c < a.n_top;
    const int ncells = (int)min<int64_t>(64, a.n_top - c0);
    float ss = 0.0f, cs = 1.0f, sa = 0.0f, ca = 1.0f;
    if (valid) {
        sincosf(a.slopes[c], &ss, &cs);
        sincosf(a.aspects[c], &sa, &ca);
    }
    for (int h0 = 0; h0 < K; h0 += kSolarHours) {
#pragma unroll
        for (int u = 0; u < kSolarHours; u += 4) {
            float out[4];
#pragma unroll
            for (int w = 0; w < 4; ++w) {
                const int h = h0 + u + w;
                out[w] = 0.0f;
                if (valid && h < a.nh) {
                    const HourConst& hc = a.hours[h];
                    out[w] = hour_irradiance(hc, cs, ss, ca, sa, svf_blocked(a.svf, c, hc.dense, a.n_total));
                }
            }
            *reinterpret_cast<float4*>(&tile[lane][u]) = make_float4(out[0], out[1], out[2], out[3]);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
#pragma unroll
        for (int r = 0; r < 8; ++r) {   // 64 cells x 8 pieces of 16 bytes
            const int j = r * 64 + lane, cell = j >> 3, part = j & 7;
            if (cell < ncells) {
                float* dst = a.hourly_top + (int64_t)row_slot((unsigned)(c0 + cell), a.row_swz) * K + h0 + part * 4;
                *reinterpret_cast<float4*>(dst) = *reinterpret_cast<const float4*>(&tile[cell][part * 4]);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");   // the tile is rewritten next round
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
}

template <int K>
__global__ __launch_bounds__(kSolarBlock) void k_solar_top(HourArgs a) {
    if constexpr (K % kSolarHours == 0) {
        __shared__ SolarLds L;
        const int wave = threadIdx.x >> 6;
        const int64_t c0 = (int64_t)blockIdx.x * kSolarBlock + wave * 64;
        if (c0 < a.n_top) solar_top_wave<K>(a, c0, L.t[wave]);
    } else {
        solar_top_cell<K>(a, (int64_t)blockIdx.x * kSolarBlock + threadIdx.x);
    }
}

struct Recv {
    float gx, gy, z, nx, ny, nz;
    float cs, ss, ca, sa;   // cos/sin of slope and aspect (the receiver's own hourly term)
};


/* Division of a cell id by the grid width via a float reciprocal and one fix-up: exact on
 * every supported grid (grid_supported: at most 2^22 rows, ids below 2^30), where the float
 * estimate is off by at most one row. */
__device__ __forceinline__ void cell_xy(int id, int W, float inv_w, int& x, int& y) {
    int q = (int)((float)id * inv_w);
    int r = id - q * W;
    if (r < 0) { --q; r += W; }
    else if (r >= W) { ++q; r -= W; }
    x = r;
    y = q;
}

/* The receiver geometry of voxel v (reflection_kernels.cu:55-63): cell, height, normal. */
__device__ __forceinline__ Recv receiver_of(const HourArgs& a, int64_t v, float inv_w) {
    Recv r;
    const int top = a.pix_parent[v];
    int gx, gy;
    cell_xy(top, a.W, inv_w, gx, gy);
    r.gx = (float)gx;
    r.gy = (float)gy;
    r.z = (v < a.n_top) ? a.grid[top] : a.wall_z[v - a.n_top];
    sincosf(a.slopes[v], &r.ss, &r.cs);
    sincosf(a.aspects[v], &r.sa, &r.ca);
    r.nx = r.ss * r.sa;
    r.ny = -r.ss * r.ca;
    r.nz = r.cs;
    return r;
}

/* Hour-invariant weight of neighbour cell `id` (0-based) for receiver r, 0 when the pair is
 * culled (reflection_kernels.cu:18-38): 0.2 cos_recv cos_send / (pi d^2), dx and dy in cells,
 * dz in metres. */
__device__ __forceinline__ float neighbour_weight(int id, const Recv& r, const float* __restrict__ grid, int W,
                                                  int cells, float inv_w) {
    if (id < 0 || id >= cells) return 0.0f;
    int nx, ny;
    cell_xy(id, W, inv_w, nx, ny);
    const float dx = (float)nx - r.gx, dy = (float)ny - r.gy, dz = grid[id] - r.z;
    const float d2 = dx * dx + dy * dy + dz * dz;
    if (d2 < 1e-4f) return 0.0f;
    const float inv_d = rsqrtf(d2);
    const float cr = (r.nx * dx + r.ny * dy + r.nz * dz) * inv_d;
    if (cr <= 0.0f) return 0.0f;
    const float cs = fmaxf(fabsf(dz) * inv_d, 0.1f);
    return (0.2f / 3.14159265f) * cr * cs * (inv_d * inv_d);
}

/* XCD-aware block order: block b runs on XCD b % 8 (nb, the grid, is a multiple of 8).
 * band == 0: each XCD walks one contiguous eighth of the voxels (its L2 holds its own
 * region); band > 0: the XCDs walk consecutive bands of `band` blocks side by side (XCD x
 * takes bands x, x + 8, x + 16, ...), so their neighbour gathers overlap in the shared
 * last-level cache.  A bijection onto [0, nb). */
__device__ __forceinline__ unsigned xcd_block(unsigned b, unsigned nb, unsigned band) {
    const unsigned per = nb / kXcds;
    const unsigned x = b % kXcds, s = b / kXcds;
    if (band == 0 || band >= per) return x * per + s;
    const unsigned full = per / band;
    if (s < full * band) return ((s / band) * kXcds + x) * band + s % band;
    const unsigned r = per - full * band;
    return full * band * kXcds + x * r + (s - full * band);
}

/* Lane-permuted copy of x (CDNA DPP); lanes the control does not write, or whose source lies
   outside their row, read 0. */
template <int CTRL, int ROW_MASK = 0xf, int BANK_MASK = 0xf>
__device__ __forceinline__ unsigned dpp0(unsigned x) {
    return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, CTRL, ROW_MASK, BANK_MASK, true);
}

/* Inclusive prefix sum (mod 2^32) over the 64 lanes: row shifts 1, 2, 3 of the input, then
   4 and 8 of the partial sums build the 16-lane row prefix; row_bcast15 / row_bcast31 carry
   the row totals into the rows above. */
__device__ __forceinline__ unsigned wave_scan_add(unsigned x) {
    unsigned s = x + dpp0<0x111>(x);        // row_shr:1
    s += dpp0<0x112>(x);                    // row_shr:2
    s += dpp0<0x113>(x);                    // row_shr:3
    s += dpp0<0x114, 0xf, 0xe>(s);          // row_shr:4, lanes 4..15 of each row
    s += dpp0<0x118, 0xf, 0xc>(s);          // row_shr:8, lanes 8..15 of each row
    s += dpp0<0x142, 0xa, 0xf>(s);          // row_bcast:15 into rows 1 and 3
    s += dpp0<0x143, 0xc, 0xf>(s);          // row_bcast:31 into rows 2 and 3
    return s;
}

/* Number of set bits of m below this lane. */
__device__ __forceinline__ int lanes_below(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

__device__ __forceinline__ int64_t uniform64(int64_t x) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

/* Each lane holds HPL consecutive hours of a row: K >= 16, one 16-byte load of 4 hours, LPE = K / 4
 * lanes covering a row and G = 64 / LPE rows per load instruction (K = 256: 1, K = 128: 2,
 * K = 64: 4); K < 16, one hour per lane, G = 64 / K rows.  Lane l adds w[i] * top[id[i]][its
 * hours] over the list entries i = l / LPE, l / LPE + G, ...; its partial sums over its hours
 * are returned.  OFF32 (the hourly buffer is below 4 GiB): rows are addressed by a 32-bit byte
 * offset from the buffer base, one shift-add per entry. */
template <int K, bool OFF32>
__device__ __forceinline__ float apply_list_at(const int* ids, const float* ws, int cnt,
                                               const float* __restrict__ top, int lane, int swz) {
    constexpr int HPL = (K >= 16) ? 4 : 1;
    constexpr int LPE = K / HPL;
    constexpr int G = 64 / LPE;
    typedef float vec __attribute__((ext_vector_type(HPL)));
    /* 16 VGPRs of rows in flight per wave: 4 loads of 16 bytes (K >= 16) or 16 of 4 bytes */
    constexpr int kUnroll = 8 / HPL;
    /* byte offset of this lane's hours within a row, and of row `id` */
    const unsigned lane_off = (unsigned)((lane % LPE) * HPL) * 4u;
    const char* lane_top = reinterpret_cast<const char*>(top) + lane_off;   // 64-bit rows: one shift-add each
    constexpr int kRowShift = __builtin_ctz(K * 4);
    auto load = [&](int cell) __attribute__((always_inline)) {
        const unsigned id = row_slot((unsigned)cell, swz);
        if constexpr (OFF32) {
            const unsigned off = (unsigned)id * (unsigned)(K * 4) + lane_off;
            return *reinterpret_cast<const vec*>(reinterpret_cast<const char*>(top) + off);
        } else {
            return *reinterpret_cast<const vec*>(lane_top + ((uint64_t)id << kRowShift));
        }
    };
    vec a = {}, b = {};
    int i = lane / LPE;
    if constexpr (!OFF32) {
        /* 64-bit rows: the batch's loads issued before its FMAs (the scheduler otherwise pairs
           them), entries still alternating between a and b in list order */
        constexpr int B = 2 * kUnroll;
        for (; i + (B - 1) * G < cnt; i += B * G) {
            vec x[B];
#pragma unroll
            for (int u = 0; u < B; ++u) x[u] = load(ids[i + u * G]);
            __builtin_amdgcn_sched_group_barrier(0x0020, B, 0);   // the VMEM reads first
#pragma unroll
            for (int u = 0; u < B; ++u) {
                const float w = ws[i + u * G];
#pragma unroll
                for (int h = 0; h < HPL; ++h) {
                    if (u & 1) b[h] = fmaf(w, x[u][h], b[h]);
                    else a[h] = fmaf(w, x[u][h], a[h]);
                }
            }
        }
    }
#pragma unroll kUnroll
    for (; i + G < cnt; i += 2 * G) {
        const vec x = load(ids[i]);
        const vec y = load(ids[i + G]);
        const float wx = ws[i], wy = ws[i + G];
#pragma unroll
        for (int u = 0; u < HPL; ++u) {
            a[u] = fmaf(wx, x[u], a[u]);
            b[u] = fmaf(wy, y[u], b[u]);
        }
    }
    if (i < cnt) {
        const vec x = load(ids[i]);
#pragma unroll
        for (int u = 0; u < HPL; ++u) a[u] = fmaf(ws[i], x[u], a[u]);
    }
    float s = 0.0f;
#pragma unroll
    for (int u = 0; u < HPL; ++u) s += a[u] + b[u];
    return s;
}

template <int K>
__device__ __forceinline__ float apply_list(const int* ids, const float* ws, int cnt, const float* __restrict__ top,
                                            int lane, bool off32, int swz) {
    static_assert(K >= 1 && K <= 256 && (K & (K - 1)) == 0, "a power of two hours per pass");
    return off32 ? apply_list_at<K, true>(ids, ws, cnt, top, lane, swz)
                 : apply_list_at<K, false>(ids, ws, cnt, top, lane, swz);
}

/* Rows slots[i] (row slots of the hourly buffer) of a merged list, added with NW = 1 or 2 weights
 * into as many sums (lanes hold hours as in apply_list_at, K >= 16: four hours per lane; a row per
 * load instruction at K = 256, its slot read into a scalar register). */
template <int K, bool OFF32, int NW>
__device__ __forceinline__ void apply_slots_at(const int* slots, const float* w0, const float* w1, int cnt,
                                               const float* __restrict__ top, int lane, float& acc0, float& acc1) {
    static_assert(K >= 16, "four hours per lane");
    constexpr int LPE = K / 4;
    constexpr int G = 64 / LPE;
    typedef float vec __attribute__((ext_vector_type(4)));
    constexpr int B = 4;   // 16 VGPRs of rows in flight
    const unsigned lane_off = (unsigned)((lane % LPE) * 4) * 4u;
    constexpr int kRowShift = __builtin_ctz(K * 4);
    auto load = [&](int slot) __attribute__((always_inline)) {
        if constexpr (G == 1) slot = __builtin_amdgcn_readfirstlane(slot);
        const char* row = reinterpret_cast<const char*>(top) +
                          (OFF32 ? (uint64_t)((unsigned)slot * (unsigned)(K * 4)) : ((uint64_t)(unsigned)slot << kRowShift));
        return *reinterpret_cast<const vec*>(row + lane_off);
    };
    vec s0 = {}, s1 = {};
    auto add = [&](const vec& x, int e) __attribute__((always_inline)) {
        const float a = w0[e];
#pragma unroll
        for (int h = 0; h < 4; ++h) s0[h] = fmaf(a, x[h], s0[h]);
        if constexpr (NW == 2) {
            const float b = w1[e];
#pragma unroll
            for (int h = 0; h < 4; ++h) s1[h] = fmaf(b, x[h], s1[h]);
        }
    };
    int i = lane / LPE;
    for (; i + (B - 1) * G < cnt; i += B * G) {
        vec x[B];
#pragma unroll
        for (int u = 0; u < B; ++u) x[u] = load(slots[i + u * G]);
        __builtin_amdgcn_sched_group_barrier(0x0020, B, 0);   // the VMEM reads first
#pragma unroll
        for (int u = 0; u < B; ++u) add(x[u], i + u * G);
    }
    for (; i < cnt; i += G) add(load(slots[i]), i);
    acc0 += (s0[0] + s0[1]) + (s0[2] + s0[3]);
    if constexpr (NW == 2) acc1 += (s1[0] + s1[1]) + (s1[2] + s1[3]);
}

/* A receiver's compacted list: row slots gathered four rows at a time (SLOTS, K >= 16), or cell
   ids by apply_list. */
template <int K, bool SLOTS>
__device__ __forceinline__ float apply_receiver_list(const int* list, const float* ws, int cnt,
                                                     const float* __restrict__ top, int lane, bool off32, int swz) {
    if constexpr (SLOTS) {
        float acc = 0.0f, unused = 0.0f;
        if (off32) apply_slots_at<K, true, 1>(list, ws, nullptr, cnt, top, lane, acc, unused);
        else apply_slots_at<K, false, 1>(list, ws, nullptr, cnt, top, lane, acc, unused);
        return acc;
    } else {
        return apply_list<K>(list, ws, cnt, top, lane, off32, swz);
    }
}

/* ba-RLE bytes staged in LDS per workgroup, split evenly over its waves. */
constexpr int kStageBytes = 8192;
/* compacted (id, weight) entries per wave; the list is applied once it may not take another 64 */
constexpr int kListCap = 128;

/* Wave-wide copy of src[lo, hi) into LDS (lo and src 16-byte aligned). */
__device__ __forceinline__ void stage_copy(uint8_t* dst, const uint8_t* __restrict__ src, int64_t lo, int64_t hi,
                                           int lane) {
    if (hi <= lo) return;
    const int n = (int)(hi - lo);
    const int full = n & ~15;
    const uint8_t* s = src + lo;
    for (int o = lane * 16; o < full; o += 64 * 16)
        *reinterpret_cast<uint4*>(dst + o) = *reinterpret_cast<const uint4*>(s + o);
    for (int o = full + lane; o < n; o += 64) dst[o] = s[o];
}

/* LDS writes of one lane visible to the other lanes of its wave (and no reordering across). */
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

/* One wave's window on its stretch of the ba-RLE stream: bytes [lo, hi) held in LDS. */
struct StreamWindow {
    uint8_t* lds;
    const uint8_t* g;   // the whole stream (16-byte aligned)
    int64_t lo, hi;
    int64_t end;        // the window may run ahead up to here (valid stream bytes)
    int cap;

    __device__ __forceinline__ bool fits(int64_t a, int64_t b) const { return b - a <= cap - 15; }
    /* make [a, b) resident (b - a <= cap - 15, b <= end) */
    __device__ __forceinline__ void cover(int64_t a, int64_t b, int lane) {
        if (a < lo || b > hi) {
            wave_lds_sync();
            lo = a & ~(int64_t)15;
            hi = min(lo + (int64_t)cap, max(end, b));
            stage_copy(lds, g, lo, hi, lane);
            wave_lds_sync();
        }
    }
};

/* The ba-RLE stream [s0, s1) of one voxel, decoded by the whole wave block by block; each
 * block is made resident in the LDS window first (a block larger than the window is read
 * from global memory).  Block layout: 3 bytes of 3-bit width codes (7 = unused), one length
 * byte per run, payload (4-bit deltas two per byte, high nibble first; wider ones big-endian). */
template <typename F>
__device__ __forceinline__ void wave_decode(StreamWindow& win, int64_t s0, int64_t s1, int lane, F&& f) {
    unsigned prev = 0;
    bool first = true;
    for (int64_t pos = s0; pos < s1;) {
        win.cover(pos, min(pos + 11, s1), lane);   // codes + at most 8 lengths
        const uint8_t* h = win.lds + (pos - win.lo);
        const unsigned bits = (unsigned)__builtin_amdgcn_readfirstlane(
            (int)(((unsigned)h[0] << 16) | ((unsigned)h[1] << 8) | (unsigned)h[2]));
        int nc = 0;
        unsigned wid[8];
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const unsigned cd = (bits >> (21 - 3 * j)) & 7u;
            wid[j] = cd;
            if (cd != 7u && nc == j) nc = j + 1;
        }
        /* run lengths, element prefix and payload offsets within the block (wave-uniform) */
        int estart[9], pstart[8];
        estart[0] = 0;
        int pay = 3 + nc;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int len = (j < nc) ? __builtin_amdgcn_readfirstlane((int)h[3 + j]) : 0;
            estart[j + 1] = estart[j] + len;
            pstart[j] = pay;
            pay += (wid[j] == 0u) ? (len + 1) >> 1 : len * (int)wid[j];
        }
        const bool resident = win.fits(pos, pos + pay);
        if (resident) win.cover(pos, pos + pay, lane);
        const uint8_t* p = resident ? win.lds + (pos - win.lo) : win.g + pos;
        const int total = estart[8];
        for (int e0 = 0; e0 < total; e0 += 64) {   // 64 elements per step
            const int e = e0 + lane;
            unsigned delta = 0;
            if (e < total) {
                /* the element's run: the last run starting at or before e (a select chain,
                   so the per-run tables stay in registers) */
                unsigned w = wid[0];
                int es = 0, ps = pstart[0];
#pragma unroll
                for (int j = 1; j < 8; ++j)
                    if (e >= estart[j]) { w = wid[j]; es = estart[j]; ps = pstart[j]; }
                const int k = e - es;
                if (w == 0u) {
                    const unsigned byte = p[ps + (k >> 1)];
                    delta = (k & 1) ? (byte & 15u) : (byte >> 4);
                } else {
                    const int q = ps + k * (int)w;
                    for (unsigned b = 0; b < w; ++b) delta = (delta << 8) | p[q + (int)b];
                }
                if (first && e == 0) delta = 0u - delta;   // first value is absolute
            }
            const unsigned val = prev - wave_scan_add(delta);   // deltas are summed mod 2^32
            f(e < total && val > 0u, (int)val - 1);   // valid: a non-zero value, id = value - 1
            prev = (unsigned)__builtin_amdgcn_readlane((int)val, 63);   // lanes past the end add zero
            first = false;
        }
        pos += pay;
    }
}

/* One wave's LDS: the stream window, the compacted (id, weight) list, decoded ids not yet
 * weighed. */
template <int FMT, int WAVES, int STAGE = kStageBytes>
struct WaveLds {
    static constexpr int kCap = (FMT == SGPU_FMT_BARLE) ? STAGE / WAVES : 16;
    int list_id[WAVES][kListCap];
    float list_w[WAVES][kListCap];
    __align__(16) uint8_t stage[WAVES][kCap];
    int pend_id[WAVES][FMT == SGPU_FMT_BARLE ? 128 : 1];
};

/* Lane l's part of a receiver's own irradiance over the pass: hours l, l + 64, ... */
template <int K>
__device__ __forceinline__ float own_part(const HourArgs& a, const HourConst* hours, const Recv& r, int64_t v,
                                          int lane) {
    float own = 0.0f;
    for (int h = lane; h < K && h < a.nh; h += 64) {
        const HourConst& hc = hours[h];
        own += hour_irradiance(hc, r.cs, r.ss, r.ca, r.sa, svf_blocked(a.svf, v, hc.dense, a.n_total));
    }
    return own;
}

/* The ba-RLE stream [s0, s1) decoded by the wave, its ids handed to take(valid, id) 64 at a time
 * (a block holds ~30 ids: decoded ids wait in pend_id, the wave's 128 slots, so that the weights
 * are computed on full 64-id chunks; the last call may have fewer valid lanes). */
template <typename F>
__device__ __forceinline__ void decode_chunks(StreamWindow& win, int64_t s0, int64_t s1, int* pend_id, int lane,
                                              F&& take) {
    int pend = 0;
    auto gather = [&](bool valid, int id) __attribute__((always_inline)) {
        const uint64_t m = __ballot(valid);
        if (valid) pend_id[pend + lanes_below(m)] = id;
        pend += __popcll(m);
        if (pend >= 64) {
            wave_lds_sync();
            const int x = pend_id[lane];
            const bool more = 64 + lane < pend;
            const int y = more ? pend_id[64 + lane] : 0;
            wave_lds_sync();
            if (more) pend_id[lane] = y;
            pend -= 64;
            take(true, x);
        }
    };
    wave_decode(win, s0, s1, lane, gather);
    if (pend > 0) {
        wave_lds_sync();
        const bool valid = lane < pend;
        take(valid, valid ? pend_id[lane] : -1);
    }
}

/* One receiver's pass, by the whole wave: lane l's part of the receiver's own irradiance (hours
 * l, l + 64, ... of the pass) and of the reflected irradiance summed over the pass's hours.
 * Shared by the static and the queued kernels, so every receiver's total does not depend on
 * which wave takes it (bit-identical annual sums under every receiver assignment).
 * list_id / list_w: the wave's kListCap compacted entries; pend_id: its 128 decoded ids not yet
 * weighed (ba-RLE). */
template <int K, int FMT>
__device__ __forceinline__ void receiver_parts(const HourArgs& a, const HourConst* hours, const Recv& r, int64_t v,
                                               int64_t s0, int64_t s1, StreamWindow& win, int* list_id, float* list_w,
                                               int* pend_id, int lane, bool off32, float inv_w, int cells, float& own,
                                               float& refl) {
    own = own_part<K>(a, hours, r, v, lane);
    /* ba-RLE lists at K >= 16 are compacted as row slots and gathered four rows at a time (on the
       2048^2 scene the CSR walk ran 3% slower that way and keeps cell ids) */
    constexpr bool kSlots = K >= 16 && FMT == SGPU_FMT_BARLE;
    float acc = 0.0f;
    int cnt = 0;
    auto take = [&](bool valid, int id) __attribute__((always_inline)) {
        const float w = valid ? neighbour_weight(id, r, a.grid, a.W, cells, inv_w) : 0.0f;
        const bool keep = w > 0.0f;
        const uint64_t m = __ballot(keep);
        if (keep) {
            const int at = cnt + lanes_below(m);
            list_id[at] = kSlots ? (int)row_slot((unsigned)id, a.row_swz) : id;
            list_w[at] = w;
        }
        cnt += __popcll(m);
        if (cnt > kListCap - 64) {
            wave_lds_sync();
            acc += apply_receiver_list<K, kSlots>(list_id, list_w, cnt, a.hourly_top, lane, off32, a.row_swz);
            wave_lds_sync();
            cnt = 0;
        }
    };
    if constexpr (FMT == SGPU_FMT_CSR) {
        for (int64_t b = s0; b < s1; b += 64) {
            const int64_t k = b + lane;
            const bool valid = k < s1;
            take(valid, valid ? a.nb_ids[k] - 1 : -1);
        }
    } else {
        decode_chunks(win, s0, s1, pend_id, lane, take);
    }
    if (cnt > 0) {
        wave_lds_sync();
        acc += apply_receiver_list<K, kSlots>(list_id, list_w, cnt, a.hourly_top, lane, off32, a.row_swz);
        wave_lds_sync();
    }
    refl = acc;
}

/* The pass's hour constants (nh <= K of them, 48 bytes each) copied into LDS by the workgroup;
   the caller synchronises before use. */
template <int K>
__device__ __forceinline__ void stage_hours(const HourArgs& a, HourConst* lds) {
    static_assert(sizeof(HourConst) == 48, "three 16-byte pieces per hour");
    const int pieces = min(a.nh, K) * 3;
    const float4* src = reinterpret_cast<const float4*>(a.hours);
    float4* dst = reinterpret_cast<float4*>(lds);
    for (int i = threadIdx.x; i < pieces; i += blockDim.x) dst[i] = src[i];
}

__device__ __forceinline__ float wave_sum(float x) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) x += __shfl_xor(x, m, 64);
    return x;
}

/* total += own + reflected of the pass, FP32 Kahan (the annual accumulator). */
__device__ __forceinline__ void kahan_add(float* sum, float* comp, int64_t q, float x) {
    float s = sum[q], c = comp[q];
    const float y = x - c;
    const float t = s + y;
    c = (t - s) - y;
    s = t;
    sum[q] = s;
    comp[q] = c;
}

/* WAVES waves per workgroup (2 or 4): the knob the autotuner picks.
 * Static assignment: each wave takes 64 receivers (receiver groups interleave them, see
 * HourArgs::group) and walks them one after another.
 * ba-RLE: the 64 voxels of a wave own one contiguous stretch of the stream; it passes through
 * an LDS window of kStageBytes / WAVES bytes filled with 16-byte loads, block by block.
 * DBG (K = 1 only): write the hour's own and reflected irradiance instead of accumulating. */
template <int K, int FMT, bool DBG, int WAVES>
__device__ __forceinline__ void reflect_block(const HourArgs& a, unsigned lb) {
    __shared__ Recv rec[WAVES][64];
    __shared__ int64_t beg[WAVES][64];   // each receiver's list [beg, fin)
    __shared__ int64_t fin[WAVES][64];
    __shared__ WaveLds<FMT, WAVES> L;
    __shared__ HourConst hrs[K];

    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t n = a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total);
    /* this wave's receivers: positions first + j * S, j = 0..63 */
    int64_t first, S;
    if (a.group > 1) {
        S = (int64_t)a.group * WAVES;
        first = (int64_t)(lb / (unsigned)a.group) * S * 64 + (int64_t)(lb % (unsigned)a.group) * WAVES + wave;
    } else {
        S = 1;
        first = ((int64_t)lb * WAVES + wave) * 64;
    }
    const int nv = (first < n) ? (int)min<int64_t>(64, (n - first + S - 1) / S) : 0;
    const int64_t p = first + (int64_t)lane * S;
    const bool live = lane < nv;
    /* v: the voxel (geometry); q: its index into nb_off, sum, comp and dbg_* -- the voxel id
       in full-scene order mode, else the position (subset / chunk-local lists) */
    int64_t v = 0, q = 0;
    if (live) {
        if (a.order) {
            v = a.order[p];
            q = v;
        } else {
            v = a.subset ? (int64_t)a.subset[p] : a.v_base + p;
            q = p;
        }
    }
    const float inv_w = 1.0f / (float)a.W;
    const int cells = a.W * a.H;
    const bool off32 = a.hourly_top_cap <= ((int64_t)1 << 30);   // below 4 GiB: 32-bit row offsets

    /* phase 1: lane = voxel */
    stage_hours<K>(a, hrs);
    {
        Recv r{};
        if (live) {
            r = receiver_of(a, v, inv_w);
            beg[wave][lane] = a.nb_off[q];
            fin[wave][lane] = a.nb_off[q + 1];
        }
        rec[wave][lane] = r;
    }
    __syncthreads();

    /* phase 2: one receiver at a time */
    float total = 0.0f, own_j = 0.0f, refl_j = 0.0f;   // lane j: receiver j's sums over the pass
    /* consecutive receivers (S == 1) mostly have neighbouring streams: the window runs ahead
       over the buffer; interleaved receivers (S > 1) each stage only their own list */
    const int64_t wend = (S == 1) ? a.nb_len : -1;
    StreamWindow win{L.stage[wave], a.nb_bytes, 0, 0, 0, WaveLds<FMT, WAVES>::kCap};
    for (int j = 0; j < nv; ++j) {
        const Recv r = rec[wave][j];
        const int64_t vj = __shfl(v, j, 64);
        const int64_t s0 = uniform64(beg[wave][j]), s1 = uniform64(fin[wave][j]);
        win.end = wend >= 0 ? wend : s1;
        float own, refl;
        receiver_parts<K, FMT>(a, hrs, r, vj, s0, s1, win, L.list_id[wave], L.list_w[wave], L.pend_id[wave], lane, off32,
                               inv_w, cells, own, refl);
        if constexpr (DBG) {
            own = wave_sum(own);
            refl = wave_sum(refl);
            if (lane == j) { own_j = own; refl_j = refl; }
        } else {
            const float t = wave_sum(own + refl);
            if (lane == j) total = t;
        }
    }

    /* phase 3: lane = voxel */
    if (!live) return;
    if constexpr (DBG) {
        a.dbg_hourly[q] = own_j;
        a.dbg_reflected[q] = refl_j;
        return;
    }
    kahan_add(a.sum, a.comp, q, total);
}

template <int K, int FMT, bool DBG, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_reflect(HourArgs a) {
    reflect_block<K, FMT, DBG, WAVES>(a, xcd_block(blockIdx.x, gridDim.x, (unsigned)a.band));
}

/* Position of the item-th receiver XCD x takes from its queue: XCD x walks bands x, x + 8,
 * x + 16, ... of `band` consecutive positions of the processing order (band == 0: one contiguous
 * eighth of the positions per XCD). */
__device__ __forceinline__ int64_t queue_position(uint64_t item, unsigned x, int64_t n, int64_t band) {
    if (band <= 0) {
        const int64_t per = (n + kXcds - 1) / kXcds;
        return item < (uint64_t)per ? (int64_t)x * per + (int64_t)item : n;
    }
    const int64_t k = (int64_t)(item / (uint64_t)band), o = (int64_t)(item % (uint64_t)band);
    return (k * kXcds + (int64_t)x) * band + o;
}

/* Queued assignment (the full scene in processing order, or a subset of voxels given in that
 * order): a persistent grid whose waves each take the next
 * receiver from a queue of their XCD (block b is on XCD b % 8), so that at any moment the
 * resident waves of an XCD work on a compact window of the processing order and share its
 * neighbours' rows in the L2.  An XCD's positions are split over kQueuesPerXcd interleaved
 * sub-queues (item i of sub-queue c is the XCD's (i * kQueuesPerXcd + c)-th position), each a
 * counter on its own 256-byte line zeroed before the launch, wave w of the XCD pulling from
 * sub-queue w % kQueuesPerXcd: a single counter per XCD serialises its atomics.  A wave asks for
 * its next item before it works on the current one.  Positions grow with the item, so a wave
 * leaves at its first position past the end: every wave reaches the exit, whatever the grid. */
/* The queued kernel's stream window: a receiver's own stream only (~300 bytes on average), 1 KB
   per wave; blocks larger than that are read from global memory. */
constexpr int kQueueStageBytes = 4096;

template <int K, int FMT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(8)))
void k_reflect_queue(HourArgs a) {
    __shared__ WaveLds<FMT, WAVES, kQueueStageBytes> L;
    __shared__ HourConst hrs[K];
    stage_hours<K>(a, hrs);
    __syncthreads();
    /* the wave index as a wave-uniform (scalar) value: the counter address must be scalar for the
       compiler to combine the lanes' atomic into one */
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const unsigned x = blockIdx.x % kXcds;
    const unsigned c = ((blockIdx.x / kXcds) * WAVES + (unsigned)wave) % kQueuesPerXcd;
    unsigned* counter = a.queue + (x * kQueuesPerXcd + c) * kQueueStride;
    /* positions: the full scene in processing order (voxel order[p]; lists and sums by voxel), or a
       subset (voxel subset[p]; lists and sums by position) */
    const int64_t n = a.order ? a.n_total : a.n_sub;
    const float inv_w = 1.0f / (float)a.W;
    const int cells = a.W * a.H;
    const bool off32 = a.hourly_top_cap <= ((int64_t)1 << 30);
    StreamWindow win{L.stage[wave], a.nb_bytes, 0, 0, 0, WaveLds<FMT, WAVES, kQueueStageBytes>::kCap};
    /* every lane takes part in the atomic (only lane 0 adds), so no lane-divergent branch
       precedes the broadcast: lane 0's old value is the wave's item.  (With the atomic inside
       `if (lane == 0)`, the compiler split the loop so that lanes 1..63 re-entered the body with
       a stale item and never left.) */
    unsigned pending = atomicAdd(counter, lane == 0 ? 1u : 0u);
    for (int64_t turn = 0; turn <= n; ++turn) {   // each turn takes a new item: at most n + 1
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)pending);
        const int64_t p = queue_position((uint64_t)item * kQueuesPerXcd + c, x, n, a.queue_band);
        if (p >= n) break;
        pending = atomicAdd(counter, lane == 0 ? 1u : 0u);   // the next item, while this one is worked
        const int64_t v = a.order ? (int64_t)a.order[p] : (int64_t)a.subset[p];
        const int64_t q = a.order ? v : p;
        const Recv r = receiver_of(a, v, inv_w);
        const int64_t s0 = uniform64(a.nb_off[q]), s1 = uniform64(a.nb_off[q + 1]);
        win.end = s1;
        float own, refl;
        receiver_parts<K, FMT>(a, hrs, r, v, s0, s1, win, L.list_id[wave], L.list_w[wave], L.pend_id[wave], lane, off32,
                               inv_w, cells, own, refl);
        const float t = wave_sum(own + refl);
        if (lane == 0) kahan_add(a.sum, a.comp, q, t);
    }
}

__global__ void k_clear_queue(unsigned* q) {
    if (threadIdx.x < kXcds * kQueuesPerXcd) q[threadIdx.x * kQueueStride] = 0u;
}

/* ---- receiver pairs: one wave walks two consecutive positions and gathers each shared row once ---- */

/* Each receiver's own irradiance over the pass, for the paired walk: own[q] = the sum over the pass's
 * hours, one thread per position (voxel v = p in full-scene order, else subset[p]; q = p). */
template <int K>
__device__ __forceinline__ float own_sum(const HourArgs& a, const HourConst* hrs, int64_t v, float inv_w) {
    const Recv r = receiver_of(a, v, inv_w);
    const int nh = min(K, a.nh);
    float s = 0.0f;
    int h = 0;
    /* eight hours at a time: their SVF bytes requested together (a bin outside the dense set reads
       byte 0 and is not blocked), then added in hour order */
    for (; h + 8 <= nh; h += 8) {
        unsigned byte[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int d = max(hrs[h + u].dense, 0);
            byte[u] = a.svf[(int64_t)(d >> 3) * a.n_total + v];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const HourConst& hc = hrs[h + u];
            const bool blocked = hc.dense >= 0 && ((byte[u] >> (hc.dense & 7)) & 1u);
            s += hour_irradiance(hc, r.cs, r.ss, r.ca, r.sa, blocked);
        }
    }
    for (; h < nh; ++h) {
        const HourConst& hc = hrs[h];
        s += hour_irradiance(hc, r.cs, r.ss, r.ca, r.sa, svf_blocked(a.svf, v, hc.dense, a.n_total));
    }
    return s;
}

template <int K>
__global__ __launch_bounds__(256) void k_own_pass(HourArgs a, float* __restrict__ own) {
    __shared__ HourConst hrs[K];
    stage_hours<K>(a, hrs);
    __syncthreads();
    const int64_t n = a.order ? a.n_total : a.n_sub;
    const float inv_w = 1.0f / (float)a.W;
    for (int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x; p < n; p += (int64_t)gridDim.x * 256)
        own[p] = own_sum<K>(a, hrs, a.order ? p : (int64_t)a.subset[p], inv_w);
}

/* Per-wave stream window of the paired kernels, and the entries of each kept list held in LDS (the
   rest go to the wave's scratch): 240 keeps 7 workgroups of 4 waves per CU for the batched kernel,
   whose waves also hold their tile's arguments in LDS. */
constexpr int kGroupStageBytes = 1024;
constexpr int kPairLds = 240;

/* One wave's LDS in the paired kernels: the first C entries of both kept lists; while the
 * streams are decoded, the stream window and the decoded ids not yet weighed; while the lists are
 * merged, their 64-entry heads and, per entry of the first head, the second receiver's weight.  A
 * pair walked receiver by receiver uses lid[0] / lw[0] as its compacted list. */
template <int C>
struct PairLdsT {
    int lid[2][C];
    float lw[2][C];
    union {
        struct {
            __align__(16) uint8_t stage[kGroupStageBytes];
            int pend_id[128];
        } dec;
        struct {
            int cid[2][64];
            float cw[2][64];
            float wx[64];
        } mrg;
    } u;
};
static_assert(kPairLds >= kListCap, "the receiver-by-receiver walk's list fits lid[0]");
static_assert(sizeof(PairLdsT<kPairLds>) == 16 * kPairLds + 1536, "LDS per wave of the paired kernels");

/* Position p of the full scene in processing order (voxel order[p], lists and sums by voxel) or
   of a subset (voxel subset[p], lists and sums by position): its voxel and its list / sum index. */
__device__ __forceinline__ void queue_voxel(const HourArgs& a, int64_t p, int64_t& v, int64_t& q) {
    v = a.order ? (int64_t)a.order[p] : (int64_t)a.subset[p];
    q = a.order ? v : p;
}

/* Index of the first of the n descending entries c[0..n) that is <= x (n <= 64): the number greater. */
__device__ __forceinline__ int rank_desc(const int* c, int n, int x) {
    int pos = 0;
#pragma unroll
    for (int st = 64; st >= 1; st >>= 1)
        if (pos + st <= n && c[pos + st - 1] > x) pos += st;
    return pos;
}

/* One wave's walk of a receiver pair (voxels v0, v1; list / sum indices q0, q1; the second absent
 * unless `two`), the body of the paired kernels:
 * Phase 1: each receiver's ba-RLE stream is decoded and weighed; its kept (id, weight) entries, in
 *   stream order (descending ids), go to list j: entries k < C in LDS, the rest to the wave's
 *   scratch (list j's part from sb_j).
 * Phase 2: the two descending lists are merged 64-entry heads at a time.  With t the larger of the
 *   64th head entries of the lists that have more than 64 left, every entry >= t of either list is
 *   in its head (later entries are smaller than their list's 64th); those are taken: the first
 *   head's with the second receiver's weight where the second head holds the id (binary search),
 *   the second head's where the first does not hold it.  The list that set t gives all 64 of its
 *   head, so every round makes progress.  Rows of the first head's entries are added into both sums
 *   (a zero weight where the second list lacks the id); the second head's remaining rows into the
 *   second receiver's only.
 * A pair whose lists overflow the scratch (or exceed a.grp_limit) is walked receiver by receiver
 * (receiver_parts, the hour constants read from global memory).  Each receiver's own irradiance is
 * a.grp_own[q]; the pass total enters the annual Kahan sum. */
template <int K, int C>
__device__ __forceinline__ void pair_walk(const HourArgs& a, int64_t v0, int64_t q0, int64_t v1, int64_t q1, bool two,
                                          PairLdsT<C>& L, StreamWindow& win, int2* scr, int cap, int lane) {
    const float inv_w = 1.0f / (float)a.W;
    const int cells = a.W * a.H;
    const bool off32 = a.hourly_top_cap <= ((int64_t)1 << 30);
    win.lo = win.hi = 0;   // a previous merge overwrote the window's bytes

    /* phase 1 */
    int nl[2] = {0, 0}, sb[2] = {0, 0};
    int spill = 0;
#pragma unroll
    for (int j = 0; j < 2; ++j) {
        if (j == 1 && !two) continue;
        const int64_t v = j == 0 ? v0 : v1, q = j == 0 ? q0 : q1;
        const Recv r = receiver_of(a, v, inv_w);
        const int64_t s0 = uniform64(a.nb_off[q]), s1 = uniform64(a.nb_off[q + 1]);
        win.end = s1;
        int cnt = 0;
        sb[j] = spill;
        auto take = [&](bool valid, int id) __attribute__((always_inline)) {
            const float w = valid ? neighbour_weight(id, r, a.grid, a.W, cells, inv_w) : 0.0f;
            const bool keep = w > 0.0f;
            const uint64_t m = __ballot(keep);
            const int at = cnt + lanes_below(m);
            if (keep) {
                if (at < C) {
                    L.lid[j][at] = id;
                    L.lw[j][at] = w;
                } else if (spill + (at - C) < cap) {
                    scr[spill + (at - C)] = make_int2(id, __float_as_int(w));
                }
            }
            cnt += __popcll(m);
        };
        decode_chunks(win, s0, s1, L.u.dec.pend_id, lane, take);
        nl[j] = cnt;
        spill += max(cnt - C, 0);
    }
    if (spill > cap || (a.grp_limit > 0 && nl[0] + nl[1] > a.grp_limit)) {
        /* walked receiver by receiver, as k_reflect_queue does (the hour constants read from
           global memory) */
        for (int j = 0; j < (two ? 2 : 1); ++j) {
            const int64_t v = j == 0 ? v0 : v1, q = j == 0 ? q0 : q1;
            const Recv r = receiver_of(a, v, inv_w);
            const int64_t s0 = uniform64(a.nb_off[q]), s1 = uniform64(a.nb_off[q + 1]);
            win.end = s1;
            float own, refl;
            wave_lds_sync();
            receiver_parts<K, SGPU_FMT_BARLE>(a, a.hours, r, v, s0, s1, win, L.lid[0], L.lw[0],
                                              L.u.dec.pend_id, lane, off32, inv_w, cells, own, refl);
            const float t = wave_sum(own + refl);
            if (lane == 0) kahan_add(a.sum, a.comp, q, t);
        }
        return;
    }

    /* phase 2 (the scratch entries and LDS lists written by other lanes are visible after the
       fence of wave_lds_sync) */
    wave_lds_sync();
    float acc0 = 0.0f, acc1 = 0.0f;
    int h0 = 0, h1 = 0;
    while (h0 < nl[0] || h1 < nl[1]) {
        int id0 = -1, id1 = -1;
        float w0 = 0.0f, w1 = 0.0f;
        {
            const int k = h0 + lane;
            if (k < nl[0]) {
                if (k < C) {
                    id0 = L.lid[0][k];
                    w0 = L.lw[0][k];
                } else {
                    const int2 e = scr[sb[0] + (k - C)];
                    id0 = e.x;
                    w0 = __int_as_float(e.y);
                }
            }
        }
        {
            const int k = h1 + lane;
            if (k < nl[1]) {
                if (k < C) {
                    id1 = L.lid[1][k];
                    w1 = L.lw[1][k];
                } else {
                    const int2 e = scr[sb[1] + (k - C)];
                    id1 = e.x;
                    w1 = __int_as_float(e.y);
                }
            }
        }
        int t = -1;
        if (nl[0] - h0 > 64) t = max(t, __builtin_amdgcn_readlane(id0, 63));
        if (nl[1] - h1 > 64) t = max(t, __builtin_amdgcn_readlane(id1, 63));
        const int tmin = max(t, 0);   // t < 0: both lists end within their heads
        const bool in0 = id0 >= tmin, in1 = id1 >= tmin;   // prefixes of the descending heads
        const int n0 = __popcll(__ballot(in0)), n1 = __popcll(__ballot(in1));
        wave_lds_sync();   // the previous round's reads of the heads are done
        L.u.mrg.cid[0][lane] = id0;
        L.u.mrg.cid[1][lane] = id1;
        L.u.mrg.cw[0][lane] = w0;
        L.u.mrg.cw[1][lane] = w1;
        wave_lds_sync();
        float wx = 0.0f;
        if (in0) {
            const int pos = rank_desc(L.u.mrg.cid[1], n1, id0);
            if (pos < n1 && L.u.mrg.cid[1][pos] == id0) wx = L.u.mrg.cw[1][pos];
        }
        bool rest = false;   // in the second head only
        if (in1) {
            const int pos = rank_desc(L.u.mrg.cid[0], n0, id1);
            rest = !(pos < n0 && L.u.mrg.cid[0][pos] == id1);
        }
        const uint64_t mr = __ballot(rest);
        const int nr = __popcll(mr);
        wave_lds_sync();   // the searches of both heads are done
        /* from here the heads hold row slots: the first's in place, the second's own compacted */
        L.u.mrg.cid[0][lane] = (int)row_slot((unsigned)id0, a.row_swz);
        L.u.mrg.wx[lane] = wx;
        if (rest) {
            const int at = lanes_below(mr);
            L.u.mrg.cid[1][at] = (int)row_slot((unsigned)id1, a.row_swz);
            L.u.mrg.cw[1][at] = w1;
        }
        wave_lds_sync();
        float unused = 0.0f;
        if (off32) {
            apply_slots_at<K, true, 2>(L.u.mrg.cid[0], L.u.mrg.cw[0], L.u.mrg.wx, n0, a.hourly_top, lane, acc0, acc1);
            apply_slots_at<K, true, 1>(L.u.mrg.cid[1], L.u.mrg.cw[1], nullptr, nr, a.hourly_top, lane, acc1, unused);
        } else {
            apply_slots_at<K, false, 2>(L.u.mrg.cid[0], L.u.mrg.cw[0], L.u.mrg.wx, n0, a.hourly_top, lane, acc0,
                                        acc1);
            apply_slots_at<K, false, 1>(L.u.mrg.cid[1], L.u.mrg.cw[1], nullptr, nr, a.hourly_top, lane, acc1,
                                        unused);
        }
        h0 += n0;
        h1 += n1;
    }
    const float r0 = wave_sum(acc0), r1 = wave_sum(acc1);
    if (lane == 0) {
        kahan_add(a.sum, a.comp, q0, a.grp_own[q0] + r0);
        if (two) kahan_add(a.sum, a.comp, q1, a.grp_own[q1] + r1);
    }
}

/* Receiver pairs, queued as k_reflect_queue but by pairs of consecutive positions (XCD bands of
 * queue_band / 2 pairs).  Receivers next to each other in the processing order share most of their
 * kept neighbours (a top cell and its walls, walls stacked on one cell, adjacent cells): a wave
 * that walks the two lists together gathers each shared row once (pair_walk).  Each receiver's own
 * irradiance comes from k_own_pass.  A receiver's sum differs from the receiver-by-receiver walk only
 * in the order of its additions, and does not depend on which wave takes the pair. */
template <int K, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(7)))
void k_reflect_pair(HourArgs a) {
    static_assert(K >= 16, "four hours per lane");
    __shared__ PairLdsT<kPairLds> PL[WAVES];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const unsigned x = blockIdx.x % kXcds;
    const unsigned c = ((blockIdx.x / kXcds) * WAVES + (unsigned)wave) % kQueuesPerXcd;
    unsigned* counter = a.queue + (x * kQueuesPerXcd + c) * kQueueStride;
    const int64_t n = a.order ? a.n_total : a.n_sub;
    const int64_t ng = (n + 1) / 2;
    const int64_t band = a.queue_band > 0 ? max<int64_t>(a.queue_band / 2, 1) : 0;
    const int cap = a.grp_cap;
    int2* scr = a.grp_scratch + ((int64_t)blockIdx.x * WAVES + wave) * cap;
    PairLdsT<kPairLds>& L = PL[wave];
    StreamWindow win{L.u.dec.stage, a.nb_bytes, 0, 0, 0, kGroupStageBytes};
    unsigned pending = atomicAdd(counter, lane == 0 ? 1u : 0u);   // every lane joins (see k_reflect_queue)
    for (int64_t turn = 0; turn <= ng; ++turn) {
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)pending);
        const int64_t g = queue_position((uint64_t)item * kQueuesPerXcd + c, x, ng, band);
        if (g >= ng) break;
        pending = atomicAdd(counter, lane == 0 ? 1u : 0u);
        const int64_t p0 = g * 2;
        const bool two = p0 + 1 < n;
        int64_t v0, q0, v1 = 0, q1 = 0;
        queue_voxel(a, p0, v0, q0);
        if (two) queue_voxel(a, p0 + 1, v1, q1);
        pair_walk<K, kPairLds>(a, v0, q0, v1, q1, two, L, win, scr, cap, lane);
    }
}

/* ---- batched tiles: one launch per pass covers every tile of a batch ---- */

/* The tile owning block b: the last tile whose first block is <= b (wave-uniform). */
__device__ __forceinline__ int tile_of_block(const BatchHour* d, int ntiles, int64_t b, bool solar) {
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if ((solar ? d[mid].top_block0 : d[mid].block0) <= b) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* The tile's arguments for pass `pass` (its hours pass * K ...); false once its hours are done. */
template <int K>
__device__ __forceinline__ bool tile_pass_args(const BatchHour& t, int pass, HourArgs& a) {
    const int h0 = pass * K;
    if (h0 >= t.nsel) return false;
    a = t.a;
    a.hours = t.a.hours + h0;
    a.nh = min(K, t.nsel - h0);
    return true;
}

template <int K>
__global__ __launch_bounds__(kSolarBlock) void k_solar_top_batch(const BatchHour* d, int ntiles, int pass) {
    const int t = tile_of_block(d, ntiles, blockIdx.x, true);
    HourArgs a;
    if (!tile_pass_args<K>(d[t], pass, a)) return;
    if constexpr (K % kSolarHours == 0) {
        __shared__ SolarLds L;
        const int wave = threadIdx.x >> 6;
        const int64_t c0 = ((int64_t)blockIdx.x - d[t].top_block0) * kSolarBlock + wave * 64;
        if (c0 < a.n_top) solar_top_wave<K>(a, c0, L.t[wave]);
    } else {
        solar_top_cell<K>(a, ((int64_t)blockIdx.x - d[t].top_block0) * kSolarBlock + threadIdx.x);
    }
}

/* XCD band of the batched reflection (xcd_block): the XCDs walk consecutive bands of 256 blocks
   side by side over the concatenated tiles (64 tiles of 128..1024 cells a side: hour loop 12.3
   -> 7.2 s against round-robin blocks; 7.4 s with one contiguous eighth per XCD). */
constexpr unsigned kBatchBand = 256;

/* The grid is a multiple of 8 (xcd_block); blocks past the last tile's own take no receivers. */
template <int K, int FMT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) void k_reflect_batch(const BatchHour* d, int ntiles, int pass) {
    const unsigned b = xcd_block(blockIdx.x, gridDim.x, kBatchBand);
    const int t = tile_of_block(d, ntiles, b, false);
    HourArgs a;
    if (!tile_pass_args<K>(d[t], pass, a)) return;   // the whole block leaves together
    reflect_block<K, FMT, false, WAVES>(a, (unsigned)((int64_t)b - d[t].block0));
}

/* The tile owning batch position p: the last tile whose first position is <= p (wave-uniform). */
__device__ __forceinline__ int tile_of_position(const BatchHour* d, int ntiles, int64_t p) {
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].pos0 <= p) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* Batched tiles, queued: k_reflect_queue over the concatenated positions of every tile (tile t owns
 * [pos0, pos0 + n_total), its voxel order[p - pos0]).  A wave takes a position from its XCD's
 * queue, finds the tile and runs the shared per-receiver routine on that tile's pass (the tile's
 * hour constants read from global memory: tiles have their own hours); tiles whose hours are done
 * are passed over.  The stream window stays valid across tiles: every tile's offsets index the
 * batch's one stream buffer. */
template <int K, int FMT, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(7)))
void k_reflect_batch_queue(const BatchHour* d, int ntiles, int pass, unsigned* queue, int64_t n, int64_t band) {
    __shared__ WaveLds<FMT, WAVES, kQueueStageBytes> L;
    __shared__ HourArgs targs[WAVES];   // each wave's current tile and pass, read field by field on use
    static_assert(sizeof(HourArgs) % 4 == 0, "copied as words");
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const unsigned x = blockIdx.x % kXcds;
    const unsigned c = ((blockIdx.x / kXcds) * WAVES + (unsigned)wave) % kQueuesPerXcd;
    unsigned* counter = queue + (x * kQueuesPerXcd + c) * kQueueStride;
    StreamWindow win{L.stage[wave], d[0].a.nb_bytes, 0, 0, 0, WaveLds<FMT, WAVES, kQueueStageBytes>::kCap};
    unsigned pending = atomicAdd(counter, lane == 0 ? 1u : 0u);   // every lane joins (see k_reflect_queue)
    /* the current tile and its positions [t_lo, t_hi), searched again only when a position falls
       outside them (consecutive items of a wave mostly stay in one tile); its pass arguments are
       copied into the wave's LDS slot then and read from there field by field (held in registers
       across the receiver routine they spilled to scratch) */
    int64_t t_lo = 0, t_hi = 0;
    bool live = false;   // the loaded tile has hours in this pass
    const HourArgs& a = targs[wave];
    for (int64_t turn = 0; turn <= n; ++turn) {
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)pending);
        const int64_t p = queue_position((uint64_t)item * kQueuesPerXcd + c, x, n, band);
        if (p >= n) break;
        pending = atomicAdd(counter, lane == 0 ? 1u : 0u);
        if (p < t_lo || p >= t_hi) {
            const int t = tile_of_position(d, ntiles, p);
            t_lo = d[t].pos0;
            t_hi = t_lo + d[t].a.n_total;
            live = pass * K < d[t].nsel;
            wave_lds_sync();   // the wave's reads of the previous tile's arguments are done
            const uint32_t* src = reinterpret_cast<const uint32_t*>(&d[t].a);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&targs[wave]);
            for (int i = lane; i < (int)(sizeof(HourArgs) / 4); i += 64) dst[i] = src[i];
            wave_lds_sync();
            if (live && lane == 0) {   // tile_pass_args: this pass's hours
                targs[wave].hours = d[t].a.hours + pass * K;
                targs[wave].nh = min(K, d[t].nsel - pass * K);
            }
            wave_lds_sync();
        }
        if (!live) continue;
        const float inv_w = 1.0f / (float)a.W;
        const int cells = a.W * a.H;
        const bool off32 = a.hourly_top_cap <= ((int64_t)1 << 30);
        const int64_t v = a.order[p - t_lo];
        const Recv r = receiver_of(a, v, inv_w);
        const int64_t s0 = uniform64(a.nb_off[v]), s1 = uniform64(a.nb_off[v + 1]);
        win.end = s1;
        float own, refl;
        receiver_parts<K, FMT>(a, a.hours, r, v, s0, s1, win, L.list_id[wave], L.list_w[wave], L.pend_id[wave], lane,
                               off32, inv_w, cells, own, refl);
        const float tt = wave_sum(own + refl);
        if (lane == 0) kahan_add(a.sum, a.comp, v, tt);
    }
}

/* Batched receiver pairs: each tile's own irradiance of the pass (k_own_pass over its voxels, with
   its own hour constants); block b belongs to the last tile whose first own block is <= b. */
template <int K>
__global__ __launch_bounds__(256) void k_own_pass_batch(const BatchHour* d, int ntiles, int pass) {
    __shared__ HourConst hrs[K];
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].own_block0 <= (int64_t)blockIdx.x) lo = mid;
        else hi = mid - 1;
    }
    HourArgs a;
    if (!tile_pass_args<K>(d[lo], pass, a)) return;   // the whole block leaves together
    stage_hours<K>(a, hrs);
    __syncthreads();
    const int64_t v = ((int64_t)blockIdx.x - d[lo].own_block0) * 256 + threadIdx.x;
    if (v < a.n_total) a.grp_own[v] = own_sum<K>(a, hrs, v, 1.0f / (float)a.W);
}

/* The tile owning pair slot g: the last tile whose first slot is <= g (wave-uniform). */
__device__ __forceinline__ int tile_of_slot(const BatchHour* d, int ntiles, int64_t g) {
    int lo = 0, hi = ntiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (d[mid].pslot0 <= g) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

/* Batched receiver pairs: k_reflect_pair over the pair slots of every tile (tile t's slot k = its
 * positions 2k and 2k + 1 in its processing order, so each tile's pairs are those of a context of
 * its own and its sums equal that context's bit for bit).  A wave takes a slot from its XCD's queue,
 * finds the tile (its pass arguments copied into the wave's LDS slot, as k_reflect_batch_queue
 * does) and walks the pair; tiles whose hours are done are passed over. */
template <int K, int WAVES>
__global__ __launch_bounds__(64 * WAVES) __attribute__((amdgpu_waves_per_eu(7)))
void k_reflect_batch_pair(const BatchHour* d, int ntiles, int pass, unsigned* queue, int64_t nslots, int64_t band,
                          int2* scratch) {
    static_assert(K >= 16, "four hours per lane");
    __shared__ PairLdsT<kPairLds> PL[WAVES];
    __shared__ HourArgs targs[WAVES];
    static_assert(sizeof(HourArgs) % 4 == 0, "copied as words");
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    const unsigned x = blockIdx.x % kXcds;
    const unsigned c = ((blockIdx.x / kXcds) * WAVES + (unsigned)wave) % kQueuesPerXcd;
    unsigned* counter = queue + (x * kQueuesPerXcd + c) * kQueueStride;
    int2* scr = scratch + ((int64_t)blockIdx.x * WAVES + wave) * kGroupCap;
    PairLdsT<kPairLds>& L = PL[wave];
    StreamWindow win{L.u.dec.stage, d[0].a.nb_bytes, 0, 0, 0, kGroupStageBytes};
    unsigned pending = atomicAdd(counter, lane == 0 ? 1u : 0u);   // every lane joins (see k_reflect_queue)
    int64_t s_lo = 0, s_hi = 0;
    bool live = false;   // the loaded tile has hours in this pass
    const HourArgs& a = targs[wave];
    for (int64_t turn = 0; turn <= nslots; ++turn) {
        const unsigned item = (unsigned)__builtin_amdgcn_readfirstlane((int)pending);
        const int64_t g = queue_position((uint64_t)item * kQueuesPerXcd + c, x, nslots, band);
        if (g >= nslots) break;
        pending = atomicAdd(counter, lane == 0 ? 1u : 0u);
        if (g < s_lo || g >= s_hi) {
            const int t = tile_of_slot(d, ntiles, g);
            s_lo = d[t].pslot0;
            s_hi = s_lo + (d[t].a.n_total + 1) / 2;
            live = pass * K < d[t].nsel;
            wave_lds_sync();   // the wave's reads of the previous tile's arguments are done
            const uint32_t* src = reinterpret_cast<const uint32_t*>(&d[t].a);
            uint32_t* dst = reinterpret_cast<uint32_t*>(&targs[wave]);
            for (int i = lane; i < (int)(sizeof(HourArgs) / 4); i += 64) dst[i] = src[i];
            wave_lds_sync();
            if (live && lane == 0) {   // tile_pass_args: this pass's hours
                targs[wave].hours = d[t].a.hours + pass * K;
                targs[wave].nh = min(K, d[t].nsel - pass * K);
            }
            wave_lds_sync();
        }
        if (!live) continue;
        const int64_t p0 = (g - s_lo) * 2;   // the pair's first position in the tile
        const bool two = p0 + 1 < a.n_total;
        const int64_t v0 = a.order[p0], v1 = two ? (int64_t)a.order[p0 + 1] : 0;
        pair_walk<K, kPairLds>(a, v0, v0, v1, v1, two, L, win, scr, kGroupCap, lane);
    }
}

template <int K>
void solar_top_k(const HourArgs& a, hipStream_t s) {
    const unsigned g = (unsigned)((a.n_top + kSolarBlock - 1) / kSolarBlock);
    hipLaunchKernelGGL(k_solar_top<K>, dim3(g), dim3(kSolarBlock), 0, s, a);
}

template <int K, bool DBG, int WAVES>
void reflect_kw(const HourArgs& a, unsigned fmt, hipStream_t s) {
    const int64_t n = a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total);
    int64_t blocks = (n + 64 * WAVES - 1) / (64 * WAVES);
    if (a.group > 1) blocks = (blocks + a.group - 1) / a.group * a.group;   // whole receiver groups
    /* round up to a multiple of the XCD count so the remap covers every block */
    const unsigned g = (unsigned)(((blocks + kXcds - 1) / kXcds) * kXcds);
    const dim3 blk(64 * WAVES);
    if (fmt == SGPU_FMT_CSR) hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_CSR, DBG, WAVES>), dim3(g), blk, 0, s, a);
    else hipLaunchKernelGGL((k_reflect<K, SGPU_FMT_BARLE, DBG, WAVES>), dim3(g), blk, 0, s, a);
}

template <int K, int WAVES>
void batch_queue_kw(const BatchHour* d, int ntiles, unsigned fmt, int pass, unsigned* queue, int64_t n, hipStream_t s);

template <int K, int WAVES>
void batch_pair_kw(const BatchHour* d, int ntiles, int pass, unsigned* queue, const BatchPairs& bp, hipStream_t s);

template <int K>
void batch_k(const BatchHour* d, int ntiles, int64_t top_blocks, int64_t blocks, unsigned fmt, int waves, int pass,
             unsigned* queue, int64_t n_all, const BatchPairs* pairs, hipStream_t s) {
    hipLaunchKernelGGL(k_solar_top_batch<K>, dim3((unsigned)top_blocks), dim3(kSolarBlock), 0, s, d, ntiles, pass);
    if constexpr (K >= 16) {
        if (queue && pairs && fmt == SGPU_FMT_BARLE) {
            if (waves == 2) batch_pair_kw<K, 2>(d, ntiles, pass, queue, *pairs, s);
            else batch_pair_kw<K, 4>(d, ntiles, pass, queue, *pairs, s);
            return;
        }
    }
    if (queue) {
        if (waves == 2) batch_queue_kw<K, 2>(d, ntiles, fmt, pass, queue, n_all, s);
        else batch_queue_kw<K, 4>(d, ntiles, fmt, pass, queue, n_all, s);
        return;
    }
    blocks = (blocks + kXcds - 1) / kXcds * kXcds;   // whole XCD sets for xcd_block
    if (fmt == SGPU_FMT_CSR) {
        if (waves == 2)
            hipLaunchKernelGGL((k_reflect_batch<K, SGPU_FMT_CSR, 2>), dim3((unsigned)blocks), dim3(128), 0, s, d,
                               ntiles, pass);
        else
            hipLaunchKernelGGL((k_reflect_batch<K, SGPU_FMT_CSR, 4>), dim3((unsigned)blocks), dim3(256), 0, s, d,
                               ntiles, pass);
    } else {
        if (waves == 2)
            hipLaunchKernelGGL((k_reflect_batch<K, SGPU_FMT_BARLE, 2>), dim3((unsigned)blocks), dim3(128), 0, s, d,
                               ntiles, pass);
        else
            hipLaunchKernelGGL((k_reflect_batch<K, SGPU_FMT_BARLE, 4>), dim3((unsigned)blocks), dim3(256), 0, s, d,
                               ntiles, pass);
    }
}

/* The persistent grid of the queued kernel: the workgroups every CU holds at once (waves take
   work from the queues, so more would only start after the work is gone), a multiple of the XCD
   count. */
struct GridShape {
    int cus, per_cu;
};
template <typename Kernel>
GridShape occupancy(Kernel kernel, int threads) {
    int dev = 0, cus = 0, per_cu = 0;
    SGPU_HIP(hipGetDevice(&dev));
    SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, 0) != hipSuccess || per_cu < 1)
        per_cu = 2;
    return {std::max(cus, 1), per_cu};
}
/* blocks_per_cu > 0 caps the resident workgroups per CU: fewer waves in flight keep the
   XCD's neighbour rows within its L2 (a knob of the autotuner) */
unsigned queue_grid(GridShape g, int blocks_per_cu) {
    const int per = blocks_per_cu > 0 ? std::min(blocks_per_cu, g.per_cu) : g.per_cu;
    const unsigned n = (unsigned)(g.cus * per);
    return (n + kXcds - 1) / kXcds * kXcds;
}

template <int K, int WAVES>
void reflect_queue_kw(const HourArgs& a, unsigned fmt, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(kXcds * kQueuesPerXcd), 0, s, a.queue);
    if (fmt == SGPU_FMT_CSR) {
        static const GridShape g = occupancy(k_reflect_queue<K, SGPU_FMT_CSR, WAVES>, 64 * WAVES);
        hipLaunchKernelGGL((k_reflect_queue<K, SGPU_FMT_CSR, WAVES>), dim3(queue_grid(g, a.queue_blocks_per_cu)),
                           dim3(64 * WAVES), 0, s, a);
    } else {
        static const GridShape g = occupancy(k_reflect_queue<K, SGPU_FMT_BARLE, WAVES>, 64 * WAVES);
        hipLaunchKernelGGL((k_reflect_queue<K, SGPU_FMT_BARLE, WAVES>), dim3(queue_grid(g, a.queue_blocks_per_cu)),
                           dim3(64 * WAVES), 0, s, a);
    }
}

template <int K, int WAVES>
void reflect_pair_kw(const HourArgs& a, hipStream_t s) {
    static const GridShape g = occupancy(k_reflect_pair<K, WAVES>, 64 * WAVES);
    const unsigned grid = queue_grid(g, a.queue_blocks_per_cu);
    if ((int64_t)grid * WAVES > a.grp_waves || a.grp_cap < 0 || a.grp_own == nullptr)
        fail(SGPU_ERR_STATE, "pair scratch holds %lld waves of %d entries; the grid has %lld waves",
             (long long)a.grp_waves, a.grp_cap, (long long)grid * WAVES);
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(kXcds * kQueuesPerXcd), 0, s, a.queue);
    hipLaunchKernelGGL((k_reflect_pair<K, WAVES>), dim3(grid), dim3(64 * WAVES), 0, s, a);
}

/* The receivers' own irradiance of the pass for the paired walk (launched with the pass's solar
   term, before the reflection). */
template <int K>
void own_pass_k(const HourArgs& a, hipStream_t s) {
    if constexpr (K >= 16) {
        const int64_t n = a.order ? a.n_total : a.n_sub;
        const unsigned blocks = (unsigned)std::min<int64_t>((n + 255) / 256, 8192);
        hipLaunchKernelGGL(k_own_pass<K>, dim3(blocks), dim3(256), 0, s, a, a.grp_own);
    }
}

/* Whether the reflection pass of these arguments walks receiver pairs: queued (the full scene in
   processing order, or a subset), ba-RLE lists, K >= 16 and the pair buffers present. */
bool walks_pairs(const HourArgs& a, int K, unsigned fmt) {
    const bool queued = a.queue != nullptr && a.dbg_hourly == nullptr &&
                        ((a.order != nullptr && a.subset == nullptr && a.n_sub == 0) ||
                         (a.order == nullptr && a.subset != nullptr && a.n_sub > 0));
    return queued && K >= 16 && fmt == SGPU_FMT_BARLE && a.grp_size == 2 && a.grp_scratch != nullptr &&
           a.grp_own != nullptr;
}

/* XCD bands of the queued batched reflection: 65536 positions (what the single-scene tuning picks
   on the 2048^2 scene: bands of 256 workgroups of 4 waves x 64) */
constexpr int64_t kBatchQueueBand = 65536;

template <int K, int WAVES>
void batch_queue_kw(const BatchHour* d, int ntiles, unsigned fmt, int pass, unsigned* queue, int64_t n, hipStream_t s) {
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(kXcds * kQueuesPerXcd), 0, s, queue);
    if (fmt == SGPU_FMT_CSR) {
        static const GridShape g = occupancy(k_reflect_batch_queue<K, SGPU_FMT_CSR, WAVES>, 64 * WAVES);
        hipLaunchKernelGGL((k_reflect_batch_queue<K, SGPU_FMT_CSR, WAVES>), dim3(queue_grid(g, 0)), dim3(64 * WAVES),
                           0, s, d, ntiles, pass, queue, n, kBatchQueueBand);
    } else {
        static const GridShape g = occupancy(k_reflect_batch_queue<K, SGPU_FMT_BARLE, WAVES>, 64 * WAVES);
        hipLaunchKernelGGL((k_reflect_batch_queue<K, SGPU_FMT_BARLE, WAVES>), dim3(queue_grid(g, 0)),
                           dim3(64 * WAVES), 0, s, d, ntiles, pass, queue, n, kBatchQueueBand);
    }
}

/* Batched receiver pairs: XCD bands of half the batched queue's band of positions, in pair slots. */
template <int K, int WAVES>
void batch_pair_kw(const BatchHour* d, int ntiles, int pass, unsigned* queue, const BatchPairs& bp, hipStream_t s) {
    static const GridShape g = occupancy(k_reflect_batch_pair<K, WAVES>, 64 * WAVES);
    const unsigned grid = queue_grid(g, 0);
    if ((int64_t)grid * WAVES > bp.waves)
        fail(SGPU_ERR_STATE, "pair scratch holds %lld waves; the grid has %lld", (long long)bp.waves,
             (long long)grid * WAVES);
    if (bp.own_blocks > 0)
        hipLaunchKernelGGL(k_own_pass_batch<K>, dim3((unsigned)bp.own_blocks), dim3(256), 0, s, d, ntiles, pass);
    hipLaunchKernelGGL(k_clear_queue, dim3(1), dim3(kXcds * kQueuesPerXcd), 0, s, queue);
    hipLaunchKernelGGL((k_reflect_batch_pair<K, WAVES>), dim3(grid), dim3(64 * WAVES), 0, s, d, ntiles, pass, queue,
                       bp.nslots, kBatchQueueBand / 2, bp.scratch);
}

template <int K>
void reflect_k(const HourArgs& a, unsigned fmt, hipStream_t s) {
    /* queued: the full scene in processing order, or a subset given by voxel ids */
    const bool queued = a.queue != nullptr && a.dbg_hourly == nullptr &&
                        ((a.order != nullptr && a.subset == nullptr && a.n_sub == 0) ||
                         (a.order == nullptr && a.subset != nullptr && a.n_sub > 0));
    if (queued) {
        if constexpr (K >= 16) {   // receiver pairs: ba-RLE lists, 4 hours per lane
            if (walks_pairs(a, K, fmt)) {
                if (a.waves == 2) reflect_pair_kw<K, 2>(a, s);
                else reflect_pair_kw<K, 4>(a, s);
                return;
            }
        }
        if (a.waves == 2) reflect_queue_kw<K, 2>(a, fmt, s);
        else reflect_queue_kw<K, 4>(a, fmt, s);
        return;
    }
    if (a.waves == 2) reflect_kw<K, false, 2>(a, fmt, s);
    else reflect_kw<K, false, 4>(a, fmt, s);
}

}  // namespace

bool supported_hours_per_pass(int K) {
    return K == 1 || K == 2 || K == 4 || K == 8 || K == 16 || K == 32 || K == 64 || K == 128 || K == 256;
}

/* The hourly irradiance rows of a pass (n_top x K floats) must fit the buffer they are written to
   (k_solar_top) and gathered from (k_reflect). */
void check_hourly_cap(const HourArgs& a, int K) {
    if (a.row_swz < 0 || a.row_swz > 30) fail(SGPU_ERR_ARG, "row swizzle shift %d out of range", a.row_swz);
    const int64_t rows = hourly_rows(a.n_top, a.row_swz);
    if (rows * (int64_t)K > a.hourly_top_cap)
        fail(SGPU_ERR_STATE, "hourly irradiance buffer holds %lld floats, the pass needs %lld",
             (long long)a.hourly_top_cap, (long long)(rows * (int64_t)K));
}

void launch_solar_top(const HourArgs& a, int K, hipStream_t s) {
    if (a.n_top <= 0) return;
    check_hourly_cap(a, K);
    switch (K) {
        case 1: solar_top_k<1>(a, s); break;
        case 2: solar_top_k<2>(a, s); break;
        case 4: solar_top_k<4>(a, s); break;
        case 8: solar_top_k<8>(a, s); break;
        case 16: solar_top_k<16>(a, s); break;
        case 32: solar_top_k<32>(a, s); break;
        case 64: solar_top_k<64>(a, s); break;
        case 128: solar_top_k<128>(a, s); break;
        case 256: solar_top_k<256>(a, s); break;
        default: fail(SGPU_ERR_ARG, "unsupported hours per pass %d", K);
    }
    /* the receivers' own irradiance, when the reflection walks pairs (the stream buffer marks ba-RLE
       lists) */
    if (walks_pairs(a, K, a.nb_bytes != nullptr ? SGPU_FMT_BARLE : SGPU_FMT_CSR)) {
        switch (K) {
            case 16: own_pass_k<16>(a, s); break;
            case 32: own_pass_k<32>(a, s); break;
            case 64: own_pass_k<64>(a, s); break;
            case 128: own_pass_k<128>(a, s); break;
            case 256: own_pass_k<256>(a, s); break;
            default: break;
        }
    }
    SGPU_HIP(hipGetLastError());
}

void launch_reflect_accum(const HourArgs& a, int K, unsigned fmt, hipStream_t s) {
    if ((a.n_sub > 0 ? a.n_sub : (a.subset ? 0 : a.n_total)) <= 0) return;
    if (fmt != SGPU_FMT_CSR && fmt != SGPU_FMT_BARLE) fail(SGPU_ERR_ARG, "format must be CSR or BARLE");
    check_hourly_cap(a, K);
    if (fmt == SGPU_FMT_BARLE && (reinterpret_cast<uintptr_t>(a.nb_bytes) & 15u) != 0)
        fail(SGPU_ERR_ARG, "ba-RLE stream must be 16-byte aligned");
    if (a.dbg_hourly != nullptr) {
        if (K != 1 || a.dbg_reflected == nullptr) fail(SGPU_ERR_ARG, "single-hour outputs need one hour per pass");
        reflect_kw<1, true, 4>(a, fmt, s);
        SGPU_HIP(hipGetLastError());
        return;
    }
    switch (K) {
        case 1: reflect_k<1>(a, fmt, s); break;
        case 2: reflect_k<2>(a, fmt, s); break;
        case 4: reflect_k<4>(a, fmt, s); break;
        case 8: reflect_k<8>(a, fmt, s); break;
        case 16: reflect_k<16>(a, fmt, s); break;
        case 32: reflect_k<32>(a, fmt, s); break;
        case 64: reflect_k<64>(a, fmt, s); break;
        case 128: reflect_k<128>(a, fmt, s); break;
        case 256: reflect_k<256>(a, fmt, s); break;
        default: fail(SGPU_ERR_ARG, "unsupported hours per pass %d", K);
    }
    SGPU_HIP(hipGetLastError());
}

void launch_hour_pass_batch(const BatchHour* d, int ntiles, int64_t top_blocks, int64_t blocks, int K,
                            unsigned fmt, int waves, int pass, unsigned* queue, int64_t n_all, const BatchPairs* pairs,
                            hipStream_t s) {
    if (ntiles <= 0 || blocks <= 0) return;
    if (fmt != SGPU_FMT_CSR && fmt != SGPU_FMT_BARLE) fail(SGPU_ERR_ARG, "format must be CSR or BARLE");
    if (waves != 2 && waves != 4) fail(SGPU_ERR_ARG, "waves per group must be 2 or 4");
    switch (K) {
        case 1: batch_k<1>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 2: batch_k<2>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 4: batch_k<4>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 8: batch_k<8>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 16: batch_k<16>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 32: batch_k<32>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 64: batch_k<64>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 128: batch_k<128>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        case 256: batch_k<256>(d, ntiles, top_blocks, blocks, fmt, waves, pass, queue, n_all, pairs, s); break;
        default: fail(SGPU_ERR_ARG, "unsupported hours per pass %d", K);
    }
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
