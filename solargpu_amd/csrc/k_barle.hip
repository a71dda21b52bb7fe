/* GPU ba-RLE encoder (format of the reference's barle_compression.h:7-44 and
 * encoder barle_compression.cpp:12-207), one wave per voxel:
 *   - the voxel's CSR ids are staged in LDS and bitonic-sorted descending by
 *     the whole wave (padded with zeros to a power of two);
 *   - duplicates are dropped with a wave-ordered compaction;
 *   - delta width classes are computed in parallel;
 *   - lane 0 runs the short serial part: run detection, the single-element
 *     absorption passes to a fixpoint, coalescing, the split at 255 and the
 *     block packing (3 code bytes, length bytes, payload).
 * Pass 1 writes per-voxel byte counts; pass 2 writes the bytes at scanned
 * offsets, so the stream is byte-identical to the reference encoder's. */
#include "sgpu_internal.h"

namespace sgpu {
namespace {

static_assert(kDirs <= 4096, "a voxel has at most one distinct hit per direction");

__device__ __forceinline__ unsigned width_class(unsigned x) {   // barle_compression.cpp:12-18
    return x <= 15u ? 0u : x <= 255u ? 1u : x <= 65535u ? 2u : x <= 16777215u ? 3u : 4u;
}

/* Run packed as w | start << 4 | len << 16 (start, len < 4096). */
__device__ __forceinline__ unsigned rw(unsigned r) { return r & 15u; }
__device__ __forceinline__ unsigned rstart(unsigned r) { return (r >> 4) & 4095u; }
__device__ __forceinline__ unsigned rlen(unsigned r) { return r >> 16; }
__device__ __forceinline__ unsigned mkrun(unsigned w, unsigned s, unsigned l) { return w | (s << 4) | (l << 16); }

/* Cursor over the runs after the split at 255 (barle_compression.cpp:29-44). */
struct SplitCursor {
    int ri;
    unsigned used;
};

__device__ __forceinline__ bool next_split(const unsigned* runs, int nruns, SplitCursor& c, unsigned& w,
                                          unsigned& start, unsigned& len) {
    if (c.ri >= nruns) return false;
    const unsigned r = runs[c.ri];
    w = rw(r);
    start = rstart(r) + c.used;
    const unsigned left = rlen(r) - c.used;
    len = left > 255u ? 255u : left;
    c.used += len;
    if (c.used == rlen(r)) { ++c.ri; c.used = 0; }
    return true;
}

/* Descending bitonic sort of keys[0, P) (P a power of two <= MAXP) by one wave. */
__device__ __forceinline__ void wave_bitonic_desc(unsigned* keys, int P, int lane) {
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = lane; i < P; i += 64) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const unsigned x = keys[i], y = keys[ixj];
                    const bool desc = (i & k) == 0;   // this k-block ends up descending
                    if (desc ? (x < y) : (x > y)) {
                        keys[i] = y;
                        keys[ixj] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
}

/* MAXP: the list capacity of this instantiation (a power of two); lists with n <= SKIP are
   left to the smaller instantiation, lists with n > MAXP never reach this one. */
template <bool WRITE, int MAXP, int SKIP>
__global__ __launch_bounds__(64) void k_barle(BarleArgs a) {
    __shared__ unsigned keys[MAXP];   // ids, then the run list
    __shared__ unsigned vals[MAXP];   // sorted-descending unique ids
    __shared__ unsigned char wcl[MAXP];

    const int64_t v = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t b = a.csr_off[v];
    const int n = (int)(a.csr_off[v + 1] - b);
    if (n <= SKIP || n > MAXP) return;
    if (n <= 0) {
        if (!WRITE && lane == 0) a.sizes[v] = 0;
        return;
    }
    int P = 64;
    while (P < n) P <<= 1;
    for (int i = lane; i < P; i += 64) keys[i] = (i < n) ? (unsigned)a.csr_ids[b + i] : 0u;   // ids >= 1
    __syncthreads();
    wave_bitonic_desc(keys, P, lane);
    /* Ordered compaction of the first of each equal group, chunks of 64 (padding zeros last). */
    int m = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool keep = (i < n) && (i == 0 || keys[i] != keys[i - 1]);
        const unsigned long long mask = __ballot(keep);
        const int pos = m + __popcll(mask & ((1ull << lane) - 1ull));
        if (keep) vals[pos] = keys[i];
        m += __popcll(mask);
    }
    __syncthreads();
    for (int i = lane; i < m; i += 64) {
        const unsigned x = vals[i];
        wcl[i] = (unsigned char)width_class(i == 0 ? x : vals[i - 1] - x);
    }
    __syncthreads();
    if (lane != 0) return;

    unsigned* runs = keys;   // sorted ids no longer needed
    /* detect_runs (barle_compression.cpp:90-105) */
    int nr = 0;
    for (int i = 0; i < m;) {
        int j = i + 1;
        while (j < m && wcl[j] == wcl[i]) ++j;
        runs[nr++] = mkrun(wcl[i], (unsigned)i, (unsigned)(j - i));
        i = j;
    }
    /* optimize_runs (barle_compression.cpp:112-152), in place */
    if (nr >= 2) {
        bool single = false;
        for (;;) {
            bool merged = false;
            int o = 0, i = 0;
            while (i + 1 < nr) {
                const unsigned r0 = runs[i], r1 = runs[i + 1];
                if (rlen(r1) == 1u && rw(r1) + 1u == rw(r0)) {
                    runs[o++] = mkrun(rw(r0), rstart(r0), rlen(r0) + 1u);
                    merged = true;
                    i += 2;
                } else {
                    runs[o++] = r0;
                    i += 1;
                }
            }
            if (i < nr) runs[o++] = runs[i];
            if (o == 1) { nr = 1; single = true; break; }
            if (!merged) break;
            nr = o;
        }
        if (!single) {
            int o = 0;
            for (int i = 0; i < nr; ++i) {
                if (o > 0 && rw(runs[o - 1]) == rw(runs[i]))
                    runs[o - 1] = mkrun(rw(runs[o - 1]), rstart(runs[o - 1]), rlen(runs[o - 1]) + rlen(runs[i]));
                else
                    runs[o++] = runs[i];
            }
            nr = o;
        }
    }

    if (!WRITE) {
        int64_t size = 0;
        int nsplit = 0;
        SplitCursor c{0, 0};
        unsigned w, s, l;
        while (next_split(runs, nr, c, w, s, l)) {
            ++nsplit;
            size += 1 + (w == 0 ? (l + 1) / 2 : l * w);
        }
        size += 3 * ((nsplit + 7) / 8);
        a.sizes[v] = (int32_t)size;
        return;
    }

    uint8_t* out = a.out + a.out_off[v];
    SplitCursor c{0, 0};
    for (;;) {
        unsigned ws[8], ss[8], ls[8];
        int cnt = 0;
        while (cnt < 8 && next_split(runs, nr, c, ws[cnt], ss[cnt], ls[cnt])) ++cnt;
        if (cnt == 0) break;
        unsigned bits = 0;
        for (int j = 0; j < 8; ++j) bits |= ((j < cnt ? ws[j] : 7u) & 7u) << ((7 - j) * 3);
        *out++ = (uint8_t)(bits >> 16);
        *out++ = (uint8_t)(bits >> 8);
        *out++ = (uint8_t)bits;
        for (int j = 0; j < cnt; ++j) *out++ = (uint8_t)ls[j];
        for (int j = 0; j < cnt; ++j) {
            const unsigned st = ss[j], ln = ls[j];
            if (ws[j] == 0) {
                for (unsigned k = 0; k < ln; k += 2) {
                    const unsigned i0 = st + k;
                    const unsigned hi = (i0 == 0) ? vals[0] : vals[i0 - 1] - vals[i0];
                    const unsigned lo = (k + 1 < ln) ? vals[i0] - vals[i0 + 1] : 0u;
                    *out++ = (uint8_t)((hi << 4) | (lo & 0x0Fu));
                }
            } else {
                for (unsigned k = 0; k < ln; ++k) {
                    const unsigned i0 = st + k;
                    const unsigned x = (i0 == 0) ? vals[0] : vals[i0 - 1] - vals[i0];
                    for (int q = (int)ws[j] - 1; q >= 0; --q) *out++ = (uint8_t)(x >> (8 * q));
                }
            }
        }
    }
}

}  // namespace

void launch_barle(const BarleArgs& a, bool write_pass, hipStream_t s) {
    if (a.n <= 0) return;
    /* most lists are short: a small-LDS instantiation for them (high occupancy), then the
       full-size one for the rest */
    const dim3 g((unsigned)a.n), blk(64);
    if (write_pass) {
        hipLaunchKernelGGL((k_barle<true, 512, -1>), g, blk, 0, s, a);
        hipLaunchKernelGGL((k_barle<true, 4096, 512>), g, blk, 0, s, a);
    } else {
        hipLaunchKernelGGL((k_barle<false, 512, -1>), g, blk, 0, s, a);
        hipLaunchKernelGGL((k_barle<false, 4096, 512>), g, blk, 0, s, a);
    }
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
