/* GPU ba-RLE encoder (format of the reference's barle_compression.h:7-44 and
 * encoder barle_compression.cpp:12-207), one wave per voxel:
 *   - the voxel's CSR ids are staged in LDS and rank-sorted descending by
 *     the whole wave (ties by index, so duplicates end up adjacent);
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

constexpr int kMaxList = kDirs;   // a voxel has at most one distinct hit per direction

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

template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle(BarleArgs a) {
    __shared__ unsigned vals[kMaxList];   // sorted-descending unique ids
    __shared__ unsigned tmp[kMaxList];    // unsorted ids, then the run list
    __shared__ unsigned char wcl[kMaxList];

    const int64_t v = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t b = a.csr_off[v];
    const int n = (int)(a.csr_off[v + 1] - b);
    if (n <= 0) {
        if (!WRITE && lane == 0) a.sizes[v] = 0;
        return;
    }
    for (int i = lane; i < n; i += 64) tmp[i] = (unsigned)a.csr_ids[b + i];
    __syncthreads();
    for (int i = lane; i < n; i += 64) {
        const unsigned x = tmp[i];
        int r = 0;
        for (int j = 0; j < n; ++j) {
            const unsigned y = tmp[j];
            r += (y > x) || (y == x && j < i);
        }
        vals[r] = x;
    }
    __syncthreads();
    /* Ordered compaction of the first of each equal group, chunks of 64. */
    int m = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool keep = (i < n) && (i == 0 || vals[i] != vals[i - 1]);
        const unsigned x = (i < n) ? vals[i] : 0u;
        const unsigned long long mask = __ballot(keep);
        const int pos = m + __popcll(mask & ((1ull << lane) - 1ull));
        __syncthreads();
        if (keep) tmp[pos] = x;
        m += __popcll(mask);
        __syncthreads();
    }
    for (int i = lane; i < m; i += 64) {
        const unsigned x = tmp[i];
        vals[i] = x;
        wcl[i] = (unsigned char)width_class(i == 0 ? x : tmp[i - 1] - x);
    }
    __syncthreads();
    if (lane != 0) return;

    unsigned* runs = tmp;   // ids no longer needed
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
    if (write_pass) hipLaunchKernelGGL(k_barle<true>, dim3((unsigned)a.n), dim3(64), 0, s, a);
    else hipLaunchKernelGGL(k_barle<false>, dim3((unsigned)a.n), dim3(64), 0, s, a);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
