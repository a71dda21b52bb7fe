/* GPU ba-RLE encoder (format of the reference's barle_compression.h:7-44 and
 * encoder barle_compression.cpp:12-207), one wave per voxel:
 *   - the voxel's CSR ids are staged in LDS and bitonic-sorted descending by
 *     the whole wave (padded with zeros to a power of two);
 *   - duplicates are dropped with a wave-ordered compaction (ballot);
 *   - delta width classes and run starts (width changes) in parallel;
 *   - lane 0 runs the short per-run part: the single-element absorption
 *     passes to a fixpoint and the coalescing (optimize_runs); the size pass
 *     totals the stream from the runs, the write pass splits them at 255 in
 *     place;
 *   - the write pass walks the blocks in order (scalar byte position) and all
 *     lanes write each block's header (3 code bytes, length bytes) and payloads.
 * Pass 1 writes per-voxel byte counts; pass 2 writes the bytes at scanned
 * offsets, so the stream is byte-identical to the reference encoder's.
 * Lists of <= 512 ids (most) take a small-LDS instantiation over all voxels;
 * longer ones are queued by pass 1 and run by a fixed grid of the full-size one. */
#include "sgpu_internal.h"

namespace sgpu {
namespace {

static_assert(kDirs <= 4096, "a voxel has at most one distinct hit per direction");
constexpr int kShort = 512;
constexpr int kLongGrid = 4096;

__device__ __forceinline__ unsigned width_class(unsigned x) {   // barle_compression.cpp:12-18
    return x <= 15u ? 0u : x <= 255u ? 1u : x <= 65535u ? 2u : x <= 16777215u ? 3u : 4u;
}

/* Run packed as w | start << 4 | len << 16 (start < 4096, len < 65536). */
__device__ __forceinline__ unsigned rw(unsigned r) { return r & 15u; }
__device__ __forceinline__ unsigned rstart(unsigned r) { return (r >> 4) & 4095u; }
__device__ __forceinline__ unsigned rlen(unsigned r) { return r >> 16; }
__device__ __forceinline__ unsigned mkrun(unsigned w, unsigned s, unsigned l) { return w | (s << 4) | (l << 16); }

__device__ __forceinline__ int lanes_below(unsigned long long m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* Descending bitonic sort of keys[0, P) (P a power of two) by one wave. */
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

/* optimize_runs (barle_compression.cpp:112-152) in place: a single-element run one width
   class below its left neighbour joins it (skipping past the pair) until nothing merges or one
   run is left; then equal-width neighbours coalesce.  Returns the new run count. */
__device__ int optimize_runs(unsigned* runs, int nr) {
    if (nr < 2) return nr;
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
        if (o == 1) return 1;
        if (!merged) break;
        nr = o;
    }
    int o = 0;
    for (int i = 0; i < nr; ++i) {
        if (o > 0 && rw(runs[o - 1]) == rw(runs[i]))
            runs[o - 1] = mkrun(rw(runs[o - 1]), rstart(runs[o - 1]), rlen(runs[o - 1]) + rlen(runs[i]));
        else
            runs[o++] = runs[i];
    }
    return o;
}

template <int MAXP>
struct EncLds {
    unsigned keys[MAXP];          // ids, then run starts, runs and (write pass) splits
    unsigned vals[MAXP];          // sorted-descending unique ids
    unsigned char wcl[MAXP];      // width class of each delta
    int ns;
};

/* Payload bytes of a split of len deltas of width class w. */
__device__ __forceinline__ int payload_bytes(unsigned w, unsigned len) {
    return (w == 0u) ? (int)((len + 1u) >> 1) : (int)(len * w);
}

__device__ __forceinline__ unsigned delta_at(const unsigned* vals, unsigned i) {
    return (i == 0) ? vals[0] : vals[i - 1] - vals[i];   // the first value is absolute
}

/* The ba-RLE stream of voxel v's list (n ids at csr_ids + b): pass 1 its size, pass 2 its bytes. */
template <bool WRITE, int MAXP>
__device__ void encode_voxel(const BarleArgs& a, EncLds<MAXP>& s, int64_t v, int64_t b, int n, int lane) {
    int P = 64;
    while (P < n) P <<= 1;
    for (int i = lane; i < P; i += 64) s.keys[i] = (i < n) ? (unsigned)a.csr_ids[b + i] : 0u;   // ids >= 1
    __syncthreads();
    wave_bitonic_desc(s.keys, P, lane);
    /* first of each equal group, in order (padding zeros sort last) */
    int m = 0;
    for (int base = 0; base < n; base += 64) {
        const int i = base + lane;
        const bool keep = (i < n) && (i == 0 || s.keys[i] != s.keys[i - 1]);
        const unsigned long long mask = __ballot(keep);
        if (keep) s.vals[m + lanes_below(mask)] = s.keys[i];
        m += __popcll(mask);
    }
    __syncthreads();
    for (int i = lane; i < m; i += 64) s.wcl[i] = (unsigned char)width_class(delta_at(s.vals, i));
    __syncthreads();
    /* detect_runs (barle_compression.cpp:90-105): a run starts where the width class changes */
    int nr = 0;
    for (int base = 0; base < m; base += 64) {
        const int i = base + lane;
        const bool st = i < m && (i == 0 || s.wcl[i] != s.wcl[i - 1]);
        const unsigned long long mask = __ballot(st);
        if (st) s.keys[nr + lanes_below(mask)] = (unsigned)i;
        nr += __popcll(mask);
    }
    __syncthreads();
    for (int base = 0; base < nr; base += 64) {
        const int r = base + lane;
        unsigned s0 = 0, s1 = 0;
        if (r < nr) {
            s0 = s.keys[r];
            s1 = (r + 1 < nr) ? s.keys[r + 1] : (unsigned)m;
        }
        __syncthreads();
        if (r < nr) s.keys[r] = mkrun(s.wcl[s0], s0, s1 - s0);
        __syncthreads();
    }
    if (!WRITE) {
        if (lane == 0) {
            /* size: per split one length byte and its payload, per block of 8 splits 3 code
               bytes; a run of len splits into ceil(len / 255) pieces of 255 and the rest */
            nr = optimize_runs(s.keys, nr);
            int ns = 0, size = 0;
            for (int r = 0; r < nr; ++r) {
                const unsigned w = rw(s.keys[r]), len = rlen(s.keys[r]);
                const unsigned full = len / 255u, rest = len % 255u;
                ns += (int)(full + (rest ? 1u : 0u));
                size += (int)(full + (rest ? 1u : 0u)) + (int)full * payload_bytes(w, 255u) + payload_bytes(w, rest);
            }
            a.sizes[v] = size + 3 * ((ns + 7) / 8);
        }
        return;
    }
    if (lane == 0) {
        nr = optimize_runs(s.keys, nr);
        /* split at 255 (barle_compression.cpp:29-44) in place, from the last run backwards
           (run r's pieces land at or after index r) */
        int ns = 0;
        for (int r = 0; r < nr; ++r) ns += (int)((rlen(s.keys[r]) + 254u) / 255u);
        int at = ns;
        for (int r = nr - 1; r >= 0; --r) {
            const unsigned w = rw(s.keys[r]), st = rstart(s.keys[r]), len = rlen(s.keys[r]);
            const int pieces = (int)((len + 254u) / 255u);
            at -= pieces;
            for (int k = 0; k < pieces; ++k) {
                const unsigned u = 255u * (unsigned)k;
                s.keys[at + k] = mkrun(w, st + u, min(255u, len - u));
            }
        }
        s.ns = ns;
    }
    __syncthreads();
    const int ns = s.ns;
    uint8_t* out = a.out + a.out_off[v];
    int pos = 0;   // wave-uniform byte position
    for (int t0 = 0; t0 < ns; t0 += 8) {
        const int cnt = min(8, ns - t0);
        if (lane < 3 + cnt) {   // block header: 3 code bytes (7 = unused slot), length bytes
            uint8_t byte;
            if (lane < 3) {
                unsigned bits = 0;
                for (int j = 0; j < 8; ++j) bits |= ((j < cnt ? rw(s.keys[t0 + j]) : 7u) & 7u) << ((7 - j) * 3);
                byte = (uint8_t)(bits >> (16 - 8 * lane));
            } else {
                byte = (uint8_t)rlen(s.keys[t0 + lane - 3]);
            }
            out[pos + lane] = byte;
        }
        pos += 3 + cnt;
        for (int t = t0; t < t0 + cnt; ++t) {   // payloads, the lanes over one split's bytes
            const unsigned sp = s.keys[t], w = rw(sp), st = rstart(sp), len = rlen(sp);
            uint8_t* dst = out + pos;
            if (w == 0u) {   // two 4-bit deltas per byte, high nibble first, odd tail low = 0
                for (unsigned k = lane; 2u * k < len; k += 64u) {
                    const unsigned i0 = st + 2u * k;
                    const unsigned hi = delta_at(s.vals, i0);
                    const unsigned lo = (2u * k + 1u < len) ? delta_at(s.vals, i0 + 1u) : 0u;
                    dst[k] = (uint8_t)((hi << 4) | (lo & 0x0Fu));
                }
            } else {         // w bytes big-endian per delta
                for (unsigned k = lane; k < len; k += 64u) {
                    const unsigned x = delta_at(s.vals, st + k);
                    for (unsigned q = 0; q < w; ++q) dst[k * w + q] = (uint8_t)(x >> (8u * (w - 1u - q)));
                }
            }
            pos += payload_bytes(w, len);
        }
    }
}

/* Lists of <= kShort ids; pass 1 queues the longer ones. */
template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle_short(BarleArgs a) {
    __shared__ EncLds<kShort> s;
    const int64_t v = blockIdx.x;
    const int64_t b = a.csr_off[v];
    const int n = (int)(a.csr_off[v + 1] - b);
    if (n > kShort) {
        if (!WRITE && threadIdx.x == 0) a.long_ids[atomicAdd(a.long_count, 1u)] = (int32_t)v;
        return;
    }
    if (n <= 0) {
        if (!WRITE && threadIdx.x == 0) a.sizes[v] = 0;
        return;
    }
    encode_voxel<WRITE, kShort>(a, s, v, b, n, threadIdx.x);
}

/* The queued long lists, a fixed grid striding over the queue. */
template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle_long(BarleArgs a) {
    __shared__ EncLds<4096> s;
    const unsigned count = *a.long_count;
    for (unsigned q = blockIdx.x; q < count; q += gridDim.x) {
        const int64_t v = a.long_ids[q];
        const int64_t b = a.csr_off[v];
        encode_voxel<WRITE, 4096>(a, s, v, b, (int)(a.csr_off[v + 1] - b), threadIdx.x);
        __syncthreads();   // LDS reused by the next voxel
    }
}

}  // namespace

void launch_barle(const BarleArgs& a, bool write_pass, hipStream_t s) {
    if (a.n <= 0) return;
    if (!a.long_ids || !a.long_count) fail(SGPU_ERR_ARG, "ba-RLE encoder needs its long-list queue");
    const dim3 g((unsigned)a.n), blk(64);
    if (write_pass) {
        hipLaunchKernelGGL(k_barle_short<true>, g, blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_long<true>, dim3(kLongGrid), blk, 0, s, a);
    } else {
        SGPU_HIP(hipMemsetAsync(a.long_count, 0, sizeof(unsigned), s));
        hipLaunchKernelGGL(k_barle_short<false>, g, blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_long<false>, dim3(kLongGrid), blk, 0, s, a);
    }
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
