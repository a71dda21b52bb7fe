/* GPU ba-RLE encoder (format of the reference's barle_compression.h:7-44 and
 * encoder barle_compression.cpp:12-207), one wave per voxel, every step
 * wave-parallel:
 *   - sort: lists of <= 1024 ids are bitonic-sorted descending in registers
 *     (E = 1..16 ids per lane, cross-lane stages by xor shuffles, stages of
 *     stride >= 64 inside each lane); longer lists in LDS;
 *   - duplicates are dropped with a wave-ordered compaction (ballot);
 *   - delta width classes and run starts (width changes) in parallel;
 *   - optimize_runs: each absorption sweep is solved with 64-bit ballot
 *     arithmetic (in a maximal chain of candidate pairs the greedy left-to-right
 *     sweep merges every other pair from the chain's start), repeated to the
 *     fixpoint, then equal-width neighbours coalesce by a boundary compaction;
 *   - pass 1 sums the stream size over the runs; pass 2 splits the runs at 255
 *     (prefix of piece counts), places every split by a prefix of payload
 *     bytes (block b of 8 splits starts at 11 b + payload bytes before it), and
 *     writes headers per split and payload bytes per delta.
 * Pass 1 writes per-voxel byte counts; pass 2 writes the bytes at scanned
 * offsets, so the stream is byte-identical to the reference encoder's.
 * Lists of <= 512 ids (most) take a small-LDS instantiation over all voxels;
 * pass 1 queues those of 513..1024 ids (registers, 13 KB of LDS) and the longer
 * ones (LDS sort, 52 KB) for two fixed grids striding over the queues. */
#include "sgpu_internal.h"

namespace sgpu {
namespace {

static_assert(kDirs <= 4096, "a voxel has at most one distinct hit per direction");
constexpr int kShort = 512;
constexpr int kMedium = 1024;
constexpr int kLongGrid = 4096;
constexpr int kHugeGrid = 1024;

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

constexpr unsigned long long kEven = 0x5555555555555555ull;
constexpr unsigned long long kOdd = 0xAAAAAAAAAAAAAAAAull;

__device__ __forceinline__ unsigned readlane(unsigned x, int l) {
    return (unsigned)__builtin_amdgcn_readlane((int)x, l);
}

/* Inclusive wave64 prefix sum. */
__device__ __forceinline__ int wave_incl_sum(int x, int lane) {
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ int wave_sum(int x) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
    return x;
}

/* Descending bitonic sort of the wave's 64 E ids, element i = r * 64 + lane in x[r]. */
template <int E, int LOGP>
__device__ __forceinline__ void reg_bitonic_desc(unsigned (&x)[E], int lane) {
#pragma unroll
    for (int lk = 1; lk <= LOGP; ++lk) {
        const int k = 1 << lk;
#pragma unroll
        for (int lj = lk - 1; lj >= 0; --lj) {
            const int j = 1 << lj;
            if (j >= 64) {   // partner in the same lane
                const int jr = j >> 6;
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    if (r & jr) continue;
                    const bool desc = ((r * 64) & k) == 0;
                    const unsigned p = x[r], q = x[r | jr];
                    x[r] = desc ? max(p, q) : min(p, q);
                    x[r | jr] = desc ? min(p, q) : max(p, q);
                }
            } else {         // partner lane ^ j; the lower index keeps the max in a descending block
                const bool lower = (lane & j) == 0;
#pragma unroll
                for (int r = 0; r < E; ++r) {
                    const unsigned y = (unsigned)__shfl_xor((int)x[r], j, 64);
                    const bool desc = ((r * 64 + lane) & k) == 0;
                    x[r] = (lower == desc) ? max(x[r], y) : min(x[r], y);
                }
            }
        }
    }
}

/* Sorted-descending unique ids of a list of n <= 64 E ids into vals; returns their count. */
template <int E, int LOGP>
__device__ __forceinline__ int sort_unique_regs(const int32_t* ids, int n, unsigned* vals, int lane) {
    unsigned x[E];
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const int i = r * 64 + lane;
        x[r] = (i < n) ? (unsigned)ids[i] : 0u;   // ids >= 1; padding zeros sort last
    }
    reg_bitonic_desc<E, LOGP>(x, lane);
    int m = 0;
#pragma unroll
    for (int r = 0; r < E; ++r) {
        const unsigned up = (unsigned)__shfl_up((int)x[r], 1, 64);
        const unsigned last = (r > 0) ? readlane(x[r > 0 ? r - 1 : 0], 63) : 0u;
        const unsigned prev = (lane == 0) ? last : up;
        const int i = r * 64 + lane;
        const bool keep = (i < n) && (i == 0 || x[r] != prev);
        const unsigned long long mask = __ballot(keep);
        if (keep) vals[m + lanes_below(mask)] = x[r];
        m += __popcll(mask);
    }
    return m;
}

/* Descending bitonic sort of keys[0, P) (P a power of two) in LDS by one wave. */
__device__ __forceinline__ void lds_bitonic_desc(unsigned* keys, int P, int lane) {
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

/* optimize_runs (barle_compression.cpp:112-152) on runs[0, nr) covering m deltas, in place,
   wave-parallel.  A sweep merges run i with a single-element run i + 1 one width class below
   it, skipping past the pair; over 64 runs the candidate mask C (with position 0 blocked when
   the previous chunk's last run merged) merges at every other bit of each maximal chain of
   ones, counted from the chain's start.  Sweeps repeat until nothing merges or one run is
   left; then equal-width neighbours coalesce.  Returns the new run count. */
__device__ int optimize_runs_wave(unsigned* runs, int nr, int m, int lane) {
    if (nr < 2) return nr;
    for (;;) {
        int o = 0;
        unsigned long long carry = 0, any = 0;
        for (int base = 0; base < nr; base += 64) {
            const int i = base + lane;
            const unsigned r0 = (i < nr) ? runs[i] : 0u;
            const unsigned r1 = (i + 1 < nr) ? runs[i + 1] : 0u;
            const bool cand = (i + 1 < nr) && rlen(r1) == 1u && rw(r1) + 1u == rw(r0);
            const unsigned long long C = __ballot(cand) & ~carry;
            const unsigned long long S = C & ~(C << 1);                 // chain starts
            const unsigned long long even_chain = C & ~(C + (S & kEven)); // chains starting at even bits
            const unsigned long long M = (even_chain & kEven) | (C & ~even_chain & kOdd);
            const unsigned long long emit = __ballot(i < nr) & ~((M << 1) | carry);
            __syncthreads();   // the chunk's reads before its writes (outputs land at <= i)
            if ((emit >> lane) & 1ull)
                runs[o + lanes_below(emit)] = ((M >> lane) & 1ull) ? mkrun(rw(r0), rstart(r0), rlen(r0) + 1u) : r0;
            o += __popcll(emit);
            carry = M >> 63;
            any |= M;
            __syncthreads();
        }
        nr = o;
        if (nr == 1) return 1;
        if (!any) break;
    }
    /* coalesce: keep the first run of every equal-width group, then lengths from the starts */
    int o = 0;
    unsigned prev_w = 0xFFFFFFFFu;
    for (int base = 0; base < nr; base += 64) {
        const int i = base + lane;
        const unsigned r0 = (i < nr) ? runs[i] : 0u;
        const unsigned up = (unsigned)__shfl_up((int)rw(r0), 1, 64);
        const unsigned pw = (lane == 0) ? prev_w : up;
        const bool first = (i < nr) && rw(r0) != pw;
        const unsigned long long B = __ballot(first);
        prev_w = readlane(rw(r0), 63);
        __syncthreads();
        if (first) runs[o + lanes_below(B)] = r0;
        o += __popcll(B);
        __syncthreads();
    }
    for (int base = 0; base < o; base += 64) {
        const int j = base + lane;
        const unsigned r0 = (j < o) ? runs[j] : 0u;
        const unsigned end = (j + 1 < o) ? rstart(runs[j + 1]) : (unsigned)m;
        __syncthreads();
        if (j < o) runs[j] = mkrun(rw(r0), rstart(r0), end - rstart(r0));
        __syncthreads();
    }
    return o;
}

template <int MAXP>
struct EncLds {
    unsigned keys[MAXP];          // LDS sort buffer (long lists), runs, then payload positions
    unsigned vals[MAXP];          // sorted-descending unique ids
    unsigned spl[MAXP];           // splits (write pass)
    unsigned char wcl[MAXP];      // width class of each delta; 0x80 marks a split start (write pass)
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
    int m;
    if constexpr (MAXP <= kShort) {
        const int32_t* ids = a.csr_ids + b;
        if (n <= 64) m = sort_unique_regs<1, 6>(ids, n, s.vals, lane);
        else if (n <= 128) m = sort_unique_regs<2, 7>(ids, n, s.vals, lane);
        else if (n <= 256) m = sort_unique_regs<4, 8>(ids, n, s.vals, lane);
        else m = sort_unique_regs<8, 9>(ids, n, s.vals, lane);
    } else if constexpr (MAXP <= kMedium) {
        m = sort_unique_regs<16, 10>(a.csr_ids + b, n, s.vals, lane);
    } else {
        int P = 64;
        while (P < n) P <<= 1;
        for (int i = lane; i < P; i += 64) s.keys[i] = (i < n) ? (unsigned)a.csr_ids[b + i] : 0u;   // ids >= 1
        __syncthreads();
        lds_bitonic_desc(s.keys, P, lane);
        m = 0;   // first of each equal group, in order (padding zeros sort last)
        for (int base = 0; base < n; base += 64) {
            const int i = base + lane;
            const bool keep = (i < n) && (i == 0 || s.keys[i] != s.keys[i - 1]);
            const unsigned long long mask = __ballot(keep);
            if (keep) s.vals[m + lanes_below(mask)] = s.keys[i];
            m += __popcll(mask);
        }
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
    nr = optimize_runs_wave(s.keys, nr, m, lane);
    if (!WRITE || a.slots) {
        /* per split one length byte and its payload, per block of 8 splits 3 code bytes; a run
           of len splits into ceil(len / 255) pieces of 255 and the rest */
        int bytes = 0, pieces = 0;
        for (int r = lane; r < nr; r += 64) {
            const unsigned w = rw(s.keys[r]), len = rlen(s.keys[r]);
            const unsigned full = len / 255u, rest = len % 255u;
            const int pc = (int)(full + (rest ? 1u : 0u));
            pieces += pc;
            bytes += pc + (int)full * payload_bytes(w, 255u) + payload_bytes(w, rest);
        }
        bytes = wave_sum(bytes);
        pieces = wave_sum(pieces);
        if (lane == 0) a.sizes[v] = bytes + 3 * ((pieces + 7) / 8);
        if (!WRITE) return;
    }
    /* split at 255 (barle_compression.cpp:29-44): run r's pieces at the prefix of piece counts */
    int ns = 0;
    for (int base = 0; base < nr; base += 64) {
        const int r = base + lane;
        const unsigned run = (r < nr) ? s.keys[r] : 0u;
        const unsigned w = rw(run), st = rstart(run), len = rlen(run);
        const int pc = (int)((len + 254u) / 255u);
        const int incl = wave_incl_sum(pc, lane);
        for (int k = 0; k < pc; ++k) {
            const unsigned u = 255u * (unsigned)k;
            s.spl[ns + incl - pc + k] = mkrun(w, st + u, min(255u, len - u));
        }
        ns += (int)readlane((unsigned)incl, 63);
    }
    __syncthreads();
    /* headers and payload positions: block b (splits 8b .. 8b+7) starts at 11 b + the payload
       bytes of the splits before it; its 3 code bytes, then cnt length bytes, then payloads */
    uint8_t* out = a.slots ? a.out + 6 * b + 8 * v : a.out + a.out_off[v];
    int pay = 0;   // payload bytes before this chunk of splits
    for (int base = 0; base < ns; base += 64) {
        const int t = base + lane;
        const bool valid = t < ns;
        const unsigned sp = valid ? s.spl[t] : 0u;
        const unsigned w = rw(sp), len = rlen(sp);
        const int pb = valid ? payload_bytes(w, len) : 0;
        const int incl = wave_incl_sum(pb, lane);
        const int before = pay + incl - pb;                  // payload bytes of splits < t
        const int at_block = __shfl(before, lane & ~7, 64);  // ... < the block's first split
        const int blk = t >> 3;
        const int cnt = min(8, ns - (blk << 3));
        if (valid) {
            const int hdr = 11 * blk + at_block;
            s.keys[t] = (unsigned)(11 * blk + 3 + cnt + before);
            out[hdr + 3 + (t & 7)] = (uint8_t)len;
            if ((t & 7) == 0) {
                unsigned bits = 0;
                for (int j = 0; j < 8; ++j) bits |= ((j < cnt ? rw(s.spl[t + j]) : 7u) & 7u) << ((7 - j) * 3);
                out[hdr] = (uint8_t)(bits >> 16);
                out[hdr + 1] = (uint8_t)(bits >> 8);
                out[hdr + 2] = (uint8_t)bits;
            }
            s.wcl[rstart(sp)] = 0x80;
        }
        pay += (int)readlane((unsigned)incl, 63);
    }
    __syncthreads();
    /* payloads, one lane per delta: its split is the count of split starts up to it, minus one */
    int starts = 0;
    for (int base = 0; base < m; base += 64) {
        const int i = base + lane;
        const bool valid = i < m;
        const unsigned long long F = __ballot(valid && (s.wcl[i] & 0x80));
        if (valid) {
            const int t = starts + lanes_below(F) + (int)((F >> lane) & 1ull) - 1;
            const unsigned sp = s.spl[t], w = rw(sp), k = (unsigned)i - rstart(sp), len = rlen(sp);
            uint8_t* dst = out + s.keys[t];
            const unsigned x = delta_at(s.vals, (unsigned)i);
            if (w == 0u) {   // two 4-bit deltas per byte, high nibble first, odd tail low = 0
                if ((k & 1u) == 0u) {
                    const unsigned lo = (k + 1u < len) ? delta_at(s.vals, (unsigned)i + 1u) : 0u;
                    dst[k >> 1] = (uint8_t)((x << 4) | (lo & 0x0Fu));
                }
            } else {         // w bytes big-endian per delta
                for (unsigned q = 0; q < w; ++q) dst[k * w + q] = (uint8_t)(x >> (8u * (w - 1u - q)));
            }
        }
        starts += __popcll(F);
    }
}

/* Lists of <= kShort ids; pass 1 queues the longer ones: up to kMedium ids at the front of
   long_ids (count long_count[0]), more from its back (count long_count[1]). */
template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle_short(BarleArgs a) {
    __shared__ EncLds<kShort> s;
    const int64_t v = blockIdx.x;
    const int64_t b = a.csr_off[v];
    const int n = (int)(a.csr_off[v + 1] - b);
    if (n > kShort) {
        if ((!WRITE || a.slots) && threadIdx.x == 0) {
            if (n <= kMedium) a.long_ids[atomicAdd(&a.long_count[0], 1u)] = (int32_t)v;
            else a.long_ids[a.n - 1 - (int64_t)atomicAdd(&a.long_count[1], 1u)] = (int32_t)v;
        }
        return;
    }
    if (n <= 0) {
        if ((!WRITE || a.slots) && threadIdx.x == 0) a.sizes[v] = 0;
        return;
    }
    encode_voxel<WRITE, kShort>(a, s, v, b, n, threadIdx.x);
}

/* The queued lists of kShort + 1 .. kMedium ids, a fixed grid striding over the queue. */
template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle_long(BarleArgs a) {
    __shared__ EncLds<kMedium> s;
    const unsigned count = a.long_count[0];
    for (unsigned q = blockIdx.x; q < count; q += gridDim.x) {
        const int64_t v = a.long_ids[q];
        const int64_t b = a.csr_off[v];
        encode_voxel<WRITE, kMedium>(a, s, v, b, (int)(a.csr_off[v + 1] - b), threadIdx.x);
        __syncthreads();   // LDS reused by the next voxel
    }
}

/* The queued lists of more than kMedium ids (back of long_ids). */
template <bool WRITE>
__global__ __launch_bounds__(64) void k_barle_huge(BarleArgs a) {
    __shared__ EncLds<4096> s;
    const unsigned count = a.long_count[1];
    for (unsigned q = blockIdx.x; q < count; q += gridDim.x) {
        const int64_t v = a.long_ids[a.n - 1 - (int64_t)q];
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
    if (write_pass && a.slots) SGPU_HIP(hipMemsetAsync(a.long_count, 0, 2 * sizeof(unsigned), s));
    if (write_pass) {
        hipLaunchKernelGGL(k_barle_short<true>, g, blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_long<true>, dim3(kLongGrid), blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_huge<true>, dim3(kHugeGrid), blk, 0, s, a);
    } else {
        SGPU_HIP(hipMemsetAsync(a.long_count, 0, 2 * sizeof(unsigned), s));
        hipLaunchKernelGGL(k_barle_short<false>, g, blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_long<false>, dim3(kLongGrid), blk, 0, s, a);
        hipLaunchKernelGGL(k_barle_huge<false>, dim3(kHugeGrid), blk, 0, s, a);
    }
    SGPU_HIP(hipGetLastError());
}

/* One wave per voxel: its slot's sizes[v] bytes to out + out_off[v]. */
__global__ __launch_bounds__(256) void k_barle_compact(const uint8_t* slots, const int64_t* csr_off,
                                                       const int32_t* sizes, const int64_t* out_off, int64_t n,
                                                       uint8_t* out) {
    const int64_t v = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (v >= n) return;
    const int lane = threadIdx.x & 63;
    const uint8_t* src = slots + 6 * csr_off[v] + 8 * v;
    uint8_t* dst = out + out_off[v];
    const int len = sizes[v];
    for (int i = lane; i < len; i += 64) dst[i] = src[i];
}

void launch_barle_compact(const uint8_t* slots, const int64_t* csr_off, const int32_t* sizes, const int64_t* out_off,
                          int64_t n, uint8_t* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_barle_compact, dim3((unsigned)((n + 3) / 4)), dim3(256), 0, s, slots, csr_off, sizes, out_off,
                       n, out);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
