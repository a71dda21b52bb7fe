/* Stage 1 on gfx950: hemispherical DDA viewshed, SVF bitmap and CSR
 * neighbour lists.
 *
 * k_trace_pool  one workgroup per (G-direction group, 256-voxel block), the
 *           direction groups of a block consecutive in the grid, so the resident
 *           workgroups trace a few consecutive blocks and their first-hit writes
 *           fall in a few tens of MB of the raw table; each
 *           wave traces the G x 64 rays of its 64 voxels on the pooled walk of
 *           dda.h (a lane takes a new ray when its own has ended).  First-hit
 *           ids go to a per-chunk raw table raw[voxel][direction] (int32,
 *           0 = no hit).
 * k_lists   one workgroup per 8 voxels, one at a time: LDS hash set of the voxel's hit cells
 *           with the smallest direction index per cell (LDS atomics), then an
 *           ordered block scan keeps the first occurrence of every cell in
 *           direction order -- the reference's append-if-absent list
 *           (viewshed_kernels.cu:50-55) -- and ORs the SVF bits.  It stores
 *           the count and leaves the kept ids, in order, at the front of the
 *           voxel's raw row; after the scan of the counts, k_list_copy moves
 *           every row's list to its offset (one wave per voxel).
 *
 * The DDA keeps the reference's operation order (viewshed_kernels.cu:27-58);
 * the library is compiled with -ffp-contract=off. */
#include <climits>

#include "dda.h"
#include "sgpu_internal.h"

namespace sgpu {
namespace {

constexpr int kTraceBlock = 256;
constexpr int kListThreads = 256;
constexpr int kDirsPerThread = (kDirs + kListThreads - 1) / kListThreads;  // 11
constexpr int kHashSize = 4096;              // > 1.5 x the 2592 possible distinct hits

/* G directions per workgroup (2592 = 162 x 16 = 81 x 32 = 54 x 48): the tile shape the autotuner
 * picks.  Each wave owns the G x 64 rays of its 64 voxels (in the lane order: blocks of ids
 * sorted by height when a.perm is given) over the workgroup's G directions, taken
 * direction-major (the 64 voxels of one direction, then the next).  A lane whose ray has
 * ended takes the next ray of the pool once kRefill lanes are idle (or none is busy), so the
 * wave does not wait for its longest ray of every direction; a refill sets up the rays of
 * all idle lanes at once. */
constexpr int kRefill = 16;

/* x as a wave-uniform value (the same in every lane): lets indexing by it use scalar loads. */
__device__ __forceinline__ int64_t uniform_i64(int64_t x) {
    const int lo = __builtin_amdgcn_readfirstlane((int)(uint32_t)x);
    const int hi = __builtin_amdgcn_readfirstlane((int)(uint32_t)((uint64_t)x >> 32));
    return (int64_t)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo);
}

/* LDS of one trace workgroup: its G directions, and per wave the origins and raw rows of its 64
   voxels. */
template <int G>
struct TraceLds {
    DirConst dir[G];
    float org[kTraceBlock / 64][3][64];   // gx, gy, base z of the wave's voxels
    int row[kTraceBlock / 64][64];        // their raw rows
};

/* One (direction group grp, voxel block blk) item of a scene's trace, by the whole workgroup:
   stage the group's directions and the block's voxel origins, then every wave runs the pooled walk
   over its 64 voxels x G directions.  Contains one workgroup barrier, reached by every thread. */
template <bool COUNT_STEPS, int G>
__device__ __forceinline__ void trace_pool_item(const TraceArgs& a, int grp, int64_t blk, TraceLds<G>& L) {
    DirConst* sdir = L.dir;
    float (*sorg)[3][64] = L.org;
    int (*srow)[64] = L.row;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = threadIdx.x; i < G; i += kTraceBlock) sdir[i] = a.dirs[grp * G + i];
    const int64_t lane_pos = blk * kTraceBlock + threadIdx.x;
    const int64_t wave_pos = lane_pos - lane;
    const int nvalid = (int)min<int64_t>(64, max<int64_t>(0, a.count - wave_pos));
    if (lane < nvalid) {
        const int64_t local = a.perm ? (int64_t)a.perm[a.v0 + lane_pos] - a.v0 : lane_pos;   // raw row
        const int64_t v = a.subset ? (int64_t)a.subset[a.v0 + local] : a.v0 + local;
        const int top = a.pix_parent[v];
        const int gx = top % a.W, gy = top / a.W;
        sorg[wave][0][lane] = (float)gx;
        sorg[wave][1][lane] = (float)gy;
        sorg[wave][2][lane] = (v < a.n_top) ? a.grid[gx + gy * a.W] : a.wall_z[v - a.n_top];
        srow[wave][lane] = (int)local;
    }
    __syncthreads();
    if (nvalid == 0) return;

    const unsigned uW = (unsigned)a.W, uH = (unsigned)a.H;
    const unsigned lim_x = 4u * uW, lim_y = 4u * uW * uH;
    const float* __restrict__ grid = a.grid;
    const float maxz = a.maxz;
    unsigned long long steps = 0;
    /* the lane's ray */
    bool active = false;
    PoolRay ray{};
    int32_t* out = nullptr;
    const int total = G * 64;
    int next = 0;   // wave-uniform pool cursor
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (next < total && (nidle >= kRefill || nidle == 64)) {
            if (!active) {
                const int j = next + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                const int k = j >> 6, vx = j & 63;
                if (j < total && vx < nvalid) {
                    pool_ray_setup(ray, sdir[k], sorg[wave][0][vx], sorg[wave][1][vx], sorg[wave][2][vx], uW, maxz);
                    out = a.raw + (int64_t)srow[wave][vx] * kDirs + grp * G + k;
                    active = true;
                }
            }
            next += nidle;
        }
        if (__ballot(active) == 0) {
            if (next >= total) break;
            continue;
        }
        if (active) {
            unsigned res;
            if (pool_ray_steps<COUNT_STEPS>(ray, grid, lim_x, lim_y, steps, res)) {
                *out = (int32_t)res;
                active = false;
            }
        }
    }
    if (COUNT_STEPS) atomicAdd(a.steps, steps);
}

template <bool COUNT_STEPS, int G>
__global__ __launch_bounds__(kTraceBlock) void k_trace_pool(TraceArgs a) {
    __shared__ TraceLds<G> L;
    trace_pool_item<COUNT_STEPS, G>(a, (int)blockIdx.x, (int64_t)blockIdx.y, L);
}

/* Inclusive wave64 scan. */
__device__ __forceinline__ int wave_incl_scan(int x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

/* Exclusive scan over the 256 threads of a block; *total gets the block sum. */
__device__ __forceinline__ int block_excl_scan(int x, int* tmp, int* total) {
    const int inc = wave_incl_scan(x);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) tmp[wave] = inc;
    __syncthreads();
    int before = 0, all = 0;
#pragma unroll
    for (int w = 0; w < kListThreads / 64; ++w) {
        const int s = tmp[w];
        if (w < wave) before += s;
        all += s;
    }
    *total = all;
    return before + inc - x;
}

struct ListsLds {
    __align__(16) int keys[kHashSize];
    __align__(16) int mins[kHashSize];
    unsigned svf_acc[kAzBins * kElBins / 8 + 1];
    int tmp[kListThreads / 64];
};

/* Thread t's directions 11 t .. 11 t + 10 of a raw row (0 past the last direction). */
__device__ __forceinline__ void load_row(const int32_t* row, int (&hv)[kDirsPerThread]) {
    const int d0 = threadIdx.x * kDirsPerThread;
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k) hv[k] = (d0 + k < kDirs) ? row[d0 + k] : 0;
}

/* Dense SVF bit of thread t's directions (-1: none). */
__device__ __forceinline__ void load_dense(const DirConst* dirs, int (&dn)[kDirsPerThread]) {
    const int d0 = threadIdx.x * kDirsPerThread;
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k) dn[k] = (d0 + k < kDirs) ? (dirs[d0 + k].info >> 8) - 1 : -1;
}

/* One voxel's first-occurrence list from its raw row hv (workgroup-wide, thread t holding
 * directions 11 t .. 11 t + 10): the count, the voxel's SVF bytes when svf is given
 * (svf[byte * n + v]), and the kept ids in direction order written over the front of the row
 * (every read of the row precedes the barriers).  A hit equal to the thread's previous
 * direction's hit cannot be a first occurrence and skips the hash set (adjacent elevations of
 * one azimuth mostly hit the same cell: about three hits in four); the others are inserted
 * with all of a thread's compare-and-swaps in flight at once, probing on only after a slot
 * held another cell.  The LDS is free again when the call returns. */
__device__ __forceinline__ void voxel_list(ListsLds& s, const int (&hv)[kDirsPerThread],
                                           const int (&dn)[kDirsPerThread], int32_t* raw, uint8_t* svf,
                                           int svf_bytes, int64_t n, int64_t v, int32_t* count_out) {
    int* keys = s.keys;
    int* mins = s.mins;
    unsigned* svf_acc = s.svf_acc;
    const int tid = threadIdx.x;
    for (int i = tid; i < kHashSize / 4; i += kListThreads) {
        reinterpret_cast<int4*>(keys)[i] = make_int4(0, 0, 0, 0);
        reinterpret_cast<int4*>(mins)[i] = make_int4(INT_MAX, INT_MAX, INT_MAX, INT_MAX);
    }
    for (int i = tid; i < svf_bytes; i += kListThreads) svf_acc[i] = 0u;
    __syncthreads();
    const int d0 = tid * kDirsPerThread;
    bool cand[kDirsPerThread];
    unsigned sl[kDirsPerThread];
    int got[kDirsPerThread];
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k) {
        cand[k] = hv[k] != 0 && (k == 0 || hv[k] != hv[k - 1]);
        sl[k] = ((unsigned)hv[k] * 2654435761u) >> (32 - 12);
    }
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k) got[k] = cand[k] ? atomicCAS(&keys[sl[k]], 0, hv[k]) : 0;
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k) {
        if (cand[k]) {
            while (got[k] != 0 && got[k] != hv[k]) {
                sl[k] = (sl[k] + 1) & (kHashSize - 1);
                got[k] = atomicCAS(&keys[sl[k]], 0, hv[k]);
            }
            atomicMin(&mins[sl[k]], d0 + k);
        }
        if (svf && hv[k] != 0 && dn[k] >= 0) atomicOr(&svf_acc[dn[k] >> 3], 1u << (dn[k] & 7));
    }
    __syncthreads();
    int kept = 0;
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k)
        if (cand[k] && mins[sl[k]] == d0 + k) ++kept;
    int total;
    const int excl = block_excl_scan(kept, s.tmp, &total);   // barriers: the row is read
    if (tid == 0) *count_out = total;
    if (svf)
        for (int i = tid; i < svf_bytes; i += kListThreads) svf[(int64_t)i * n + v] = (uint8_t)svf_acc[i];
    int32_t* dst = raw + excl;
#pragma unroll
    for (int k = 0; k < kDirsPerThread; ++k)
        if (cand[k] && mins[sl[k]] == d0 + k) *dst++ = hv[k];
    __syncthreads();   // every LDS read of this voxel is done
}

/* kListVoxels consecutive voxels per workgroup; the next voxel's raw row is loaded into
 * registers while the current one goes through the hash set. */
constexpr int kListVoxels = 8;
__global__ __launch_bounds__(kListThreads) void k_lists(ListArgs a) {
    __shared__ ListsLds s;
    const int64_t first = (int64_t)blockIdx.x * kListVoxels;
    const int nvox = (int)min<int64_t>(kListVoxels, a.count - first);
    int hv[kDirsPerThread], nx[kDirsPerThread], dn[kDirsPerThread];
    load_dense(a.dirs, dn);
    load_row(a.raw + first * kDirs, hv);
    for (int i = 0; i < nvox; ++i) {
        const int64_t local = first + i;
        if (i + 1 < nvox) load_row(a.raw + (local + 1) * kDirs, nx);
        voxel_list(s, hv, dn, a.raw + local * kDirs, a.svf, a.svf_bytes, a.n_total, a.v0 + local,
                   a.counts + local);
#pragma unroll
        for (int k = 0; k < kDirsPerThread; ++k) hv[k] = nx[k];
    }
}

/* The same over the concatenated voxels of a batch: each voxel's tile (binary search on the
 * tile bases) gives its SVF bitmap and direction table. */
__global__ __launch_bounds__(kListThreads) void k_lists_batch(BatchListArgs a) {
    __shared__ ListsLds s;
    __shared__ int tile_s;
    const int64_t first = (int64_t)blockIdx.x * kListVoxels;
    const int nvox = (int)min<int64_t>(kListVoxels, a.count - first);
    int hv[kDirsPerThread], nx[kDirsPerThread], dn[kDirsPerThread];
    int cur = -1;   // the tile dn belongs to
    load_row(a.raw + first * kDirs, hv);
    for (int i = 0; i < nvox; ++i) {
        const int64_t local = first + i;
        const int64_t g = a.v0 + local;
        if (threadIdx.x == 0) {   // last tile with vbase <= g
            int lo = 0;
            if (i == 0) {
                int hi = a.ntiles - 1;
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (a.tiles[mid].vbase <= g) lo = mid;
                    else hi = mid - 1;
                }
            } else {   // the voxels are consecutive: from the previous voxel's tile on
                lo = tile_s;
                while (lo + 1 < a.ntiles && a.tiles[lo + 1].vbase <= g) ++lo;
            }
            tile_s = lo;   // the previous voxel's reads of tile_s precede voxel_list's barriers
        }
        __syncthreads();
        const int ti = tile_s;
        const TileDesc& t = a.tiles[ti];
        if (ti != cur) {
            load_dense(t.dirs, dn);
            cur = ti;
        }
        if (i + 1 < nvox) load_row(a.raw + (local + 1) * kDirs, nx);
        voxel_list(s, hv, dn, a.raw + local * kDirs, t.svf, t.svf_bytes, t.n_total, g - t.vbase, a.counts + local);
#pragma unroll
        for (int k = 0; k < kDirsPerThread; ++k) hv[k] = nx[k];
    }
}

/* Every voxel's kept ids (the front of its raw row) to its scanned offset; one wave per voxel. */
constexpr int kCopyWaves = 4;
__global__ __launch_bounds__(64 * kCopyWaves) void k_list_copy(const int32_t* raw, const int32_t* counts,
                                                               const int64_t* offsets, int32_t* ids, int64_t count) {
    const int64_t local = (int64_t)blockIdx.x * kCopyWaves + (threadIdx.x >> 6);
    if (local >= count) return;
    const int lane = threadIdx.x & 63;
    const int n = counts[local];
    const int32_t* src = raw + local * kDirs;
    int32_t* dst = ids + offsets[local];
    for (int i = lane; i < n; i += 64) dst[i] = src[i];
}

void launch_list_copy(const int32_t* raw, const int32_t* counts, const int64_t* offsets, int32_t* ids, int64_t count,
                      hipStream_t s) {
    const unsigned g = (unsigned)((count + kCopyWaves - 1) / kCopyWaves);
    hipLaunchKernelGGL(k_list_copy, dim3(g), dim3(64 * kCopyWaves), 0, s, raw, counts, offsets, ids, count);
}

/* Persistent multi-tile DDA: every workgroup takes (voxel block, direction group) items from a
 * global counter until the chunk is exhausted.  A chunk is a list of slices, each a run of one
 * tile's voxels (in the tile's height-sorted lane order) with its own TraceArgs; a block lies in
 * one slice, so the tile's heightmap, direction table and bounds are uniform over the item and the
 * item runs exactly the single-scene trace body (trace_pool_item).  Items are block-major: the
 * direction groups of a block are consecutive, like the single-scene grid. */
__global__ __launch_bounds__(kTraceBlock) void k_trace_batch_pool(BatchTraceArgs b) {
    constexpr int G = kPersistentDirs;
    constexpr int64_t groups = kDirs / G;
    __shared__ TraceLds<G> L;
    __shared__ long long item_s;
    for (;;) {
        if (threadIdx.x == 0) item_s = (long long)atomicAdd(b.next, 1ull);
        __syncthreads();   // also: every wave is done with the previous item's LDS
        const int64_t item = uniform_i64(item_s);
        if (item >= b.items) break;   // every wave of the workgroup leaves together
        const int64_t blk = item / groups;
        const int grp = (int)(item % groups);
        const int si = b.block_slice[blk];
        trace_pool_item<false, G>(b.slices[si], grp, (int64_t)b.block_local[blk], L);
    }
}

/* ---- int32 counts -> int64 exclusive offsets ---- */
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;
constexpr int kScanTile = kScanThreads * kScanItems;

__device__ __forceinline__ long long wave_incl_scan64(long long x) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        long long y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    return x;
}

__device__ __forceinline__ long long block_excl_scan64(long long x, long long* tmp, long long* total) {
    const long long inc = wave_incl_scan64(x);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 63) tmp[wave] = inc;
    __syncthreads();
    long long before = 0, all = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if (w < wave) before += tmp[w];
        all += tmp[w];
    }
    __syncthreads();
    *total = all;
    return before + inc - x;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_sums(const int32_t* c, int64_t n, int64_t* sums) {
    __shared__ long long tmp[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k)
        if (base + k < n) s += c[base + k];
    long long total;
    block_excl_scan64(s, tmp, &total);
    if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

/* Single block: exclusive scan of the tile sums in place; writes the grand total. */
__global__ __launch_bounds__(1024) void k_scan_tile_sums(int64_t* sums, int64_t ntiles, int64_t* total_out) {
    __shared__ long long tmp[1024 / 64];
    long long carry = 0;
    for (int64_t b = 0; b < ntiles; b += 1024) {
        const int64_t i = b + threadIdx.x;
        const long long x = (i < ntiles) ? sums[i] : 0;
        long long tot;
        const long long ex = block_excl_scan64(x, tmp, &tot);
        if (i < ntiles) sums[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *total_out = carry;
}

__global__ __launch_bounds__(kScanThreads) void k_tile_scan(const int32_t* c, int64_t n, const int64_t* sums,
                                                           int64_t* off) {
    __shared__ long long tmp[kScanThreads / 64];
    const int64_t base = (int64_t)blockIdx.x * kScanTile + (int64_t)threadIdx.x * kScanItems;
    int v[kScanItems];
    long long s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        v[k] = (base + k < n) ? c[base + k] : 0;
        s += v[k];
    }
    long long total;
    long long run = sums[blockIdx.x] + block_excl_scan64(s, tmp, &total);
    for (int k = 0; k < kScanItems; ++k) {
        if (base + k < n) off[base + k] = run;
        run += v[k];
    }
}

}  // namespace

void launch_trace(const TraceArgs& a, hipStream_t s) {
    if (a.count <= 0) return;
    if (!grid_off32((unsigned)a.W, (unsigned)a.H)) fail(SGPU_ERR_ARG, "grid beyond 32-bit byte offsets");
    const int G = a.group ? a.group : 32;
    const unsigned by = (unsigned)((a.count + kTraceBlock - 1) / kTraceBlock);
    dim3 grid((unsigned)(kDirs / G), by);
    switch (G) {
        case 16:
            if (a.steps) hipLaunchKernelGGL((k_trace_pool<true, 16>), grid, dim3(kTraceBlock), 0, s, a);
            else hipLaunchKernelGGL((k_trace_pool<false, 16>), grid, dim3(kTraceBlock), 0, s, a);
            break;
        case 32:
            if (a.steps) hipLaunchKernelGGL((k_trace_pool<true, 32>), grid, dim3(kTraceBlock), 0, s, a);
            else hipLaunchKernelGGL((k_trace_pool<false, 32>), grid, dim3(kTraceBlock), 0, s, a);
            break;
        case 48:
            if (a.steps) hipLaunchKernelGGL((k_trace_pool<true, 48>), grid, dim3(kTraceBlock), 0, s, a);
            else hipLaunchKernelGGL((k_trace_pool<false, 48>), grid, dim3(kTraceBlock), 0, s, a);
            break;
        default: fail(SGPU_ERR_ARG, "directions per group must be 16, 32 or 48");
    }
    SGPU_HIP(hipGetLastError());
}

void launch_lists(const ListArgs& a, bool write_pass, hipStream_t s) {
    if (a.count <= 0) return;
    if (a.svf_bytes > kAzBins * kElBins / 8 + 1) fail(SGPU_ERR_ARG, "svf_bytes %d out of range", a.svf_bytes);
    if (write_pass) launch_list_copy(a.raw, a.counts, a.offsets, a.ids, a.count, s);
    else hipLaunchKernelGGL(k_lists, dim3((unsigned)((a.count + kListVoxels - 1) / kListVoxels)), dim3(kListThreads), 0,
                            s, a);
    SGPU_HIP(hipGetLastError());
}

void launch_trace_persistent(const BatchTraceArgs& a, int workgroups, hipStream_t s) {
    if (a.items <= 0) return;
    hipLaunchKernelGGL(k_trace_batch_pool, dim3((unsigned)std::max(workgroups, 1)), dim3(kTraceBlock), 0, s, a);
    SGPU_HIP(hipGetLastError());
}

void launch_lists_batch(const BatchListArgs& a, bool write_pass, hipStream_t s) {
    if (a.count <= 0) return;
    if (write_pass) launch_list_copy(a.raw, a.counts, a.offsets, a.ids, a.count, s);
    else hipLaunchKernelGGL(k_lists_batch, dim3((unsigned)((a.count + kListVoxels - 1) / kListVoxels)),
                            dim3(kListThreads), 0, s, a);
    SGPU_HIP(hipGetLastError());
}

size_t scan_scratch_elems(int64_t n) { return (size_t)((n + kScanTile - 1) / kScanTile) + 1; }

void launch_scan_counts(const int32_t* counts, int64_t n, int64_t* off, int64_t* scratch, hipStream_t s) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    if (tiles == 0) {
        SGPU_HIP(hipMemsetAsync(off, 0, sizeof(int64_t), s));
        return;
    }
    hipLaunchKernelGGL(k_tile_sums, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, counts, n, scratch);
    hipLaunchKernelGGL(k_scan_tile_sums, dim3(1), dim3(1024), 0, s, scratch, tiles, off + n);
    hipLaunchKernelGGL(k_tile_scan, dim3((unsigned)tiles), dim3(kScanThreads), 0, s, counts, n, scratch, off);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
