/* Voxelization on the device: the wall voxels and top-cell slope/aspect of
 * voxelization.cpp:10-97 and the voxel array assembly of main.cpp:112-127, for
 * scenes set from a heightmap (sgpu_set_scene_heights).
 *
 *   k_vox_count  lane = cell: the cell's wall count (drop to its lowest
 *                4-neighbour over the voxel size, an off-grid side counting as
 *                no drop);
 *   (int64 exclusive scan of the counts: the walls' offsets, in the
 *    reference's (gy, gx, k) order)
 *   k_vox_write  lane = cell: the top voxel (parent = itself, slope and aspect
 *                from central differences clamped at the border) and its wall
 *                voxels at the scanned offsets (z = z0 - (k + 1/2) voxel, slope
 *                pi/2, aspect from the openness of the four sides, an off-grid
 *                side fully open).
 *
 * Same expressions as the host restatement (std::min / std::max as the ternaries
 * they are, correctly rounded division, FP contraction off), so parents and wall
 * heights are bit-identical; slope and aspect go through the device atan2f.
 *
 * The scene's two processing orders, as the host builds them (host_voxel.cpp):
 *   k_order_cells   lane = cell: its rank in Hilbert-tile order (kOrderTile-cell tiles
 *                   along a Hilbert curve, raster inside a tile; tile ranks and
 *                   first ranks from the host) and its voxel count 1 + walls;
 *   (int64 scan of the counts in that order)
 *   k_order_write   lane = rank: the cell, then its walls (contiguous ids);
 *   k_trace_perm    workgroup = block of 512 voxel ids: bitonic sort by (parent
 *                   cell height, voxel height, id) -- the host's stable sort. */
#include <climits>
#include <vector>

#include "sgpu_internal.h"

namespace sgpu {
namespace {

constexpr float kOpen = -1e30f;
constexpr int kVoxBlock = 256;

struct CellNb {
    float z0, nb[4];   // the cell's height; east, west, north, south (kOpen off the grid)
};

__device__ __forceinline__ CellNb cell_nb(const float* H, int w, int h, int gx, int gy) {
    CellNb c;
    const size_t i = (size_t)gy * w + gx;
    c.z0 = H[i];
    c.nb[0] = gx + 1 < w ? H[i + 1] : kOpen;
    c.nb[1] = gx - 1 >= 0 ? H[i - 1] : kOpen;
    c.nb[2] = gy - 1 >= 0 ? H[i - w] : kOpen;
    c.nb[3] = gy + 1 < h ? H[i + w] : kOpen;
    return c;
}

/* voxelization.cpp:18-33: walls of the cell (0 when the drop is at most one voxel). */
__device__ __forceinline__ int wall_count(const CellNb& c, float voxel) {
    float lowest = c.z0;
#pragma unroll
    for (int s = 0; s < 4; ++s) {
        const float v = (c.nb[s] <= kOpen * 0.5f) ? c.z0 : c.nb[s];
        lowest = (s == 0) ? v : ((v < lowest) ? v : lowest);   // std::min(lowest, v)
    }
    const float drop = c.z0 - lowest;
    if (drop <= voxel) return 0;
    return (int)(drop / voxel);
}

__device__ __forceinline__ float max0(float x) { return (0.0f < x) ? x : 0.0f; }   // std::max(0.0f, x)

__global__ __launch_bounds__(kVoxBlock) void k_vox_count(const float* H, int w, int h, float voxel, int32_t* counts) {
    const int64_t i = (int64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (i >= (int64_t)w * h) return;
    const int gx = (int)(i % w), gy = (int)(i / w);
    counts[i] = wall_count(cell_nb(H, w, h, gx, gy), voxel);
}

__global__ __launch_bounds__(kVoxBlock) void k_vox_write(const float* H, int w, int h, float sx, float sy,
                                                         float voxel, const int64_t* off, int32_t* pp,
                                                         float* slopes, float* aspects, float* wall_z) {
    const int64_t i = (int64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    const int64_t n_top = (int64_t)w * h;
    if (i >= n_top) return;
    const int gx = (int)(i % w), gy = (int)(i / w);
    const float two_pi = 2.0f * (float)3.14159265358979323846;
    /* voxelization.cpp:66-97: top voxel */
    {
        const int up = max(gy - 1, 0), dn = min(gy + 1, h - 1);
        const int lf = max(gx - 1, 0), rt = min(gx + 1, w - 1);
        const float dx = (H[(size_t)gy * w + rt] - H[(size_t)gy * w + lf]) / ((float)(rt - lf) * sx);
        const float dy = (H[(size_t)dn * w + gx] - H[(size_t)up * w + gx]) / ((float)(dn - up) * sy);
        pp[i] = (int32_t)i;
        slopes[i] = atan2f(sqrtf(dx * dx + dy * dy), 1.0f);
        float a = atan2f(-dx, dy);
        if (a < 0.0f) a += two_pi;
        aspects[i] = a;
    }
    /* voxelization.cpp:34-62: wall voxels, main.cpp:118-126: appended after the tops */
    const CellNb c = cell_nb(H, w, h, gx, gy);
    const int count = wall_count(c, voxel);
    const int64_t base = off[i];
    for (int k = 0; k < count; ++k) {
        const float zz = c.z0 - ((float)k + 0.5f) * voxel;
        const float oe = max0(zz - c.nb[0]), ow = max0(zz - c.nb[1]);
        const float on = max0(zz - c.nb[2]), os = max0(zz - c.nb[3]);
        float a = atan2f(oe - ow, on - os);
        if (a < 0.0f) a += two_pi;
        const int64_t j = base + k;
        pp[n_top + j] = (int32_t)i;
        wall_z[j] = zz;
        slopes[n_top + j] = (float)3.14159265358979323846 * 0.5f;
        aspects[n_top + j] = a;
    }
}

__global__ __launch_bounds__(kVoxBlock) void k_order_cells(int w, int h, int tile, int tw, const int32_t* tile_rank,
                                                           const int64_t* tile_first, const int64_t* wall_off,
                                                           int32_t* cell_at, int32_t* vcount) {
    const int64_t c = (int64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (c >= (int64_t)w * h) return;
    const int x = (int)(c % w), y = (int)(c / w);
    const int tx = x / tile, ty = y / tile;
    const int x0 = tx * tile, y0 = ty * tile, wt = min(tile, w - x0);
    const int64_t pos = tile_first[tile_rank[ty * tw + tx]] + (int64_t)(y - y0) * wt + (x - x0);
    cell_at[pos] = (int32_t)c;
    vcount[pos] = 1 + (int32_t)(wall_off[c + 1] - wall_off[c]);
}

__global__ __launch_bounds__(kVoxBlock) void k_order_write(int64_t n_top, const int32_t* cell_at, const int64_t* vpos,
                                                           const int64_t* wall_off, int32_t* order) {
    const int64_t j = (int64_t)blockIdx.x * kVoxBlock + threadIdx.x;
    if (j >= n_top) return;
    const int32_t c = cell_at[j];
    const int64_t p = vpos[j];
    order[p] = c;
    const int64_t w0 = wall_off[c], nw = wall_off[c + 1] - w0;
    for (int64_t k = 0; k < nw; ++k) order[p + 1 + k] = (int32_t)(n_top + w0 + k);
}

constexpr int kPermBlock = 512;

__device__ __forceinline__ bool perm_less(float ga, float za, int ia, float gb, float zb, int ib) {
    return ga < gb || (ga == gb && (za < zb || (za == zb && ia < ib)));
}

__global__ __launch_bounds__(256) void k_trace_perm(const float* grid, const int32_t* pp, const float* wall_z,
                                                    int64_t n_top, int64_t n_total, int32_t* perm) {
    __shared__ float sg[kPermBlock], sz[kPermBlock];
    __shared__ int si[kPermBlock];
    const int64_t b0 = (int64_t)blockIdx.x * kPermBlock;
    for (int i = threadIdx.x; i < kPermBlock; i += 256) {
        const int64_t v = b0 + i;
        if (v < n_total) {
            const float g = grid[pp[v]];
            sg[i] = g;
            sz[i] = (v < n_top) ? g : wall_z[v - n_top];
            si[i] = (int)v;
        } else {   // padding sorts last
            sg[i] = INFINITY;
            sz[i] = INFINITY;
            si[i] = INT_MAX;
        }
    }
    __syncthreads();
    for (int k = 2; k <= kPermBlock; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int i = threadIdx.x; i < kPermBlock; i += 256) {
                const int ixj = i ^ j;
                if (ixj > i) {
                    const bool up = (i & k) == 0;   // this k-block ends up ascending
                    const bool gt = perm_less(sg[ixj], sz[ixj], si[ixj], sg[i], sz[i], si[i]);
                    if (up == gt) {
                        const float tg = sg[i], tz = sz[i];
                        const int ti = si[i];
                        sg[i] = sg[ixj];
                        sz[i] = sz[ixj];
                        si[i] = si[ixj];
                        sg[ixj] = tg;
                        sz[ixj] = tz;
                        si[ixj] = ti;
                    }
                }
            }
            __syncthreads();
        }
    }
    for (int i = threadIdx.x; i < kPermBlock; i += 256)
        if (b0 + i < n_total) perm[b0 + i] = si[i];
}

}  // namespace

void orders_device(hipStream_t s, const float* heights_dev, int w, int h, const DeviceVoxels& v, int64_t n_total,
                   int tile, int block, int32_t* order, int32_t* perm) {
    if (block != kPermBlock) fail(SGPU_ERR_ARG, "lane-order blocks of %d ids", kPermBlock);
    const int64_t n_top = (int64_t)w * h;
    const int tw = (w + tile - 1) / tile, th = (h + tile - 1) / tile;
    const std::vector<std::pair<int, int>> tiles = hilbert_tiles(tw, th);
    std::vector<int32_t> rank((size_t)tw * th);
    std::vector<int64_t> first(tiles.size());
    int64_t at = 0;
    for (size_t r = 0; r < tiles.size(); ++r) {
        const int tx = tiles[r].first, ty = tiles[r].second;
        rank[(size_t)ty * tw + tx] = (int32_t)r;
        first[r] = at;
        at += (int64_t)std::min(tile, w - tx * tile) * std::min(tile, h - ty * tile);
    }
    int32_t *rank_d = nullptr, *cell_at = nullptr, *vcount = nullptr;
    int64_t *first_d = nullptr, *vpos = nullptr, *scratch = nullptr;
    auto release = [&] {
        for (void* p : {(void*)rank_d, (void*)cell_at, (void*)vcount, (void*)first_d, (void*)vpos, (void*)scratch})
            if (p) (void)hipFree(p);
    };
    try {
        SGPU_HIP(hipMalloc(&rank_d, rank.size() * sizeof(int32_t)));
        SGPU_HIP(hipMalloc(&first_d, first.size() * sizeof(int64_t)));
        SGPU_HIP(hipMalloc(&cell_at, (size_t)n_top * sizeof(int32_t)));
        SGPU_HIP(hipMalloc(&vcount, (size_t)n_top * sizeof(int32_t)));
        SGPU_HIP(hipMalloc(&vpos, (size_t)(n_top + 1) * sizeof(int64_t)));
        SGPU_HIP(hipMalloc(&scratch, scan_scratch_elems(n_top) * sizeof(int64_t)));
        SGPU_HIP(hipMemcpyAsync(rank_d, rank.data(), rank.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(first_d, first.data(), first.size() * sizeof(int64_t), hipMemcpyHostToDevice, s));
        const unsigned g = (unsigned)((n_top + kVoxBlock - 1) / kVoxBlock);
        hipLaunchKernelGGL(k_order_cells, dim3(g), dim3(kVoxBlock), 0, s, w, h, tile, tw, rank_d, first_d,
                           v.wall_off, cell_at, vcount);
        SGPU_HIP(hipGetLastError());
        launch_scan_counts(vcount, n_top, vpos, scratch, s);
        hipLaunchKernelGGL(k_order_write, dim3(g), dim3(kVoxBlock), 0, s, n_top, cell_at, vpos, v.wall_off, order);
        SGPU_HIP(hipGetLastError());
        const unsigned gb = (unsigned)((n_total + kPermBlock - 1) / kPermBlock);
        hipLaunchKernelGGL(k_trace_perm, dim3(gb), dim3(256), 0, s, heights_dev, v.pix_parent, v.wall_z, n_top,
                           n_total, perm);
        SGPU_HIP(hipGetLastError());
        SGPU_HIP(hipStreamSynchronize(s));   // the host tables above must outlive the copies
    } catch (...) {
        release();
        throw;
    }
    release();
}

int64_t voxelize_device(hipStream_t s, const float* heights_dev, int w, int h, double sx, double sy,
                        DeviceVoxels* out) {
    const int64_t n_top = (int64_t)w * h;
    const float voxel = (float)sx;   // main.cpp:113: the voxel size is the x pixel scale
    const unsigned g = (unsigned)((n_top + kVoxBlock - 1) / kVoxBlock);
    int32_t* counts = nullptr;
    int64_t* scratch = nullptr;
    int64_t n_walls = 0;
    auto release_tmp = [&] {
        if (counts) (void)hipFree(counts);
        if (scratch) (void)hipFree(scratch);
    };
    try {
        SGPU_HIP(hipMalloc(&counts, (size_t)n_top * sizeof(int32_t)));
        SGPU_HIP(hipMalloc(&out->wall_off, (size_t)(n_top + 1) * sizeof(int64_t)));
        int64_t* off = out->wall_off;
        SGPU_HIP(hipMalloc(&scratch, scan_scratch_elems(n_top) * sizeof(int64_t)));
        hipLaunchKernelGGL(k_vox_count, dim3(g), dim3(kVoxBlock), 0, s, heights_dev, w, h, voxel, counts);
        SGPU_HIP(hipGetLastError());
        launch_scan_counts(counts, n_top, off, scratch, s);
        SGPU_HIP(hipMemcpyAsync(&n_walls, off + n_top, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        SGPU_HIP(hipStreamSynchronize(s));
        if (n_top + n_walls >= ((int64_t)1 << 31)) fail(SGPU_ERR_ARG, "num_total_voxels out of range");
        const int64_t n_total = n_top + n_walls;
        SGPU_HIP(hipMalloc(&out->pix_parent, (size_t)n_total * sizeof(int32_t)));
        SGPU_HIP(hipMalloc(&out->slopes, (size_t)n_total * sizeof(float)));
        SGPU_HIP(hipMalloc(&out->aspects, (size_t)n_total * sizeof(float)));
        SGPU_HIP(hipMalloc(&out->wall_z, (size_t)std::max<int64_t>(n_walls, 1) * sizeof(float)));
        hipLaunchKernelGGL(k_vox_write, dim3(g), dim3(kVoxBlock), 0, s, heights_dev, w, h, (float)sx, (float)sy,
                           voxel, off, out->pix_parent, out->slopes, out->aspects, out->wall_z);
        SGPU_HIP(hipGetLastError());
        SGPU_HIP(hipStreamSynchronize(s));
    } catch (...) {
        release_tmp();
        throw;
    }
    release_tmp();
    return n_top + n_walls;
}

}  // namespace sgpu
