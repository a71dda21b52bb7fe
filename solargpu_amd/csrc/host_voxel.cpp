/* Voxelization contract (host C++): column-drop wall voxels and top-cell
 * slope/aspect, as in the reference's voxelization.cpp:10-97, plus the voxel
 * array assembly of main.cpp:112-127.  Built with -ffp-contract=off. */
#include <algorithm>
#include <cmath>
#include <thread>
#include <vector>

#include "sgpu_internal.h"

namespace sgpu {

/* voxelization.cpp:10-64.  A cell whose drop to its lowest 4-neighbour (an
 * off-grid side counts as no drop) exceeds one voxel gets floor(drop/voxel)
 * stacked wall voxels; each faces the side(s) open at its height, where an
 * off-grid side is fully open (-1e30 sentinel, not clamped). */
Walls detect_walls(const float* H, int w, int h, float voxel) {
    Walls out;
    const float kOpen = -1e30f;
    for (int gy = 0; gy < h; ++gy) {
        for (int gx = 0; gx < w; ++gx) {
            const float z0 = H[(size_t)gy * w + gx];
            const float nb[4] = {
                gx + 1 < w ? H[(size_t)gy * w + gx + 1] : kOpen,   // east
                gx - 1 >= 0 ? H[(size_t)gy * w + gx - 1] : kOpen,  // west
                gy - 1 >= 0 ? H[(size_t)(gy - 1) * w + gx] : kOpen, // north
                gy + 1 < h ? H[(size_t)(gy + 1) * w + gx] : kOpen,  // south
            };
            float lowest = z0;
            bool first = true;
            for (float v : nb) {
                const float c = (v <= kOpen * 0.5f) ? z0 : v;
                lowest = first ? c : std::min(lowest, c);
                first = false;
            }
            const float drop = z0 - lowest;
            if (drop <= voxel) continue;
            const int count = (int)(drop / voxel);
            for (int k = 0; k < count; ++k) {
                const float zz = z0 - (k + 0.5f) * voxel;
                const float oe = std::max(0.0f, zz - nb[0]), ow = std::max(0.0f, zz - nb[1]);
                const float on = std::max(0.0f, zz - nb[2]), os = std::max(0.0f, zz - nb[3]);
                float a = std::atan2(oe - ow, on - os);
                if (a < 0.0f) a += 2.0f * (float)3.14159265358979323846;
                out.parent.push_back(gy * w + gx);
                out.z.push_back(zz);
                out.aspect.push_back(a);
            }
        }
    }
    return out;
}

/* voxelization.cpp:66-97: central differences clamped at the border; slope is
 * atan(|grad|), aspect atan2(-dx, dy) wrapped to [0, 2pi). */
void top_slope_aspect(const float* H, int w, int h, double sx, double sy, float* slope, float* aspect) {
    for (int gy = 0; gy < h; ++gy) {
        const int up = std::max(gy - 1, 0), dn = std::min(gy + 1, h - 1);
        for (int gx = 0; gx < w; ++gx) {
            const int lf = std::max(gx - 1, 0), rt = std::min(gx + 1, w - 1);
            const float dx = (H[(size_t)gy * w + rt] - H[(size_t)gy * w + lf]) / ((rt - lf) * (float)sx);
            const float dy = (H[(size_t)dn * w + gx] - H[(size_t)up * w + gx]) / ((dn - up) * (float)sy);
            const size_t i = (size_t)gy * w + gx;
            slope[i] = std::atan2(std::sqrt(dx * dx + dy * dy), 1.0f);
            float a = std::atan2(-dx, dy);
            if (a < 0.0f) a += 2.0f * (float)3.14159265358979323846;
            aspect[i] = a;
        }
    }
}

std::vector<int32_t> trace_order(int64_t n_total, const float* grid, const int32_t* pp, const float* wall_z,
                                 int64_t n_top, int block) {
    std::vector<int32_t> perm((size_t)n_total);
    const int64_t nblocks = (n_total + block - 1) / block;
    const int nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(16, std::min<int64_t>(
                                                        std::thread::hardware_concurrency(), nblocks / 64)));
    auto work = [&](int t) {
        std::vector<std::pair<std::pair<float, float>, int32_t>> key;
        for (int64_t blk = t; blk < nblocks; blk += nthreads) {
            const int64_t b = blk * block, e = std::min<int64_t>(n_total, b + block);
            key.clear();
            for (int64_t v = b; v < e; ++v) {
                const float ground = grid[pp[v]];
                key.push_back({{ground, v < n_top ? ground : wall_z[v - n_top]}, (int32_t)v});
            }
            std::stable_sort(key.begin(), key.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
            for (int64_t i = b; i < e; ++i) perm[(size_t)i] = key[(size_t)(i - b)].second;
        }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthreads; ++t) pool.emplace_back(work, t);
    work(0);
    for (std::thread& th : pool) th.join();
    return perm;
}

void class_order(int64_t n_total, int W, const int32_t* pp, std::vector<int32_t>& perm, std::vector<ClassRun>& runs) {
    perm.assign((size_t)n_total, 0);
    runs.clear();
    std::vector<uint8_t> key((size_t)std::min<int64_t>(n_total, kClassBlock));
    for (int64_t b = 0; b < n_total; b += kClassBlock) {
        const int64_t e = std::min<int64_t>(n_total, b + kClassBlock);
        int64_t count[256] = {0};
        for (int64_t v = b; v < e; ++v) {
            const int k = class_key(pp[v] % W, pp[v] / W);
            key[(size_t)(v - b)] = (uint8_t)k;
            ++count[k];
        }
        int64_t start[256];
        int64_t at = b;
        for (int k = 0; k < 256; ++k) {   // a counting sort: stable, so raster order within a class
            start[k] = at;
            if (count[k]) runs.push_back(ClassRun{at, (int32_t)count[k], k});
            at += count[k];
        }
        for (int64_t v = b; v < e; ++v) perm[(size_t)start[key[(size_t)(v - b)]]++] = (int32_t)v;
    }
}

std::vector<std::pair<int, int>> hilbert_tiles(int tw, int th) {
    int side = 1;
    while (side < std::max(tw, th)) side <<= 1;
    std::vector<std::pair<int, int>> tiles;
    tiles.reserve((size_t)tw * th);
    for (int64_t d = 0; d < (int64_t)side * side; ++d) {
        int x = 0, y = 0;
        int64_t t = d;
        for (int s = 1; s < side; s <<= 1) {   // Hilbert index -> (x, y)
            const int rx = (int)(1 & (t / 2)), ry = (int)(1 & (t ^ rx));
            if (ry == 0) {
                if (rx == 1) {
                    x = s - 1 - x;
                    y = s - 1 - y;
                }
                std::swap(x, y);
            }
            x += s * rx;
            y += s * ry;
            t /= 4;
        }
        if (x < tw && y < th) tiles.push_back({x, y});
    }
    return tiles;
}

std::vector<int32_t> processing_order(int w, int h, int64_t n_total, const int32_t* pp, int tile) {
    const int64_t n_top = (int64_t)w * h;
    /* walls grouped by parent cell, stable (counting sort) */
    std::vector<int64_t> start(n_top + 1, 0);
    for (int64_t v = n_top; v < n_total; ++v) ++start[(int64_t)pp[v] + 1];
    for (int64_t c = 0; c < n_top; ++c) start[c + 1] += start[c];
    std::vector<int32_t> walls((size_t)(n_total - n_top));
    {
        std::vector<int64_t> fill(start.begin(), start.end() - 1);
        for (int64_t v = n_top; v < n_total; ++v) walls[(size_t)fill[pp[v]]++] = (int32_t)v;
    }
    std::vector<int32_t> order;
    order.reserve((size_t)n_total);
    const std::vector<std::pair<int, int>> tiles = hilbert_tiles((w + tile - 1) / tile, (h + tile - 1) / tile);
    for (const auto& tt : tiles) {
        const int tx = tt.first * tile, ty = tt.second * tile;
        for (int y = ty; y < std::min(h, ty + tile); ++y)
            for (int x = tx; x < std::min(w, tx + tile); ++x) {
                const int64_t c = (int64_t)y * w + x;
                order.push_back((int32_t)c);
                for (int64_t k = start[c]; k < start[c + 1]; ++k) order.push_back(walls[(size_t)k]);
            }
    }
    return order;
}

}  // namespace sgpu
