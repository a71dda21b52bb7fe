/* The hemispherical-viewshed DDA step loop shared by every kernel that traces
 * rays (plain viewshed, surrogate sun-bin trace), so their hits are identical.
 * Operation order of viewshed_kernels.cu:27-58; compiled with
 * -ffp-contract=off and IEEE division. */
#pragma once

#include "sgpu_internal.h"

namespace sgpu {

/* Steps whose cells and heights are fetched before any of them is tested: the
 * traversal does not depend on the heights, so kDdaUnroll loads are in flight at
 * once; the tests then run in step order, so the first exit wins as before. */
constexpr int kDdaUnroll = 4;

/* First hit (cell id + 1) of the ray from integer cell (gx, gy) at height
 * base_z in direction dc, or 0.  `steps` counts DDA steps when COUNT. */
template <bool COUNT>
__device__ __forceinline__ int dda_first_hit(const float* __restrict__ grid, int W, unsigned uW, unsigned uH,
                                             float maxz, float gxf, float gyf, float base_z, const DirConst& dc,
                                             unsigned long long& steps) {
    const bool upper = dc.info & 1;
    const float ox = gxf + dc.cx * 2.0f;
    const float oy = gyf + dc.cy * 2.0f;
    const float oz = base_z + dc.cz * 2.0f;
    int ix = (int)floorf(ox);
    int iy = (int)floorf(oy);
    float tmx = ((dc.cx >= 0.0f) ? ((float)ix + 1.0f - ox) : (ox - (float)ix)) / dc.ax;
    float tmy = ((dc.cy >= 0.0f) ? ((float)iy + 1.0f - oy) : (oy - (float)iy)) / dc.ay;
    const int sx = (dc.cx >= 0.0f) ? 1 : -1;
    const int sy = (dc.cy >= 0.0f) ? 1 : -1;
    for (;;) {
        float t[kDdaUnroll], h[kDdaUnroll];
        int cell[kDdaUnroll];
        bool inside[kDdaUnroll];
#pragma unroll
        for (int u = 0; u < kDdaUnroll; ++u) {
            const bool xs = tmx < tmy;   // a tie steps in y
            t[u] = xs ? tmx : tmy;
            ix += xs ? sx : 0;
            iy += xs ? 0 : sy;
            tmx = xs ? tmx + dc.rx : tmx;
            tmy = xs ? tmy : tmy + dc.ry;
            inside[u] = (unsigned)ix < uW && (unsigned)iy < uH;
            cell[u] = inside[u] ? ix + iy * W : 0;   // off-grid steps read a valid cell, unused
            h[u] = grid[cell[u]];
        }
#pragma unroll
        for (int u = 0; u < kDdaUnroll; ++u) {
            if (COUNT) ++steps;
            if (!inside[u]) return 0;
            const float zr = oz + dc.cz * t[u];
            if (upper ? (zr > maxz) : (zr < 0.0f)) return 0;
            if (h[u] - zr > 0.01f) return cell[u] + 1;
        }
    }
}

}  // namespace sgpu
