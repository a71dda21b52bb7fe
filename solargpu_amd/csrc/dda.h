/* The hemispherical-viewshed DDA shared by every kernel that traces rays (plain
 * viewshed, batched viewshed, surrogate sun-bin trace), so their hits are
 * identical.  Operation order of viewshed_kernels.cu:27-58 (origin two units
 * along the direction from the integer cell, floor, first-edge distances over
 * |c|, strict tMaxX < tMaxY else a y step, exit off the grid or through maxz /
 * z = 0, hit when grid - z_ray > 0.01); compiled with -ffp-contract=off.
 *
 * Kernels run the walk pooled: the lanes of a wave are on different rays, so a
 * ray's direction (step signs, hemisphere) is per-lane state.  The grid is
 * addressed by 32-bit byte offsets (grids with W (H + 2) <= 2^30, checked when a
 * scene is set): the column and row byte offsets (4 ix, 4 W iy) step as unsigned
 * values, a column or row one or two steps off the grid wraps to a value >= 4 W,
 * resp. >= 4 W H, so one unsigned compare each is the bounds test and their sum
 * is the cell's byte offset. */
#pragma once

#include "sgpu_internal.h"

namespace sgpu {

/* a / b correctly rounded, from rb = RN(1 / b) (the direction table's rx / ry): q0 = RN(a rb),
 * the exact remainder a - q0 b by one FMA, then one FMA correction (Markstein's theorem for a
 * reciprocal within half an ulp).  Three instructions instead of the IEEE division sequence;
 * for every ray set-up numerator of every direction up to 32768 cells per axis it equals the
 * IEEE quotient bit for bit (checked exhaustively against the oracle's division). */
__device__ __forceinline__ float div_by(float a, float b, float rb) {
    const float q0 = a * rb;
    const float r = __builtin_fmaf(-q0, b, a);
    return __builtin_fmaf(r, rb, q0);
}

/* Steps whose cells and heights are fetched before any of them is tested: the
 * traversal does not depend on the heights, so kDdaUnroll loads are in flight at
 * once; the tests then run in step order, so the first exit wins. */
constexpr int kDdaUnroll = 3;

/* One ray of a pooled walk. */
struct PoolRay {
    unsigned cx4, ry4, sx4, sy4;   // column / row byte offsets and their steps
    float tmx, tmy, oz, cz, rx, ry;
    float zhi, zlo;                // leaves above zhi (upper) or below zlo (lower)
};

/* Start cell (ix, iy) and first-edge distances (tmx, tmy) of the ray from integer cell (gx, gy)
   along dc (viewshed_kernels.cu:27-35); the walk and the step tables both set up through it. */
__device__ __forceinline__ void ray_start(const DirConst& dc, float gxf, float gyf, int& ix, int& iy, float& tmx,
                                          float& tmy) {
    const float ox = gxf + dc.cx * 2.0f;
    const float oy = gyf + dc.cy * 2.0f;
    ix = (int)floorf(ox);
    iy = (int)floorf(oy);
    tmx = div_by((dc.cx >= 0.0f) ? ((float)ix + 1.0f - ox) : (ox - (float)ix), dc.ax, dc.rx);
    tmy = div_by((dc.cy >= 0.0f) ? ((float)iy + 1.0f - oy) : (oy - (float)iy), dc.ay, dc.ry);
}

/* Set-up of the ray from integer cell (gx, gy) at base_z along dc (viewshed_kernels.cu:27-35). */
__device__ __forceinline__ void pool_ray_setup(PoolRay& r, const DirConst& dc, float gxf, float gyf, float base_z,
                                               unsigned uW, float maxz) {
    int ix, iy;
    ray_start(dc, gxf, gyf, ix, iy, r.tmx, r.tmy);
    r.oz = base_z + dc.cz * 2.0f;
    r.sx4 = (dc.cx >= 0.0f) ? 4u : (unsigned)-4;
    r.sy4 = (dc.cy >= 0.0f) ? 4u * uW : (unsigned)-(int)(4u * uW);
    r.cx4 = 4u * (unsigned)ix;
    r.ry4 = 4u * uW * (unsigned)iy;
    r.cz = dc.cz;
    r.rx = dc.rx;
    r.ry = dc.ry;
    const bool upper = dc.info & 1;
    r.zhi = upper ? maxz : INFINITY;
    r.zlo = upper ? -INFINITY : 0.0f;
}

/* kDdaUnroll steps of a ray (viewshed_kernels.cu:37-58), branch-free: the x/y step is a select,
 * the cells' heights are fetched first (an off-grid step reads cell 0, unused), the exit tests
 * run in step order as selects.  True when the ray has ended, with *res the hit cell id + 1 or
 * 0; `steps` counts the steps taken when COUNT. */
template <bool COUNT>
__device__ __forceinline__ bool pool_ray_steps(PoolRay& r, const float* __restrict__ grid, unsigned lim_x,
                                               unsigned lim_y, unsigned long long& steps, unsigned& res) {
    float t[kDdaUnroll], h[kDdaUnroll];
    unsigned off[kDdaUnroll];
    bool inside[kDdaUnroll];
#pragma unroll
    for (int u = 0; u < kDdaUnroll; ++u) {
        const bool xs = r.tmx < r.tmy;   // a tie steps in y
        t[u] = xs ? r.tmx : r.tmy;
        r.cx4 = xs ? r.cx4 + r.sx4 : r.cx4;
        r.ry4 = xs ? r.ry4 : r.ry4 + r.sy4;
        r.tmx = xs ? r.tmx + r.rx : r.tmx;
        r.tmy = xs ? r.tmy : r.tmy + r.ry;
        inside[u] = r.cx4 < lim_x && r.ry4 < lim_y;
        off[u] = inside[u] ? r.cx4 + r.ry4 : 0u;
        h[u] = *(const float*)((const char*)grid + off[u]);
    }
    /* the stopping step's cell offset and whether it is a hit (inside, below the exits); its id is
       formed once, after the steps */
    unsigned stop_off = 0;
    bool hit = false, done = false;
#pragma unroll
    for (int u = 0; u < kDdaUnroll; ++u) {
        if (COUNT) steps += done ? 0 : 1;
        const float zr = r.oz + r.cz * t[u];
        const bool zout = (zr > r.zhi) | (zr < r.zlo);
        const bool stop = !inside[u] | zout | (h[u] - zr > 0.01f);
        const bool first = stop & !done;   // at most one step of the ray's pass is its first stop
        stop_off = first ? off[u] : stop_off;
        hit = hit | (first & inside[u] & !zout);
        done |= stop;
    }
    res = hit ? (stop_off >> 2) + 1u : 0u;
    return done;
}

}  // namespace sgpu
