/* The hemispherical-viewshed DDA step loop shared by every kernel that traces
 * rays (plain viewshed, surrogate sun-bin trace), so their hits are identical.
 * Operation order of viewshed_kernels.cu:27-58; compiled with
 * -ffp-contract=off and IEEE division. */
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
 * once; the tests then run in step order, so the first exit wins as before. */
constexpr int kDdaUnroll = 2;

/* The walk for one ray, specialised on the direction's hemisphere (UPPER: the ray
 * leaves through maxz, else through z = 0) and on whether 32-bit byte offsets
 * address the grid (OFF32: W (H + 2) <= 2^30).  Exit tests are selects, not
 * branches.  OFF32 steps the column and row byte offsets (4 ix, 4 W iy) as
 * unsigned values: a column or row one or two steps off the grid wraps to a value
 * >= 4 W, resp. >= 4 W H, so one unsigned compare each is the bounds test and their
 * sum is the cell's byte offset. */
template <bool COUNT, bool UPPER, bool OFF32>
__device__ __forceinline__ int dda_walk(const float* __restrict__ grid, int W, unsigned uW, unsigned uH, float maxz,
                                        float gxf, float gyf, float base_z, const DirConst& dc,
                                        unsigned long long& steps) {
    const float ox = gxf + dc.cx * 2.0f;
    const float oy = gyf + dc.cy * 2.0f;
    const float oz = base_z + dc.cz * 2.0f;
    int ix = (int)floorf(ox);
    int iy = (int)floorf(oy);
    float tmx = div_by((dc.cx >= 0.0f) ? ((float)ix + 1.0f - ox) : (ox - (float)ix), dc.ax, dc.rx);
    float tmy = div_by((dc.cy >= 0.0f) ? ((float)iy + 1.0f - oy) : (oy - (float)iy), dc.ay, dc.ry);
    const int sx = (dc.cx >= 0.0f) ? 1 : -1;
    const int sy = (dc.cy >= 0.0f) ? 1 : -1;
    if constexpr (OFF32) {
        unsigned sx4 = (unsigned)(4 * sx), sy4 = (unsigned)sy * 4u * uW;   // wave-uniform
        /* keep the steps in VGPRs: a select between one and a wave-uniform step would
           otherwise copy the step into a VGPR on every iteration */
        asm volatile("" : "+v"(sx4), "+v"(sy4));
        const unsigned lim_x = 4u * uW, lim_y = 4u * uW * uH;
        unsigned cx4 = 4u * (unsigned)ix, ry4 = 4u * uW * (unsigned)iy;
        for (;;) {
            float t[kDdaUnroll], h[kDdaUnroll];
            unsigned off[kDdaUnroll];
            bool inside[kDdaUnroll];
#pragma unroll
            for (int u = 0; u < kDdaUnroll; ++u) {
                const bool xs = tmx < tmy;   // a tie steps in y
                t[u] = xs ? tmx : tmy;
                cx4 = xs ? cx4 + sx4 : cx4;
                ry4 = xs ? ry4 : ry4 + sy4;
                tmx = xs ? tmx + dc.rx : tmx;
                tmy = xs ? tmy : tmy + dc.ry;
                inside[u] = cx4 < lim_x && ry4 < lim_y;
                off[u] = inside[u] ? cx4 + ry4 : 0u;   // off-grid steps read a valid cell, unused
                h[u] = *(const float*)((const char*)grid + off[u]);
            }
            unsigned res = 0;
            bool done = false;
#pragma unroll
            for (int u = 0; u < kDdaUnroll; ++u) {
                if (COUNT) steps += done ? 0 : 1;
                const float zr = oz + dc.cz * t[u];
                const bool zout = UPPER ? (zr > maxz) : (zr < 0.0f);
                const bool stop = !inside[u] | zout | (h[u] - zr > 0.01f);
                res = (stop & !done) ? ((inside[u] & !zout) ? (off[u] >> 2) + 1u : 0u) : res;
                done |= stop;
            }
            if (done) return (int)res;
        }
    }
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
        int res = 0;
        bool done = false;
#pragma unroll
        for (int u = 0; u < kDdaUnroll; ++u) {
            if (COUNT) steps += done ? 0 : 1;
            const float zr = oz + dc.cz * t[u];
            const bool zout = UPPER ? (zr > maxz) : (zr < 0.0f);
            const bool stop = !inside[u] | zout | (h[u] - zr > 0.01f);
            res = (stop & !done) ? ((inside[u] & !zout) ? cell[u] + 1 : 0) : res;
            done |= stop;
        }
        if (done) return res;
    }
}

/* One ray of a pooled walk (lanes of a wave on different rays): dda_walk's OFF32 state
 * with the direction's step signs and hemisphere as per-lane values. */
struct PoolRay {
    unsigned cx4, ry4, sx4, sy4;   // column / row byte offsets and their steps
    float tmx, tmy, oz, cz, rx, ry;
    float zhi, zlo;                // leaves above zhi (upper) or below zlo (lower)
};

/* dda_walk's set-up of the ray from integer cell (gx, gy) at base_z along dc. */
__device__ __forceinline__ void pool_ray_setup(PoolRay& r, const DirConst& dc, float gxf, float gyf, float base_z,
                                               unsigned uW, float maxz) {
    const float ox = gxf + dc.cx * 2.0f;
    const float oy = gyf + dc.cy * 2.0f;
    r.oz = base_z + dc.cz * 2.0f;
    const int ix = (int)floorf(ox);
    const int iy = (int)floorf(oy);
    r.tmx = div_by((dc.cx >= 0.0f) ? ((float)ix + 1.0f - ox) : (ox - (float)ix), dc.ax, dc.rx);
    r.tmy = div_by((dc.cy >= 0.0f) ? ((float)iy + 1.0f - oy) : (oy - (float)iy), dc.ay, dc.ry);
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

/* kDdaUnroll steps of a pooled ray (dda_walk's OFF32 steps and tests); true when the ray has
 * ended, with *res the hit cell id + 1 or 0. */
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
        off[u] = inside[u] ? r.cx4 + r.ry4 : 0u;   // off-grid steps read a valid cell, unused
        h[u] = *(const float*)((const char*)grid + off[u]);
    }
    unsigned out = 0;
    bool done = false;
#pragma unroll
    for (int u = 0; u < kDdaUnroll; ++u) {
        if (COUNT) steps += done ? 0 : 1;
        const float zr = r.oz + r.cz * t[u];
        const bool zout = (zr > r.zhi) | (zr < r.zlo);
        const bool stop = !inside[u] | zout | (h[u] - zr > 0.01f);
        out = (stop & !done) ? ((inside[u] & !zout) ? (off[u] >> 2) + 1u : 0u) : out;
        done |= stop;
    }
    res = out;
    return done;
}

/* First hit (cell id + 1) of the ray from integer cell (gx, gy) at height
 * base_z in direction dc, or 0.  `steps` counts DDA steps when COUNT.  The
 * direction is the same for every lane of a wave, so the dispatch is uniform. */
template <bool COUNT>
__device__ __forceinline__ int dda_first_hit(const float* __restrict__ grid, int W, unsigned uW, unsigned uH,
                                             float maxz, float gxf, float gyf, float base_z, const DirConst& dc,
                                             unsigned long long& steps) {
    const bool off32 = grid_off32(uW, uH);
    if (dc.info & 1) {
        return off32 ? dda_walk<COUNT, true, true>(grid, W, uW, uH, maxz, gxf, gyf, base_z, dc, steps)
                     : dda_walk<COUNT, true, false>(grid, W, uW, uH, maxz, gxf, gyf, base_z, dc, steps);
    }
    return off32 ? dda_walk<COUNT, false, true>(grid, W, uW, uH, maxz, gxf, gyf, base_z, dc, steps)
                 : dda_walk<COUNT, false, false>(grid, W, uW, uH, maxz, gxf, gyf, base_z, dc, steps);
}

}  // namespace sgpu
