/* Surrogate path kernels on gfx950 (no reference counterpart):
 *   k_sun_svf_pool  the SVF bitmap from the ~265 sun-bin directions only (the
 *                plain tracer's pooled DDA, so the bits are identical);
 *   k_annual_dd  exact annual direct + diffuse per voxel (FP32 Kahan over all
 *                sunlit hours) and the diffuse part alone -- the solar kernel
 *                of solar_kernel.cu:54-77 without the reflection term;
 *   k_horizon_pool  32-sector horizon march per voxel, pooled over voxels and sectors;
 *   k_features   the 53 features of the MLP from the march results, bf16;
 *   k_gate       surrogate estimate, predicted error, fallback flags;
 *   k_combine    scatter of the exact fallback totals. */
#include <cmath>

#include "dda.h"
#include "sgpu_internal.h"
#include "sgpu_surrogate.h"

namespace sgpu {
namespace {

__device__ __forceinline__ uint16_t f2bf(float x) {
    const uint32_t u = __float_as_uint(x);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

__device__ __forceinline__ void voxel_origin(const SurArgs& a, int64_t v, int& top, int& gx, int& gy, float& z) {
    top = a.pix_parent[v];
    gx = top % a.W;
    gy = top / a.W;
    z = (v < a.n_top) ? a.grid[top] : a.wall_z[v - a.n_top];
}

/* Sun-bin SVF on the pooled walk (dda.h): each wave owns the n_sun_dirs x 64 rays
 * of its 64 voxels, direction-major; a lane whose ray has ended takes the next one once
 * kSunRefill lanes are idle (or none is busy).  A hit sets the voxel's bit in LDS words (bit d
 * of word d / 32 is SVF byte d / 8, bit d % 8); the bytes and the hit count are written at
 * the end.  The rays are the plain viewshed's, so the bits equal its SVF. */
constexpr int kSunRefill = 16;

__global__ __launch_bounds__(256) void k_sun_svf_pool(SurArgs a) {
    extern __shared__ unsigned sbits[];   // [4 waves][64 voxels][nw words]
    __shared__ float sorg[4][3][64];
    const int nw = (a.n_sun_dirs + 31) >> 5;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    for (int i = threadIdx.x; i < 4 * 64 * nw; i += 256) sbits[i] = 0u;
    const int nvalid = (int)min<int64_t>(64, max<int64_t>(0, a.n_total - (v - lane)));
    if (lane < nvalid) {
        int top, gx, gy;
        float z;
        voxel_origin(a, v, top, gx, gy, z);
        sorg[wave][0][lane] = (float)gx;
        sorg[wave][1][lane] = (float)gy;
        sorg[wave][2][lane] = z;
    }
    __syncthreads();
    unsigned* wbits = sbits + wave * 64 * nw;
    if (nvalid > 0) {
        const unsigned uW = (unsigned)a.W;
        const unsigned lim_x = 4u * uW, lim_y = 4u * uW * (unsigned)a.H;
        unsigned long long steps = 0;
        bool active = false;
        PoolRay ray{};
        int rd = 0, rvx = 0;   // the lane's ray: sun bin, voxel
        const int total = a.n_sun_dirs * 64;
        int next = 0;          // wave-uniform pool cursor
        for (;;) {
            const uint64_t idle = __ballot(!active);
            const int nidle = __popcll(idle);
            if (next < total && (nidle >= kSunRefill || nidle == 64)) {
                if (!active) {
                    const int j = next + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32),
                                                                       __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                    const int d = j >> 6, vx = j & 63;
                    if (j < total && vx < nvalid) {
                        pool_ray_setup(ray, a.dirs[a.sun_dirs[d]], sorg[wave][0][vx], sorg[wave][1][vx],
                                       sorg[wave][2][vx], uW, a.maxz);
                        rd = d;
                        rvx = vx;
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
                if (pool_ray_steps<false>(ray, a.grid, lim_x, lim_y, steps, res)) {
                    if (res) atomicOr(&wbits[rvx * nw + (rd >> 5)], 1u << (rd & 31));
                    active = false;
                }
            }
        }
    }
    __syncthreads();
    if (lane < nvalid) {
        const unsigned* vb = wbits + lane * nw;
        for (int b = 0; b < a.svf_bytes; ++b) a.svf[(int64_t)b * a.n_total + v] = (uint8_t)(vb[b >> 2] >> (8 * (b & 3)));
        int hits = 0;
        for (int w = 0; w < nw; ++w) hits += __popc(vb[w]);
        a.sun_hits[v] = (uint16_t)hits;
    }
}

constexpr int kDdVoxels = 2;   // voxels per lane in k_annual_dd

__global__ __launch_bounds__(256) void k_annual_dd(SurArgs a) {
    // Each lane sums kDdVoxels voxels 256 apart, so every hour's wave-uniform constants serve
    // several independent Kahan chains; each voxel's sums keep hour order.
    int64_t v[kDdVoxels];
    bool ok[kDdVoxels];
    float ss[kDdVoxels], cs[kDdVoxels], sa[kDdVoxels], ca[kDdVoxels];
    float s[kDdVoxels], c[kDdVoxels], sd[kDdVoxels], cd[kDdVoxels];
#pragma unroll
    for (int i = 0; i < kDdVoxels; ++i) {
        v[i] = (int64_t)blockIdx.x * (256 * kDdVoxels) + i * 256 + threadIdx.x;
        ok[i] = v[i] < a.n_total;
        const int64_t vl = ok[i] ? v[i] : 0;   // a voxel past the end computes voxel 0, unstored
        sincosf(a.slopes[vl], &ss[i], &cs[i]);
        sincosf(a.aspects[vl], &sa[i], &ca[i]);
        v[i] = vl;
        s[i] = c[i] = sd[i] = cd[i] = 0.0f;
    }
    for (int h = 0; h < a.n_hours; ++h) {
        const HourConst& hc = a.hours[h];   // wave-uniform
#pragma unroll
        for (int i = 0; i < kDdVoxels; ++i) {
            const float cdiff = hc.cos_az * ca[i] + hc.sin_az * sa[i];
            const float cos_aoi = hc.cos_z * cs[i] + hc.sin_z * ss[i] * cdiff;
            bool shaded = false;
            if (hc.dense >= 0) shaded = (a.svf[(int64_t)(hc.dense >> 3) * a.n_total + v[i]] >> (hc.dense & 7)) & 1;
            float dir = 0.0f;
            if (cos_aoi > 0.0f && hc.dense >= 0 && !shaded) dir = fmaxf(0.0f, hc.dni_eff * cos_aoi);
            const float aa = fmaxf(0.0f, cos_aoi);
            const float dif =
                fmaxf(0.0f, hc.dhi * ((1 - hc.f1) * ((1 + cs[i]) / 2) + hc.f1 * (aa * hc.inv_b) + hc.f2 * ss[i]));
            const float x = dir + dif;
            float y = x - c[i], t = s[i] + y;
            c[i] = (t - s[i]) - y;
            s[i] = t;
            y = dif - cd[i];
            t = sd[i] + y;
            cd[i] = (t - sd[i]) - y;
            sd[i] = t;
        }
    }
#pragma unroll
    for (int i = 0; i < kDdVoxels; ++i) {
        if (ok[i]) {
            a.dd[v[i]] = s[i];
            a.diff[v[i]] = sd[i];
        }
    }
}

/* Horizon of every (voxel, azimuth sector) of a chunk: march cell by cell from the cell centre
 * along (sin az, -cos az) (azimuth clockwise from north = -y, as the ray directions), tracking
 * the steepest tan(elevation) = (h - z0) / distance (distance in cells, heights in metres, the
 * geometry of the viewshed rays); stop at the grid edge, at 1024 cells or 2048 steps, or when
 * even the highest cell could no longer raise it.  hz gets (best tangent bits, its cell; -1 when
 * no cell was passed).  The marches are pooled like the DDA: each wave owns the 32 x 64 marches
 * of its 64 voxels, sector-major, and a lane whose march has ended takes the next once
 * kSunRefill lanes are idle (or none is busy), so the wave does not wait for its longest march
 * of every sector. */
constexpr int kHorizonUnroll = 8;

__global__ __launch_bounds__(256) void k_horizon_pool(SurArgs a) {
    __shared__ int sgx[4][64], sgy[4][64];
    __shared__ float sz[4][64];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t vwave = a.hz_v0 + (int64_t)blockIdx.x * 256 + wave * 64;   // the wave's first voxel
    const int nvalid = (int)min<int64_t>(64, max<int64_t>(0, a.hz_v0 + a.hz_count - vwave));
    if (lane < nvalid) {
        int top, gx, gy;
        float z;
        voxel_origin(a, vwave + lane, top, gx, gy, z);
        sgx[wave][lane] = gx;
        sgy[wave][lane] = gy;
        sz[wave][lane] = z;
    }
    __syncthreads();
    if (nvalid <= 0) return;
    bool active = false;
    int ix = 0, iy = 0, stx = 0, sty = 0, k = 0, bcell = -1, rvx = 0, rs = 0;
    float tmx = 0.0f, tmy = 0.0f, rx = 0.0f, ry = 0.0f, best = 0.0f, z0 = 0.0f, top = 0.0f;
    const int total = kSectors * 64;
    int next = 0;   // wave-uniform pool cursor
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (next < total && (nidle >= kSunRefill || nidle == 64)) {
            if (!active) {
                const int j = next + (int)__builtin_amdgcn_mbcnt_hi((unsigned)(idle >> 32),
                                                                   __builtin_amdgcn_mbcnt_lo((unsigned)idle, 0u));
                const int sec = j >> 6, vx = j & 63;
                if (j < total && vx < nvalid) {
                    const float az = (sec + 0.5f) * (6.28318531f / kSectors);
                    float sx, cy;
                    sincosf(az, &sx, &cy);
                    const float sy = -cy;
                    const int gx = sgx[wave][vx], gy = sgy[wave][vx];
                    z0 = sz[wave][vx];
                    const float ox = gx + 0.5f, oy = gy + 0.5f;
                    ix = gx;
                    iy = gy;
                    const float ax = fmaxf(fabsf(sx), 1e-8f), ay = fmaxf(fabsf(sy), 1e-8f);
                    tmx = ((sx >= 0.0f) ? (ix + 1.0f - ox) : (ox - ix)) / ax;
                    tmy = ((sy >= 0.0f) ? (iy + 1.0f - oy) : (oy - iy)) / ay;
                    rx = 1.0f / ax;
                    ry = 1.0f / ay;
                    stx = (sx >= 0.0f) ? 1 : -1;
                    sty = (sy >= 0.0f) ? 1 : -1;
                    best = -1e30f;
                    bcell = -1;
                    top = a.maxz - z0;
                    k = 0;
                    rvx = vx;
                    rs = sec;
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
            // kHorizonUnroll steps per pool turn: the cells depend only on the traversal, so the
            // heights of all of them are fetched before the first is tested (an off-grid or
            // too-far step fetches cell 0, unused); the tests then run in step order, so the
            // result equals that of one step per turn.
            int cells[kHorizonUnroll];
            float ts[kHorizonUnroll], hs[kHorizonUnroll];
            bool off[kHorizonUnroll];
#pragma unroll
            for (int u = 0; u < kHorizonUnroll; ++u) {
                if (tmx < tmy) {
                    ts[u] = tmx;
                    ix += stx;
                    tmx += rx;
                } else {
                    ts[u] = tmy;
                    iy += sty;
                    tmy += ry;
                }
                off[u] = (unsigned)ix >= (unsigned)a.W || (unsigned)iy >= (unsigned)a.H || ts[u] > 1024.0f;
                cells[u] = off[u] ? 0 : ix + iy * a.W;
            }
#pragma unroll
            for (int u = 0; u < kHorizonUnroll; ++u) hs[u] = a.grid[cells[u]];
            bool end = false;
#pragma unroll
            for (int u = 0; u < kHorizonUnroll; ++u) {
                if (!end) {
                    if (off[u]) {
                        end = true;
                    } else {
                        const float tn = (hs[u] - z0) / fmaxf(ts[u], 0.5f);
                        if (tn > best) {
                            best = tn;
                            bcell = cells[u];
                        }
                        end = top <= best * ts[u];
                    }
                    if (++k >= 2048) end = true;
                }
            }
            if (end) {
                a.hz[(vwave - a.hz_v0 + rvx) * kSectors + rs] = make_int2(__float_as_int(best), bcell);
                active = false;
            }
        }
    }
}

__global__ __launch_bounds__(256) void k_features(SurArgs a) {
    const int64_t v = a.hz_v0 + (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= a.hz_v0 + a.hz_count) return;
    int top, gx, gy;
    float z;
    voxel_origin(a, v, top, gx, gy, z);
    float ss, cs, sa, ca;
    sincosf(a.slopes[v], &ss, &cs);
    sincosf(a.aspects[v], &sa, &ca);
    uint16_t* row = a.X + v * kSurInPad;
    float svf_h = 0.0f;
    float quad[4] = {0.f, 0.f, 0.f, 0.f};
    float lamb = 0.0f;   // Lambertian estimate from the horizon cells (weights of the reflection pass)
    const float aspect = a.aspects[v];
    const float nx = ss * sa, ny = -ss * ca, nz = cs;
    for (int s = 0; s < kSectors; ++s) {
        const float az = (s + 0.5f) * (6.28318531f / kSectors);
        float sx, cy;
        sincosf(az, &sx, &cy);
        const int2 hit = a.hz[(v - a.hz_v0) * kSectors + s];   // the sector's march (k_horizon_pool)
        const float best = __int_as_float(hit.x);
        const int cell = hit.y;
        const float ang = (cell >= 0) ? atanf(best) : 0.0f;
        row[s] = f2bf(ang * 0.63661977f);
        const float se = sinf(fmaxf(ang, 0.0f));
        svf_h += se * se;
        if (cell >= 0 && ang > 0.0f) {
            /* quadrant of the sector relative to the facing direction: 0 front, 1 right, 2 back, 3 left */
            float rel = az - aspect;
            rel -= 6.28318531f * floorf(rel * 0.15915494f);
            const int q = (int)((rel + 0.78539816f) * 0.63661977f) & 3;
            quad[q] += a.dd[cell] * se;
        }
        if (cell >= 0) {   /* 0.2 cos_recv cos_send / (pi d^2), dx dy in cells, dz in metres */
            const float dx = (float)(cell % a.W - gx), dy = (float)(cell / a.W - gy), dz = a.grid[cell] - z;
            const float d2 = dx * dx + dy * dy + dz * dz;
            if (d2 >= 1e-4f) {
                const float inv_d = rsqrtf(d2);
                const float cr = (nx * dx + ny * dy + nz * dz) * inv_d;
                if (cr > 0.0f) lamb += 0.0636619772f * cr * fmaxf(fabsf(dz) * inv_d, 0.1f) * inv_d * inv_d * a.dd[cell];
            }
        }
    }
    const bool wall = v >= a.n_top;
    const float dd = a.dd[v];
    row[32] = f2bf(ss * sa);
    row[33] = f2bf(-ss * ca);
    row[34] = f2bf(cs);
    row[35] = f2bf((a.maxz - z) * 0.02f);
    row[36] = f2bf(log1pf((float)a.sun_hits[v]) * 0.179f);
    row[37] = f2bf(1.0f - svf_h / kSectors);
    row[38] = f2bf(a.lat / 90.0f);
    row[39] = f2bf(a.albedo);
    row[40] = f2bf(wall ? 1.0f : 0.0f);
    row[41] = f2bf(a.slopes[v] * 0.63661977f);
    row[42] = f2bf(dd * 1e-6f);
    row[43] = f2bf(dd > 0.0f ? a.diff[v] / dd : 0.0f);
#pragma unroll
    for (int q = 0; q < 4; ++q) row[44 + q] = f2bf(quad[q] * (1e-6f / 8.0f));
    row[48] = f2bf(a.cell_m);   // horizontal distances of the ray geometry are in cells
    row[49] = f2bf(lamb * 1e-5f);
    /* the ground in front of a wall (one cell along its outward normal) and the height above it */
    float front_dd = 0.0f, front_h = 0.0f;
    if (wall) {
        const int fx = min(max(gx + (int)rintf(sa), 0), a.W - 1), fy = min(max(gy - (int)rintf(ca), 0), a.H - 1);
        const int fc = fx + fy * a.W;
        front_dd = a.dd[fc];
        front_h = z - a.grid[fc];
    }
    row[50] = f2bf(front_dd * 1e-6f);
    row[51] = f2bf(front_h * 0.02f);
    /* mean annual direct + diffuse of the 8 cells around the voxel's cell */
    float ring = 0.0f;
    int nring = 0;
    for (int oy = -1; oy <= 1; ++oy)
        for (int ox = -1; ox <= 1; ++ox) {
            const int cx = gx + ox, cy = gy + oy;
            if ((ox || oy) && cx >= 0 && cy >= 0 && cx < a.W && cy < a.H) {
                ring += a.dd[cx + cy * a.W];
                ++nring;
            }
        }
    row[52] = f2bf(nring ? ring / (float)nring * 1e-6f : 0.0f);
#pragma unroll
    for (int f = kSurFeatures; f < kSurInPad; ++f) row[f] = 0;
}

/* Surrogate total = exact direct + diffuse + predicted reflected share x (direct + diffuse).
 * Predicted relative error of that total: with an error network, 10^(its first output), a
 * regression of the log10 per-voxel error the surrogate makes on held-out tiles; without one,
 * a score -- the calibrated relative error of the reflected share times the predicted
 * reflected fraction of the total, plus the miss on the diffuse share (known exactly here; a
 * large miss marks an input the network does not represent).  Voxels whose predicted error
 * exceeds tau go to the exact path. */
__global__ __launch_bounds__(256) void k_gate(SurArgs a) {
    const int64_t v = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (v >= a.n_total) return;
    const float dd = a.dd[v];
    const float share = fmaxf(a.Y[2 * v], 0.0f);
    const float refl = share * dd;
    const float total = dd + refl;
    float err;
    if (a.Yerr != nullptr) {
        err = exp10f(fminf(a.Yerr[2 * v], 30.0f));
    } else {
        const float exact_dshare = dd > 0.0f ? a.diff[v] / dd : 0.0f;
        const float dmiss = fabsf(a.Y[2 * v + 1] - exact_dshare);
        const float frac = total > 0.0f ? refl / total : 0.0f;
        err = a.refl_rel_err * frac + a.diff_miss_weight * dmiss + a.share_abs_err * dd / fmaxf(total, 1e-3f);
    }
    a.out[v] = total;
    a.fallback[v] = (!(err <= a.tau) || !(total > 0.0f)) ? 1 : 0;
}

__global__ __launch_bounds__(256) void k_scatter(const int32_t* ids, const float* vals, int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[ids[i]] = vals[i];
}

/* out[ids[i]] = base[ids[i]] + vals[i] (exact direct + diffuse plus the fallback's reflected sum) */
__global__ __launch_bounds__(256) void k_scatter_sum(const int32_t* ids, const float* vals, const float* base,
                                                     int64_t n, float* out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) {
        const int32_t v = ids[i];
        out[v] = vals[i] + base[v];
    }
}

/* flags by position p: flags of voxel order[p] (order NULL: p itself) */
__global__ __launch_bounds__(256) void k_permute_flags(const int32_t* flags, const int32_t* order, int64_t n,
                                                       int32_t* pflags) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n) pflags[p] = flags[order ? (int64_t)order[p] : p];
}

/* ids[off[p]] = the voxel at position p for every flagged position */
__global__ __launch_bounds__(256) void k_compact(const int32_t* pflags, const int32_t* order, const int64_t* off,
                                                 int64_t n, int32_t* ids) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p < n && pflags[p]) ids[off[p]] = order ? order[p] : (int32_t)p;
}

}  // namespace

void launch_sun_svf(const SurArgs& a, hipStream_t s) {
    const dim3 g((unsigned)((a.n_total + 255) / 256));
    if (!grid_off32((unsigned)a.W, (unsigned)a.H)) fail(SGPU_ERR_ARG, "grid beyond 32-bit byte offsets");
    if (a.n_sun_dirs > kAzBins * kElBins) fail(SGPU_ERR_ARG, "more sun bins than upper directions");
    const size_t lds = (size_t)4 * 64 * ((a.n_sun_dirs + 31) / 32) * sizeof(unsigned);
    hipLaunchKernelGGL(k_sun_svf_pool, g, dim3(256), lds, s, a);
    SGPU_HIP(hipGetLastError());
}
void launch_annual_dd(const SurArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_annual_dd, dim3((unsigned)((a.n_total + 256 * kDdVoxels - 1) / (256 * kDdVoxels))), dim3(256), 0,
                       s, a);
    SGPU_HIP(hipGetLastError());
}
void launch_features(const SurArgs& a, int2* hz, int64_t hz_cap, hipStream_t s) {
    if (hz_cap <= 0) fail(SGPU_ERR_ARG, "horizon buffer missing");
    for (int64_t v0 = 0; v0 < a.n_total; v0 += hz_cap) {
        SurArgs c = a;
        c.hz = hz;
        c.hz_v0 = v0;
        c.hz_count = std::min<int64_t>(hz_cap, a.n_total - v0);
        const dim3 g((unsigned)((c.hz_count + 255) / 256));
        hipLaunchKernelGGL(k_horizon_pool, g, dim3(256), 0, s, c);
        hipLaunchKernelGGL(k_features, g, dim3(256), 0, s, c);
        SGPU_HIP(hipGetLastError());
    }
}
void launch_gate(const SurArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_gate, dim3((unsigned)((a.n_total + 255) / 256)), dim3(256), 0, s, a);
    SGPU_HIP(hipGetLastError());
}
void launch_scatter(const int32_t* ids, const float* vals, int64_t n, float* out, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ids, vals, n, out);
    SGPU_HIP(hipGetLastError());
}
void launch_scatter_sum(const int32_t* ids, const float* vals, const float* base, int64_t n, float* out,
                        hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter_sum, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, ids, vals, base, n, out);
    SGPU_HIP(hipGetLastError());
}
void launch_permute_flags(const int32_t* flags, const int32_t* order, int64_t n, int32_t* pflags, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_permute_flags, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, flags, order, n, pflags);
    SGPU_HIP(hipGetLastError());
}
void launch_compact(const int32_t* pflags, const int32_t* order, const int64_t* off, int64_t n, int32_t* ids,
                    hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_compact, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, pflags, order, off, n, ids);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
