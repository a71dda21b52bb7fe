/* Stage 1's trace by step tables on gfx950: the hemispherical DDA of viewshed_kernels.cu:27-58
 * for the voxels of one traversal class at a time (class_key in sgpu_internal.h), with the same
 * operation order and the same results as the pooled walk of dda.h.
 *
 * k_table_records one record per chunk position of the class-sorted order (cell, base z, raw row).
 * k_trace_table   one workgroup per (G-direction group, item of up to kTableItemVoxels
 *                 class-sorted voxels of one class).  Lanes 0..G-1 of wave 0 first walk the
 *                 class's representative ray of each direction for S steps and keep, per step,
 *                 the cell's byte offset from the start cell and RN(cz t), and per axis the step
 *                 of its i-th move (so a ray's exit through the grid edge is one lookup at
 *                 set-up).  Every ray of the class along that direction takes the same steps from
 *                 its own start cell, so a ray step is then: the table entry, the cell's height,
 *                 z = oz + RN(cz t) and the three exit tests.  Quads of 4 lanes walk 4 consecutive
 *                 voxels of the class-sorted order (adjacent cells of a row, or the stacked wall
 *                 voxels of a cell) along one direction in lockstep, so their height loads share
 *                 a cache line; a quad takes the wave's next (direction, quad of voxels) job once
 *                 kRefillQuads quads are idle.  Rays still going after S steps are appended to a
 *                 pending list.
 * k_trace_pending the pooled walk of dda.h over every direction of the irregular voxels (cells
 *                 near a binade end) and the pending rays, these from the class's state after
 *                 the table's S steps. */
#include "dda.h"
#include "sgpu_internal.h"

namespace sgpu {
namespace {

constexpr int kTableBlock = 256;
constexpr int kQuads = 16;          // lockstep quads of a wave: 4 lanes, 4 consecutive voxels, one direction
/* Steps a ray may skip at once: the next kSkipSteps cells lie in the (kSkipSteps + 1)^2 cells ahead of
   its current cell, whose highest height the quadrant's max map holds. */
constexpr int kSkipSteps = kMaxMapWindow - 1;
constexpr int kRefillQuads = 4;     // refill once this many quads are idle (or all of them)
constexpr int kPendRefill = 16;
constexpr int kPendPerWave = 2048;  // pending-walk rays per wave

/* Per-direction constants of a table: the start cell's byte offset from the origin cell, and the
   terms of the grid-edge lookups: a ray from cell (gx, gy) can take lx = mx * gx + cx x moves
   (and likewise y moves) before it leaves the grid, plus one (the lookup index, clamped). */
struct TabDir {
    int dbase;
    int mx, cx, my;
    int cy;
    float c2z;        // RN(2 cz): the origin's z offset
    float zhi, zlo;   // a ray leaves above zhi or below zlo
    int quad;         // the ray's quadrant: bit 0 toward +x, bit 1 toward +y (its max map)
    int lowest;       // the step of a skip window whose z is lowest: 0 (rising ray) or kSkipSteps - 1
    int pad[2];
};

template <int G, int S>
struct TableLds {
    int2 ent[G * S];         // direction k, step s + 1 at k * S + s: (byte offset of its cell from the start cell, bits of RN(cz t))
    uint16_t xi[G][S + 2];   // xi[0] = 1; xi[i] = the step of the i-th x move; S + 1 past the table
    uint16_t yi[G][S + 2];
    TabDir td[G];
    DirConst dir[G];
};

/* The tables of the group's G directions for class `key` (lanes k < G of one wave): the
   representative origin cell (2^ex + 2, 2^ey + 2) lies inside its class, and the walk is
   pool_ray_steps's traversal (tie steps in y). */
template <int G, int S>
__device__ void build_tables(TableLds<G, S>& L, int key, int W, int H, float maxz, TableResume* resume) {
    const int k = threadIdx.x;
    if (k >= G) return;
    const DirConst dc = L.dir[k];
    const int gxr = (1 << (3 + (key >> 4))) + 2, gyr = (1 << (3 + (key & 15))) + 2;
    int ix, iy;
    float tmx, tmy;
    ray_start(dc, (float)gxr, (float)gyr, ix, iy, tmx, tmy);
    const int dix = ix - gxr, diy = iy - gyr;
    const bool px = dc.cx >= 0.0f, py = dc.cy >= 0.0f;
    const bool upper = dc.info & 1;
    /* x moves left inside: W - 1 - (gx + dix) toward +x, gx + dix toward -x; + 1 for the lookup */
    L.td[k] = TabDir{4 * dix + 4 * W * diy, px ? -1 : 1, px ? W - dix : dix + 1, py ? -1 : 1, py ? H - diy : diy + 1,
                     dc.cz * 2.0f, upper ? maxz : INFINITY, upper ? -INFINITY : 0.0f, (px ? 1 : 0) | (py ? 2 : 0),
                     dc.cz >= 0.0f ? 0 : kSkipSteps - 1, {0, 0}};
    const int sx4 = px ? 4 : -4;
    const int sy4 = py ? 4 * W : -4 * W;
    int mx = 0, my = 0, off = 0;
    L.xi[k][0] = 1;
    L.yi[k][0] = 1;
    for (int s = 1; s <= S; ++s) {
        const bool xs = tmx < tmy;
        const float t = xs ? tmx : tmy;
        if (xs) {
            ++mx;
            tmx = tmx + dc.rx;
            off += sx4;
            L.xi[k][mx] = (uint16_t)s;
        } else {
            ++my;
            tmy = tmy + dc.ry;
            off += sy4;
            L.yi[k][my] = (uint16_t)s;
        }
        L.ent[k * S + s - 1] = make_int2(off, __float_as_int(dc.cz * t));
    }
    for (int i = mx + 1; i <= S + 1; ++i) L.xi[k][i] = (uint16_t)(S + 1);
    for (int i = my + 1; i <= S + 1; ++i) L.yi[k][i] = (uint16_t)(S + 1);
    /* every workgroup of this class and group writes the same values */
    resume[k] = TableResume{tmx, tmy, mx, my};
}

__device__ __forceinline__ int mbcnt64(uint64_t m) {
    return (int)__builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

/* The quadrant max maps, in two separable passes: along x, the highest of cells x, x + sx, ...,
   x + (kMaxMapWindow - 1) sx of the row (sx = -1 for quadrants 0 and 2, +1 for 1 and 3); then along y
   likewise over those maxima.  Cells off the grid do not count. */
__global__ __launch_bounds__(256) void k_max_map_x(const float* grid, int W, int H, float* tmp) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t cells = (int64_t)W * H;
    if (i >= 2 * cells) return;
    const int sx = (i < cells) ? -1 : 1;
    const int64_t c = i % cells;
    const int x = (int)(c % W);
    float m = -INFINITY;
    for (int d = 0; d < kMaxMapWindow; ++d) {
        const int xx = x + sx * d;
        if (xx >= 0 && xx < W) m = fmaxf(m, grid[c - x + xx]);
    }
    tmp[i] = m;
}
__global__ __launch_bounds__(256) void k_max_map_y(const float* tmp, int W, int H, float* qmax) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    const int64_t cells = (int64_t)W * H;
    if (i >= 4 * cells) return;
    const int q = (int)(i / cells);
    const int64_t c = i % cells;
    const float* src = tmp + ((q & 1) ? cells : 0);
    const int sy = (q & 2) ? 1 : -1;
    const int y = (int)(c / W);
    float m = -INFINITY;
    for (int d = 0; d < kMaxMapWindow; ++d) {
        const int yy = y + sy * d;
        if (yy >= 0 && yy < H) m = fmaxf(m, src[c + (int64_t)(yy - y) * W]);
    }
    qmax[i] = m;
}

/* One voxel of the class-sorted order: its cell (gx | gy << 16; grids below 2^16 a side), base
   height and raw row. */
__global__ __launch_bounds__(256) void k_table_records(const int32_t* perm, int64_t v0, int64_t n,
                                                       const int32_t* pix_parent, const float* grid,
                                                       const float* wall_z, int64_t n_top, int W, int4* rec) {
    const int64_t p = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= n) return;
    const int64_t v = perm[v0 + p];
    const int top = pix_parent[v];
    const float bz = (v < n_top) ? grid[top] : wall_z[v - n_top];
    rec[p] = make_int4((top % W) | ((top / W) << 16), __float_as_int(bz), (int)(v - v0), 0);
}

template <bool COUNT, int G, int S>
__global__ __launch_bounds__(kTableBlock) void k_trace_table(TableArgs a) {
    static_assert(S % 3 == 0, "the walk takes three steps at a time");
    __shared__ TableLds<G, S> L;
    __shared__ int4 srec[kTableBlock / 64][64];   // each wave's current set of 64 voxel records
    const int grp = blockIdx.x;
    const TableItem it = a.items[blockIdx.y];
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6)), lane = threadIdx.x & 63;
    if (threadIdx.x < G) L.dir[threadIdx.x] = a.dirs[grp * G + threadIdx.x];
    __syncthreads();
    if (wave == 0) build_tables<G, S>(L, it.key, a.W, a.H, a.maxz, a.resume + (int64_t)it.key * kDirs + grp * G);
    __syncthreads();

    const int W = a.W;
    const char* gridb = (const char*)a.grid;
    /* the wave's sets of 64 voxels: wave, wave + 4, ...; jobs (direction k, quad q of the set),
       direction-major */
    const int nsets = (it.count + 63) >> 6;
    const int my_sets = nsets > wave ? (nsets - wave + 3) >> 2 : 0;
    constexpr int kJobsPerSet = kQuads * G;
    const int total = my_sets * kJobsPerSet;
    const int64_t rowbase = (int64_t)grp * G;
    int next = 0, cur = -1;
    bool active = false, push = false;
    int k = 0, ti = 0, rem = 0, prow = 0;   // direction, table index k * S + s, steps left inside the grid - 1
    unsigned base = 0;                      // byte offset of the start cell (unsigned: grids reach 2^32 bytes)
    unsigned pos = 0;                       // byte offset of the cell of the ray's last step
    float oz = 0.0f, zhi = 0.0f, zlo = 0.0f;
    int32_t* out = nullptr;
    unsigned long long steps = 0;
    for (;;) {
        /* rays that outlived the table: one slot each (lanes that add 0 keep the counter) */
        if (__ballot(push)) {
            const unsigned long long slot = atomicAdd(a.pend_count, push ? 1ull : 0ull);
            if (push && slot < (unsigned long long)a.pend_cap)
                a.pend[slot] = make_uint2((unsigned)prow, (unsigned)(grp * G + k) | ((unsigned)it.key << 12));
            push = false;
        }
        /* idle quads: bit 4p of qi */
        const uint64_t idle = __ballot(!active);
        uint64_t qi = idle & (idle >> 1);
        qi &= qi >> 2;
        qi &= 0x1111111111111111ull;
        const int nq = __popcll(qi);
        if (next < total && (nq >= kRefillQuads || qi == 0x1111111111111111ull)) {
            const int set = next / kJobsPerSet;
            if (set != cur) {   // stage the wave's next 64 voxel records
                cur = set;
                const int p = (wave + 4 * set) * 64 + lane;
                srec[wave][lane] = p < it.count ? a.rec[it.pos + p] : make_int4(0, 0, -1, 0);
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            }
            const int in_set = next - set * kJobsPerSet;
            const int take = min(nq, kJobsPerSet - in_set);
            const int qp = lane & ~3;   // first lane of the lane's quad
            const int rank = __popcll(qi & ((1ull << qp) - 1ull));
            if (((qi >> qp) & 1ull) && rank < take) {
                const int job = in_set + rank;
                const int4 r = srec[wave][(job % kQuads) * 4 + (lane & 3)];
                if (r.z >= 0) {
                    k = job / kQuads;
                    const TabDir t = L.td[k];
                    const int gx = r.x & 0xFFFF, gy = (int)((unsigned)r.x >> 16);
                    const int jx = min(max(t.mx * gx + t.cx, 0), S + 1);
                    const int jy = min(max(t.my * gy + t.cy, 0), S + 1);
                    rem = min((int)L.xi[k][jx], (int)L.yi[k][jy]) - 1;
                    base = 4u * (unsigned)(gx + W * gy) + (unsigned)t.dbase;
                    oz = __int_as_float(r.y) + t.c2z;
                    zhi = t.zhi;
                    zlo = t.zlo;
                    prow = r.z;
                    out = a.raw + (int64_t)r.z * kDirs + rowbase + k;
                    ti = k * S;
                    pos = base;
                    active = true;
                }
            }
            next += take;
        }
        if (__ballot(active) == 0) {
            if (next >= total) break;
            continue;
        }
        if (active) {
            /* the next three steps' heights, and with them (same round trip) the max map of the cells
               ahead: the next kSkipSteps cells cannot stop the ray as a hit when the highest of them is
               no more than 0.01 above its lowest z in those steps (z is monotone along the ray, and so
               is rounding); an exit among them is seen by the next step, which is then off the grid or
               out of z too.  A ray whose last cell is off the grid (rem < 0) or out of z has left: no
               skip, so the next step stops it. */
            int2 e[3];
            float h[3];
            bool ins[3];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                e[u] = L.ent[ti + u];
                ins[u] = rem > u;
                const unsigned off = ins[u] ? base + (unsigned)e[u].x : 0u;
                h[u] = *(const float*)(gridb + off);
            }
            bool skip = false;
            if (!COUNT && a.qmax && ti != k * S && ti + kSkipSteps <= k * S + S && rem >= 0) {
                const TabDir& t = L.td[k];
                const float m = *(const float*)((const char*)(a.qmax + (int64_t)t.quad * a.map_cells) + pos);
                const float zpos = oz + __int_as_float(L.ent[ti - 1].y);
                const float zref = oz + __int_as_float(L.ent[ti + t.lowest].y);
                skip = !(zpos > zhi) && !(zpos < zlo) && !(m - zref > 0.01f);
            }
            if (skip) {
                pos = base + (unsigned)L.ent[ti + kSkipSteps - 1].x;
                ti += kSkipSteps;
                rem -= kSkipSteps;
                if (ti - k * S >= S) {
                    push = true;
                    active = false;
                }
                continue;
            }
            pos = base + (unsigned)e[2].x;   // the last step's cell (inside the grid unless the ray stops)
            /* the first stopping step: its cell's byte offset when it is a hit, else ~0 (never a
               multiple of 4) */
            bool st[3];
            unsigned r[3];
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const float zr = oz + __int_as_float(e[u].y);
                const bool zout = (zr > zhi) | (zr < zlo);
                st[u] = !ins[u] | zout | (h[u] - zr > 0.01f);
                r[u] = (ins[u] & !zout) ? base + (unsigned)e[u].x : ~0u;
            }
            const unsigned res = st[0] ? r[0] : (st[1] ? r[1] : r[2]);
            const bool done = st[0] | st[1] | st[2];
            if (COUNT) steps += st[0] ? 1 : (st[1] ? 2 : 3);
            ti += 3;
            rem -= 3;
            if (done) {
                *out = res == ~0u ? 0 : (int32_t)((res >> 2) + 1u);
                active = false;
            } else if (ti - k * S >= S) {
                push = true;
                active = false;
            }
        }
    }
    if (COUNT) atomicAdd(a.steps, steps);
}

/* The pooled walk over the chunk's remaining rays: ray i < n_irregular * kDirs is direction
   i % kDirs of irregular position i / kDirs, the others are pending rays resumed after the
   table's steps. */
template <bool COUNT>
__global__ __launch_bounds__(256) void k_trace_pending(PendingArgs a) {
    const int64_t n_irr = a.n_irregular * kDirs;
    const int64_t n = n_irr + a.n_pend;
    const int64_t w0 = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * kPendPerWave;
    const int total = (int)max<int64_t>(0, min<int64_t>(kPendPerWave, n - w0));
    const unsigned uW = (unsigned)a.W, uH = (unsigned)a.H;
    const unsigned lim_x = 4u * uW, lim_y = 4u * uW * uH;
    unsigned long long steps = 0;
    bool active = false;
    PoolRay ray{};
    int32_t* out = nullptr;
    int next = 0;
    for (;;) {
        const uint64_t idle = __ballot(!active);
        const int nidle = __popcll(idle);
        if (next < total && (nidle >= kPendRefill || nidle == 64)) {
            if (!active) {
                const int j = next + mbcnt64(idle);
                if (j < total) {
                    const int64_t i = w0 + j;
                    int row, d, key = -1;
                    if (i < n_irr) {
                        row = (int)(a.perm[a.v0 + a.irregular[i / kDirs]] - a.v0);
                        d = (int)(i % kDirs);
                    } else {
                        const uint2 e = a.pend[i - n_irr];
                        row = (int)e.x;
                        d = (int)(e.y & 0xFFFu);
                        key = (int)(e.y >> 12);
                    }
                    const int64_t v = a.v0 + row;
                    const int top = a.pix_parent[v];
                    const float bz = (v < a.n_top) ? a.grid[top] : a.wall_z[v - a.n_top];
                    const DirConst dc = a.dirs[d];
                    pool_ray_setup(ray, dc, (float)(top % a.W), (float)(top / a.W), bz, uW, a.maxz);
                    if (key >= 0) {   // the class's state after the table's steps
                        const TableResume r = a.resume[(int64_t)key * kDirs + d];
                        ray.cx4 += (unsigned)r.mx * ray.sx4;
                        ray.ry4 += (unsigned)r.my * ray.sy4;
                        ray.tmx = r.tmx;
                        ray.tmy = r.tmy;
                    }
                    out = a.raw + (int64_t)row * kDirs + d;
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
            if (pool_ray_steps<COUNT>(ray, a.grid, lim_x, lim_y, steps, res)) {
                *out = (int32_t)res;
                active = false;
            }
        }
    }
    if (COUNT) atomicAdd(a.steps, steps);
}

}  // namespace

void launch_trace_table(const TableArgs& a, int n_items, int dirs, int steps, hipStream_t s) {
    if (n_items <= 0) return;
    if (!grid_off32((unsigned)a.W, (unsigned)a.H)) fail(SGPU_ERR_ARG, "grid beyond 32-bit byte offsets");
    const dim3 block(kTableBlock);
#define SGPU_TABLE_CASE(G, S)                                                                                   \
    if (dirs == G && steps == S) {                                                                           \
        const dim3 grid((unsigned)(kDirs / G), (unsigned)n_items);                                           \
        if (a.steps) hipLaunchKernelGGL((k_trace_table<true, G, S>), grid, block, 0, s, a);                  \
        else hipLaunchKernelGGL((k_trace_table<false, G, S>), grid, block, 0, s, a);                         \
        SGPU_HIP(hipGetLastError());                                                                         \
        return;                                                                                              \
    }
    SGPU_TABLE_CASE(8, 255)
    SGPU_TABLE_CASE(4, 255)
    SGPU_TABLE_CASE(4, 510)
    SGPU_TABLE_CASE(8, 126)
    SGPU_TABLE_CASE(4, 126)
    SGPU_TABLE_CASE(16, 126)
    SGPU_TABLE_CASE(8, 63)
#undef SGPU_TABLE_CASE
    fail(SGPU_ERR_ARG, "step-table shape %d directions x %d steps is not built", dirs, steps);
}

void launch_max_maps(const float* grid, int W, int H, float* tmp, float* qmax, hipStream_t s) {
    const int64_t cells = (int64_t)W * H;
    hipLaunchKernelGGL(k_max_map_x, dim3((unsigned)((2 * cells + 255) / 256)), dim3(256), 0, s, grid, W, H, tmp);
    hipLaunchKernelGGL(k_max_map_y, dim3((unsigned)((4 * cells + 255) / 256)), dim3(256), 0, s, tmp, W, H, qmax);
    SGPU_HIP(hipGetLastError());
}

void launch_table_records(const int32_t* perm, int64_t v0, int64_t n, const int32_t* pix_parent, const float* grid,
                          const float* wall_z, int64_t n_top, int W, int4* rec, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_table_records, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, perm, v0, n, pix_parent,
                       grid, wall_z, n_top, W, rec);
    SGPU_HIP(hipGetLastError());
}

void launch_trace_pending(const PendingArgs& a, hipStream_t s) {
    const int64_t n = a.n_irregular * kDirs + a.n_pend;
    if (n <= 0) return;
    const int64_t per_block = 4 * (int64_t)kPendPerWave;
    const dim3 grid((unsigned)((n + per_block - 1) / per_block));
    if (a.steps) hipLaunchKernelGGL(k_trace_pending<true>, grid, dim3(256), 0, s, a);
    else hipLaunchKernelGGL(k_trace_pending<false>, grid, dim3(256), 0, s, a);
    SGPU_HIP(hipGetLastError());
}

}  // namespace sgpu
