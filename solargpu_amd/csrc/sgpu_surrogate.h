/* Surrogate path: kernel arguments and launchers (k_surrogate.hip). */
#pragma once

#include <cstdint>

#include "sgpu_internal.h"

namespace sgpu {

constexpr int kSectors = 32;
constexpr int kSurFeatures = 53;
constexpr int kSurInPad = 64;

struct SurArgs {
    const float* grid;
    int W, H;
    int64_t n_total, n_top;
    const int32_t* pix_parent;
    const float* wall_z;
    const float* slopes;
    const float* aspects;
    float maxz;
    const DirConst* dirs;
    const int16_t* sun_dirs;   // direction index of every dense sun bin
    int n_sun_dirs;
    int svf_bytes;
    uint8_t* svf;
    uint16_t* sun_hits;
    const HourConst* hours;
    int n_hours;
    float* dd;                 // annual direct + diffuse
    float* diff;               // annual diffuse
    uint16_t* X;               // features [n_total][64] bf16
    float lat, albedo, cell_m;
    const float* Y;            // MLP output [n_total][2]
    const float* Yerr;         // error network output [n_total][2] (log10 of the predicted error), or null
    float* out;                // surrogate annual total
    int32_t* fallback;         // 1 = take the exact plain result
    float tau, refl_rel_err, diff_miss_weight, share_abs_err;
    /* horizon march results of voxels [hz_v0, hz_v0 + hz_count): (best tangent bits, cell) per
       sector, set by launch_features per chunk */
    int2* hz;
    int64_t hz_v0, hz_count;
};

void launch_sun_svf(const SurArgs& a, hipStream_t s);
void launch_annual_dd(const SurArgs& a, hipStream_t s);
/* The 32-sector horizon march (pooled over voxels and sectors) and the feature rows, over chunks
   of at most hz_cap voxels whose march results go to hz (hz_cap x 32 entries). */
void launch_features(const SurArgs& a, int2* hz, int64_t hz_cap, hipStream_t s);
void launch_gate(const SurArgs& a, hipStream_t s);
void launch_scatter(const int32_t* ids, const float* vals, int64_t n, float* out, hipStream_t s);
/* ids[off[v]] = v for every flagged v (off = exclusive scan of the flags). */
void launch_scatter_sum(const int32_t* ids, const float* vals, const float* base, int64_t n, float* out,
                        hipStream_t s);
/* The flagged voxels in processing order (order NULL: voxel order): pflags[p] = flags[order[p]];
   after the scan of pflags into off, ids[off[p]] = order[p] for every flagged position. */
void launch_permute_flags(const int32_t* flags, const int32_t* order, int64_t n, int32_t* pflags, hipStream_t s);
void launch_compact(const int32_t* pflags, const int32_t* order, const int64_t* off, int64_t n, int32_t* ids,
                    hipStream_t s);

}  // namespace sgpu
