/* solar_gpu_api.h -- drop-in C++ entry points of the SolarGPU pipeline on
 * MI355X.  The declarations match the SolarGPU reference's
 * include/solar_gpu_api.h:37-54 exactly; both are implemented over the C ABI
 * in sgpu.h (sgpu_uncompressed_irradiance / sgpu_compressed_irradiance).
 *
 *   grid_data          heightmap, grid_width * grid_height floats, row-major
 *   num_total_voxels   top + wall voxels; num_top_voxels == grid_width * grid_height,
 *                      wall voxels are indices [num_top_voxels, num_total_voxels)
 *   pixel_parent       per-voxel top-pixel index (num_total_voxels)
 *   wall_z             per-wall-voxel height (num_total - num_top), may be nullptr without walls
 *   max_height         max height in the grid (upward ray termination)
 *   sun_positions      (azimuth_deg, altitude_deg), 8760 entries
 *   direct_irradiance, diffuse_irradiance   TMY W/m^2, 8760 entries each
 *   slopes, aspects    per-voxel radians (num_total_voxels)
 *   irradiance_out     [out] annual W*h/m^2 per voxel (num_total_voxels)
 *
 * Returns 0 on success and a nonzero sgpu_status on failure (the reference
 * always returns 0 and exits the process on errors; this never exits).
 * uncompressed_gpu_irradiance keeps CSR neighbour lists; compressed_gpu_irradiance
 * keeps ba-RLE streams decoded inside the reflection kernel. */
#pragma once

#include <vector>

#include "sun_position.h"

namespace solar_gpu {

int uncompressed_gpu_irradiance(const float* grid_data, int grid_width, int grid_height,
    int num_total_voxels, int num_top_voxels,
    const int* pixel_parent, const float* wall_z, float max_height,
    const std::vector<SunSample>& sun_positions,
    const std::vector<float>& direct_irradiance, const std::vector<float>& diffuse_irradiance,
    const float* slopes, const float* aspects, float* irradiance_out);

int compressed_gpu_irradiance(const float* grid_data, int grid_width, int grid_height,
    int num_total_voxels, int num_top_voxels,
    const int* pixel_parent, const float* wall_z, float max_height,
    const std::vector<SunSample>& sun_positions,
    const std::vector<float>& direct_irradiance, const std::vector<float>& diffuse_irradiance,
    const float* slopes, const float* aspects, float* irradiance_out);

}  // namespace solar_gpu
