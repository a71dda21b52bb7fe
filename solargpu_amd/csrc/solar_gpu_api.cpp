/* C++ drop-in entry points (solar_gpu_api.h) over the C ABI. */
#include "solar_gpu_api.h"

#include <cstdio>

#include "sgpu.h"

namespace solar_gpu {
namespace {

using Entry = int (*)(const float*, int, int, int, int, const int32_t*, const float*, float, const float*,
                      const float*, const float*, const float*, const float*, float*);

int run(Entry fn, const char* name, const float* grid, int w, int h, int n_total, int n_top, const int* pp,
        const float* wall_z, float maxz, const std::vector<SunSample>& sun, const std::vector<float>& dni,
        const std::vector<float>& dhi, const float* slopes, const float* aspects, float* out) {
    if (sun.size() != SGPU_HOURS_PER_YEAR || dni.size() != SGPU_HOURS_PER_YEAR ||
        dhi.size() != SGPU_HOURS_PER_YEAR) {
        std::fprintf(stderr, "[%s] sun/direct/diffuse series must have %d entries\n", name, SGPU_HOURS_PER_YEAR);
        return SGPU_ERR_ARG;
    }
    std::vector<float> az_alt(2 * SGPU_HOURS_PER_YEAR);
    for (int k = 0; k < SGPU_HOURS_PER_YEAR; ++k) {
        az_alt[2 * k] = std::get<0>(sun[k]);
        az_alt[2 * k + 1] = std::get<1>(sun[k]);
    }
    int st = fn(grid, w, h, n_total, n_top, pp, wall_z, maxz, az_alt.data(), dni.data(), dhi.data(), slopes,
                aspects, out);
    if (st) std::fprintf(stderr, "[%s] failed: %s\n", name, sgpu_last_error());
    return st;
}

}  // namespace

int uncompressed_gpu_irradiance(const float* grid_data, int grid_width, int grid_height, int num_total_voxels,
                                int num_top_voxels, const int* pixel_parent, const float* wall_z, float max_height,
                                const std::vector<SunSample>& sun_positions,
                                const std::vector<float>& direct_irradiance,
                                const std::vector<float>& diffuse_irradiance, const float* slopes,
                                const float* aspects, float* irradiance_out) {
    return run(sgpu_uncompressed_irradiance, "uncompressed", grid_data, grid_width, grid_height, num_total_voxels,
               num_top_voxels, pixel_parent, wall_z, max_height, sun_positions, direct_irradiance,
               diffuse_irradiance, slopes, aspects, irradiance_out);
}

int compressed_gpu_irradiance(const float* grid_data, int grid_width, int grid_height, int num_total_voxels,
                              int num_top_voxels, const int* pixel_parent, const float* wall_z, float max_height,
                              const std::vector<SunSample>& sun_positions, const std::vector<float>& direct_irradiance,
                              const std::vector<float>& diffuse_irradiance, const float* slopes, const float* aspects,
                              float* irradiance_out) {
    return run(sgpu_compressed_irradiance, "compressed", grid_data, grid_width, grid_height, num_total_voxels,
               num_top_voxels, pixel_parent, wall_z, max_height, sun_positions, direct_irradiance,
               diffuse_irradiance, slopes, aspects, irradiance_out);
}

}  // namespace solar_gpu
