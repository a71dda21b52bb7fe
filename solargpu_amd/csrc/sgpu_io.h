/* Native I/O of the CLI contract: GeoTIFF, TMY series, voxel dump. */
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "sgpu_internal.h"

namespace sgpu {

struct GeoTiff {
    int width = 0, height = 0;
    int compression = 1;
    int rows_per_strip = 0;
    std::vector<double> pixel_scale;   // 33550
    std::vector<double> tiepoints;     // 33922
    std::vector<uint16_t> geo_keys;    // 34735
    std::vector<double> geo_doubles;   // 34736
    std::string geo_ascii;             // 34737
    std::string nodata;                // 42113 (read only; not written, as in the reference)
    std::vector<float> data;
};

GeoTiff read_geotiff(const std::string& path);
void write_geotiff(const std::string& path, const float* data, int w, int h, const GeoTiff& meta);
std::vector<float> read_tmy(const std::string& path, size_t n);
void write_voxel_dump(const std::string& path, int n_total, int n_top, int w, const int32_t* pixel_parent,
                      const float* grid, const float* wall_z, const float* irradiance);

}  // namespace sgpu
