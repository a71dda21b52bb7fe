/* C-ABI shim over the SolarGPU reference's host-only sources, compiled from
 * where they lie (see oracle/Makefile, target `ref`) into oracle/_ref/.
 * TEST INFRASTRUCTURE ONLY: used to write and re-check golden vectors for
 * sun position (sun_position.h:24-107), sky bins (sky_bins.cpp:9-52),
 * voxelization (voxelization.cpp:10-97) and the ba-RLE encoder
 * (barle_compression.cpp:156-207).  The reference's CUDA sources are not
 * buildable here (no nvcc) and are not part of this shim. */
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <vector>

#include "barle_compression.h"
#include "sky_bins.h"
#include "sun_position.h"
#include "voxelization.h"

using namespace solar_gpu;

extern "C" {

void ref_sun_positions(float lon, float lat, int tz, float* out) {
    setenv("TZ", "UTC", 1);  // mktime date normalisation must not see DST
    tzset();
    SunPositionCalculator calc(lon, lat, tz);
    std::vector<SunSample> v;
    calc.calc(v);
    for (size_t k = 0; k < v.size(); ++k) {
        out[2 * k] = std::get<0>(v[k]);
        out[2 * k + 1] = std::get<1>(v[k]);
    }
}

int ref_sky_bins(const float* az_alt, float* nudged, int* f2d, int* sdi, int* hgs) {
    std::vector<SunSample> v(kHoursPerYear);
    for (int k = 0; k < kHoursPerYear; ++k) v[k] = SunSample(az_alt[2 * k], az_alt[2 * k + 1]);
    SkyBins b = compute_sky_bins(v);
    for (int k = 0; k < kHoursPerYear; ++k) {
        nudged[2 * k] = std::get<0>(b.sun_azm_alt[k]);
        nudged[2 * k + 1] = std::get<1>(b.sun_azm_alt[k]);
        sdi[k] = b.sun_dense_idx[k];
    }
    for (size_t i = 0; i < b.full_to_dense.size(); ++i) f2d[i] = b.full_to_dense[i];
    *hgs = (int)b.hgrid_size_s;
    return (int)b.active_bin_count;
}

int64_t ref_detect_walls(const float* H, int w, int h, double sx, double sy, float voxel, int* parent, float* z,
                         float* aspect, int64_t cap) {
    HeightGrid g;
    g.width = w; g.height = h; g.scaleX = sx; g.scaleY = sy;
    g.heights.assign(H, H + (size_t)w * h);
    std::vector<WallVoxel> walls = detect_wall_voxels(g, voxel);
    for (size_t i = 0; i < walls.size() && (int64_t)i < cap; ++i) {
        parent[i] = walls[i].parentIdx; z[i] = walls[i].z; aspect[i] = walls[i].aspect;
    }
    return (int64_t)walls.size();
}

void ref_top_slope_aspect(const float* H, int w, int h, double sx, double sy, float* slope, float* aspect) {
    HeightGrid g;
    g.width = w; g.height = h; g.scaleX = sx; g.scaleY = sy;
    g.heights.assign(H, H + (size_t)w * h);
    SlopeAspect sa = compute_top_slope_aspect(g);
    std::memcpy(slope, sa.slope.data(), sa.slope.size() * sizeof(float));
    std::memcpy(aspect, sa.aspect.data(), sa.aspect.size() * sizeof(float));
}

int64_t ref_barle_compress_voxel(const int* raw, int max_neighbors, uint8_t* out, int64_t cap) {
    int raw_bytes = 0;
    std::vector<barle::Byte> enc = barle::compress_voxel_neighbors(raw, max_neighbors, raw_bytes);
    if ((int64_t)enc.size() > cap) return -(int64_t)enc.size();
    if (!enc.empty()) std::memcpy(out, enc.data(), enc.size());
    return (int64_t)enc.size();
}

int64_t ref_barle_compress_all(const int* raw, int n, int max_neighbors, int* offsets, uint8_t* flat, int64_t cap) {
    std::vector<int> off;
    std::vector<barle::Byte> data;
    std::size_t raw_bytes = 0;
    barle::compress_all_neighbors(raw, n, max_neighbors, off, data, raw_bytes);
    for (int i = 0; i <= n; ++i) offsets[i] = off[i];
    if ((int64_t)data.size() > cap) return -(int64_t)data.size();
    if (!data.empty()) std::memcpy(flat, data.data(), data.size());
    return (int64_t)data.size();
}

}  // extern "C"
