/* solargpu: drop-in for the reference CLI (src/main.cpp:71-183) on MI355X.
 *
 *   solargpu <dsm.tiff> <lon> <lat> <tz> <direct.txt> <diffuse.txt>
 *
 * Same positional arguments and defaults, same outputs: the annual irradiance
 * of the top voxels as <name>_solar_potential.tiff (geo tags of the input) and
 * every voxel as <name>_voxels.bin, both from the ba-RLE pipeline.  Both
 * pipelines run through the reference's C++ entry points (solar_gpu_api.h). */
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "sgpu.h"
#include "solar_gpu_api.h"

namespace {

[[noreturn]] void die(const char* what) {
    std::fprintf(stderr, "%s: %s\n", what, sgpu_last_error());
    std::exit(EXIT_FAILURE);
}

std::string base_name(const std::string& p) {
    const size_t s = p.find_last_of('/');
    return s == std::string::npos ? p : p.substr(s + 1);
}

}  // namespace

int main(int argc, char** argv) {
    const std::string dsm = argc > 1 ? argv[1] : "test_data/dsm_heightmap.tiff";
    const float lon = argc > 2 ? (float)std::atof(argv[2]) : 15.592f;
    const float lat = argc > 3 ? (float)std::atof(argv[3]) : 46.543f;
    const int tz = argc > 4 ? std::atoi(argv[4]) : (int)std::lround(lon / 15.0f);
    const std::string direct_path = argc > 5 ? argv[5] : "test_data/tmy_direct.txt";
    const std::string diffuse_path = argc > 6 ? argv[6] : "test_data/tmy_diffuse.txt";

    std::vector<float> dni(SGPU_HOURS_PER_YEAR), dhi(SGPU_HOURS_PER_YEAR);
    if (sgpu_read_tmy(direct_path.c_str(), dni.data(), SGPU_HOURS_PER_YEAR)) die("direct irradiance");
    if (sgpu_read_tmy(diffuse_path.c_str(), dhi.data(), SGPU_HOURS_PER_YEAR)) die("diffuse irradiance");

    sgpu_geotiff* tif = nullptr;
    if (sgpu_geotiff_open(dsm.c_str(), &tif)) die("heightmap");
    int w = 0, h = 0, comp = 0;
    double sx = 0, sy = 0;
    sgpu_geotiff_info(tif, &w, &h, &sx, &sy, &comp);
    std::vector<float> heights((size_t)w * h);
    sgpu_geotiff_read(tif, heights.data());
    std::printf("Heightmap: %s (%dx%d) scale=(%g, %g)\n", dsm.c_str(), w, h, sx, sy);
    const float maxz = *std::max_element(heights.begin(), heights.end());

    std::vector<solar_gpu::SunSample> sun;
    solar_gpu::SunPositionCalculator(lon, lat, tz).calc(sun);

    int64_t n_total = 0;
    if (sgpu_assemble_voxels(heights.data(), w, h, sx, sy, 0, &n_total, nullptr, nullptr, nullptr, nullptr))
        die("voxelization");
    const int n_top = w * h;
    std::vector<int32_t> pp(n_total);
    std::vector<float> slopes(n_total), aspects(n_total), wall_z(std::max<int64_t>(n_total - n_top, 0));
    if (sgpu_assemble_voxels(heights.data(), w, h, sx, sy, n_total, &n_total, pp.data(), slopes.data(),
                             aspects.data(), wall_z.data()))
        die("voxelization");
    std::printf("Top voxels: %d, wall voxels: %lld, total voxels: %lld\n", n_top, (long long)(n_total - n_top),
                (long long)n_total);

    std::vector<float> irr_csr(n_total), irr_barle(n_total);
    const float* wz = wall_z.empty() ? nullptr : wall_z.data();
    std::printf("\n=== Uncompressed (CSR) pipeline ===\n");
    if (solar_gpu::uncompressed_gpu_irradiance(heights.data(), w, h, (int)n_total, n_top, pp.data(), wz, maxz, sun,
                                               dni, dhi, slopes.data(), aspects.data(), irr_csr.data()))
        return EXIT_FAILURE;
    std::printf("\n=== Compressed (ba-RLE) pipeline ===\n");
    if (solar_gpu::compressed_gpu_irradiance(heights.data(), w, h, (int)n_total, n_top, pp.data(), wz, maxz, sun, dni,
                                             dhi, slopes.data(), aspects.data(), irr_barle.data()))
        return EXIT_FAILURE;

    double max_abs = 0.0, sum_abs = 0.0;
    for (int64_t i = 0; i < n_total; ++i) {
        const double d = std::fabs((double)irr_csr[i] - (double)irr_barle[i]);
        max_abs = std::max(max_abs, d);
        sum_abs += d;
    }
    std::printf("\nCompressed vs uncompressed: max abs diff = %g, mean abs diff = %g\n", max_abs,
                n_total ? sum_abs / n_total : 0.0);

    const std::string name = base_name(dsm);
    if (sgpu_geotiff_write(tif, (name + "_solar_potential.tiff").c_str(), irr_barle.data(), w, h))
        die("output raster");
    if (sgpu_write_voxel_dump((name + "_voxels.bin").c_str(), (int)n_total, n_top, w, pp.data(), heights.data(), wz,
                              irr_barle.data()))
        std::fprintf(stderr, "Warning: cannot write %s_voxels.bin: %s\n", name.c_str(), sgpu_last_error());
    sgpu_geotiff_close(tif);
    return 0;
}
