/* C ABI of the native I/O (include/sgpu.h: sgpu_geotiff_*, sgpu_read_tmy,
 * sgpu_write_voxel_dump). */
#include <cstring>
#include <memory>

#include "sgpu_ctx.h"
#include "sgpu_io.h"

struct sgpu_geotiff {
    sgpu::GeoTiff g;
};

using namespace sgpu;

extern "C" {

int sgpu_geotiff_open(const char* path, sgpu_geotiff** out) {
    return guarded([&] {
        need(path && out, SGPU_ERR_ARG, "null argument");
        *out = nullptr;
        std::unique_ptr<sgpu_geotiff> h(new sgpu_geotiff());
        h->g = read_geotiff(path);
        if (h->g.pixel_scale.size() < 2) fail(SGPU_ERR_FORMAT, "TIFF %s missing GeoTIFF pixel-scale tag", path);
        *out = h.release();
    });
}

int sgpu_geotiff_info(const sgpu_geotiff* h, int* w, int* hh, double* sx, double* sy, int* comp) {
    return guarded([&] {
        need(h != nullptr, SGPU_ERR_ARG, "null handle");
        if (w) *w = h->g.width;
        if (hh) *hh = h->g.height;
        if (sx) *sx = h->g.pixel_scale.size() > 0 ? h->g.pixel_scale[0] : 0.0;
        if (sy) *sy = h->g.pixel_scale.size() > 1 ? h->g.pixel_scale[1] : 0.0;
        if (comp) *comp = h->g.compression;
    });
}

int sgpu_geotiff_read(const sgpu_geotiff* h, float* out) {
    return guarded([&] {
        need(h && out, SGPU_ERR_ARG, "null argument");
        std::memcpy(out, h->g.data.data(), h->g.data.size() * sizeof(float));
    });
}

int sgpu_geotiff_write(const sgpu_geotiff* meta, const char* path, const float* data, int w, int hh) {
    return guarded([&] {
        need(meta && path && data && w > 0 && hh > 0, SGPU_ERR_ARG, "bad arguments");
        write_geotiff(path, data, w, hh, meta->g);
    });
}

int sgpu_geotiff_close(sgpu_geotiff* h) {
    return guarded([&] { delete h; });
}

int sgpu_read_tmy(const char* path, float* out, int n) {
    return guarded([&] {
        need(path && out && n > 0, SGPU_ERR_ARG, "bad arguments");
        std::vector<float> v = read_tmy(path, (size_t)n);
        std::memcpy(out, v.data(), v.size() * sizeof(float));
    });
}

int sgpu_write_voxel_dump(const char* path, int n_total, int n_top, int w, const int32_t* pp, const float* grid,
                          const float* wall_z, const float* irr) {
    return guarded([&] {
        need(path && pp && grid && irr && n_total >= n_top && w > 0, SGPU_ERR_ARG, "bad arguments");
        need(n_total == n_top || wall_z != nullptr, SGPU_ERR_ARG, "wall_z required");
        write_voxel_dump(path, n_total, n_top, w, pp, grid, wall_z, irr);
    });
}

}  // extern "C"
