/* SolarGPU CPU oracle -- TEST INFRASTRUCTURE ONLY.
 *
 * A scalar/multithreaded C++ restatement of the SolarGPU reference method
 * (SAMPA-Project/SolarGPU).  It exists to check the HIP product path: only the
 * test suite, the smoke check and the benchmark's CPU-baseline leg may load it,
 * and only as the checker or the timed CPU baseline, never as a product path.
 *
 * Every function cites the reference file:line it restates.  Float arithmetic
 * follows the reference's expression order; the library is built with
 * -ffp-contract=off and without fast-math.
 *
 * How each function is pinned (what would catch a plausible mistake):
 *   orc_sun_positions   golden vectors from the reference's own sun_position.h
 *                       (built under oracle/_ref) + independent NOAA/Spencer formula
 *   orc_sky_bins        golden vectors from the reference's sky_bins.cpp
 *   orc_detect_walls /
 *   orc_top_slope_aspect golden vectors from the reference's voxelization.cpp
 *   orc_directions      bit-identical to the product's host table; closed forms
 *                       through the flat-plane and single-box cases
 *   orc_trace_ray /
 *   orc_trace_raw /
 *   orc_viewshed        an independent vectorised numpy float32 restatement of the
 *                       DDA (tests/test_oracle_pins.py), list = first occurrence of
 *                       raw hits, special cases (flat plane, a single box with
 *                       hand-derived hits).  No reference run exists (CUDA only):
 *                       exact ids are pinned by the contraction-off restatement, as
 *                       the parity contract defines.
 *   orc_barle_encode_*  golden streams from the reference's barle_compression.cpp
 *   orc_barle_decode    round trip on the golden streams and on random lists
 *   orc_solar_hour      closed forms (flat unshaded voxel, shaded voxel = diffuse
 *                       only) against an independent float64 Perez; tilted planes
 *                       (tests/test_oracle_tilted.py): the beam on walls and slopes
 *                       facing the sun, away from it and due south at noon from the
 *                       geometric angle of incidence, and the Perez F1/F2 recovered
 *                       per hour from two vertical walls and fitted per clearness
 *                       bin against the published Perez (1990) coefficient table
 *   orc_reflect_hour_*  closed-form two-cell case; receivers tilted and facing the
 *                       four compass directions, normals written from geometry
 *   orc_annual_fp32 /
 *   orc_annual_fp32_range  agreement with the FP64 oracle within 1e-4
 *   orc_annual_fp64     linearity identity vs the per-hour FP64 sum on tiny scenes;
 *                       agreement with the pinned FP32 terms within 1e-4 on scenes
 *                       with walls (slope pi/2, every aspect)
 * Planted mistakes (degrees for radians in the incidence angle, a swapped or
 * negated normal component, a sign or a 1e-3 typo in the F2 row, the tilt
 * factor of the sky term) were each run against these tests and each failed.
 */
#pragma once
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HOURS 8760

/* PSA sun position (sun_position.h:24-40, 58-107): out[2k] = azimuth deg, out[2k+1] = altitude deg. */
void orc_sun_positions(float lon, float lat, int tz, float* az_alt_out);

/* Sky bins (sky_bins.cpp:9-52).  full_to_dense has hsize*vsize entries.
   Returns active bin count; *hgrid_size_s receives SVF bytes per voxel. */
int orc_sky_bins(const float* az_alt, float h_res, float v_res, float* nudged_out,
                 int* full_to_dense, int* sun_dense_idx, int* hsize, int* vsize, int* hgrid_size_s);

/* Ray directions (viewshed_stage.cuh:64-80): cdir[3d..3d+2], is_upper[d], dense[d] (-1 if none). */
int orc_directions(int hsize, int vsize, float hres, float vres, const int* full_to_dense,
                   float* cdir, int* is_upper, int* dense);

/* Wall voxels (voxelization.cpp:10-64).  Returns the count; writes at most cap entries. */
int64_t orc_detect_walls(const float* heights, int w, int h, float voxel, int* parent, float* z,
                         float* aspect, int64_t cap);
/* Top slope/aspect (voxelization.cpp:66-97). */
void orc_top_slope_aspect(const float* heights, int w, int h, double sx, double sy, float* slope,
                          float* aspect);

/* Pin of the product's ray set-up: for every direction (cdir, ndir) and axis and every integer
   origin coordinate g < gmax, the set-up quotient (fraction to the first cell edge) / |c| by
   IEEE division against q0 = RN(num * rx), q = fma(fma(-q0, a, num), rx, q0) with
   rx = RN(1 / a) (what the GPU computes).  Returns the number of quotients that differ. */
int64_t orc_check_setup_division(const float* cdir, int ndir, int gmax);

/* One DDA ray (viewshed_kernels.cu:7-59).  Returns hit cell id + 1, or 0.  *steps gets the step count. */
int orc_trace_ray(const float* grid, int w, int h, float maxz, float ox_cell, float oy_cell, float base_z,
                  const float* cdir3, int is_upper, int* steps);

/* Whole viewshed (viewshed_stage.cuh:36-125 + gpu_irradiance_uncompressed.cu:79-95).
   Writes svf (hgrid_size_s * n_total bytes, word-major), CSR offsets (n_total+1) and ids
   (1-based cell ids in first-occurrence direction order).  Returns nnz, or -(needed) if ids_cap
   is too small (svf and offsets are still valid then).  total_steps may be NULL. */
int64_t orc_viewshed(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                     const float* wall_z, float maxz, const float* cdir, const int* is_upper,
                     const int* dense, int ndir, int hgrid_size_s, int nthreads, uint8_t* svf,
                     int64_t* csr_off, int32_t* ids, int64_t ids_cap, int64_t* total_steps);
/* Viewshed of an arbitrary voxel list: SVF bits at svf[i + byte * n] (svf may be NULL), CSR
 * lists by position; returns nnz (-nnz when above cap). */
int64_t orc_viewshed_subset(const float* grid, int w, int h, int n_top, const int* pix_parent, const float* wall_z,
                            float maxz, const float* cdir, const int* up, const int* dense, int ndir, int hgs,
                            const int32_t* voxels, int64_t n, int nthreads, uint8_t* svf, int64_t* off,
                            int32_t* ids, int64_t cap, int64_t* total_steps);

/* Raw per-direction hits for voxels [v0, v0+count): raw[(v-v0)*ndir + d] = id+1 or 0. */
void orc_trace_raw(const float* grid, int w, int h, int n_top, const int* pix_parent, const float* wall_z,
                   float maxz, const float* cdir, const int* is_upper, int ndir, int v0, int count,
                   int nthreads, int32_t* raw);

/* ba-RLE encoder for one voxel (barle_compression.cpp:156-187).  ids: n values (any order,
   zero entries end the list as in the reference).  Returns byte count, or -(needed). */
int64_t orc_barle_encode_voxel(const int32_t* ids, int n, uint8_t* out, int64_t cap);
/* All voxels from a CSR list (barle_compression.cpp:189-207 with int64 offsets). */
int64_t orc_barle_encode_all(const int64_t* csr_off, const int32_t* ids, int n_total, int nthreads,
                             int64_t* off_out, uint8_t* bytes, int64_t cap);
/* Decoder walk order of reflection_kernels.cu:132-180: emits every decoded value (including
   zeros) in stream order.  Returns the count, or -1 on a malformed stream. */
int orc_barle_decode(const uint8_t* bytes, int64_t size, uint32_t* vals, int cap);

/* Per-hour constants of the sunlit hours (gpu_irradiance_uncompressed.cu:114-117 and the
   hour-only part of solar_kernel.cu:8-50).  Returns the number of sunlit hours; hours_out
   receives their indices day*24+hour. */
int orc_sunlit_hours(const float* nudged_az_alt, int* hours_out);

/* Perez sky clearness epsilon of one hour in FP32 (solar_kernel.cu:23-40: zenith from the
   nudged altitude, max(1, eps)), into *eps; returns its bin 0..7 (the discrete decision the
   FP64 oracle shares with the FP32 path). */
int orc_hour_eps(float alt, float dif, float dir, float* eps);

/* Solar kernel for one hour (solar_kernel.cu:54-77) on voxels [0, n). */
void orc_solar_hour(const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                    const float* dhi, int hour_index, int n, const float* slopes, const float* aspects,
                    const uint8_t* svf, int n_total, int nthreads, float* hourly);

/* Reflection for one hour (reflection_kernels.cu:18-89 CSR order, or :103-183 ba-RLE order). */
void orc_reflect_hour_csr(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                          const float* wall_z, const float* slopes, const float* aspects,
                          const int64_t* off, const int32_t* ids, const float* hourly, int nthreads,
                          float* reflected);
void orc_reflect_hour_barle(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                            const float* wall_z, const float* slopes, const float* aspects,
                            const int64_t* off, const uint8_t* bytes, const float* hourly, int nthreads,
                            float* reflected);

/* Reference-structured FP32 annual loop (gpu_irradiance_uncompressed.cu:114-134 /
   gpu_irradiance_compressed.cu:123-143).  mode 0 = CSR, 1 = ba-RLE.  hour_stride > 1 runs only
   every hour_stride-th sunlit hour (bounded CPU-baseline samples).  kahan=1 replaces the naive
   accumulate with a compensated sum.  Returns the number of hours computed. */
int orc_annual_fp32(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                    const float* wall_z, const float* slopes, const float* aspects,
                    const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                    const float* dhi, const uint8_t* svf, const int64_t* off, const int32_t* ids,
                    const uint8_t* bytes, int mode, int hour_stride, int kahan, int nthreads,
                    float* out);

/* Stage-2 CPU baseline sample: the reference-structured FP32 hour loop for voxels
   [v0, v0+count) only (hourly of every top cell is still computed, since any voxel may
   see any cell).  Lists are given for the range: off has count+1 entries (off[0] may be
   nonzero), ids is indexed by off.  t_top_s / t_range_s (may be NULL) receive the seconds
   spent on the all-top-cell solar term and on the range.  Returns the number of hours computed. */
int orc_annual_fp32_range(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                          const float* wall_z, const float* slopes, const float* aspects,
                          const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                          const float* dhi, const uint8_t* svf, int64_t v0, int64_t count,
                          const int64_t* off, const int32_t* ids, int hour_stride, int nthreads, float* out,
                          double* t_top_s, double* t_range_s);
/* Same loop for an arbitrary voxel list (voxels[0..count)), lists by position (the CPU
 * baseline's random sample). */
int orc_annual_fp32_subset(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                           const float* wall_z, const float* slopes, const float* aspects,
                           const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                           const float* dhi, const uint8_t* svf, const int32_t* voxels, int64_t count,
                           const int64_t* off, const int32_t* ids, int stride, int nthreads, float* out,
                           double* t_top_s, double* t_range_s);

/* FP64 oracle.  Same discrete decisions as the FP32 path (sunlit-hour set, SVF bits, neighbour
   sets, Perez epsilon bin taken from the FP32 hour constants); all continuous arithmetic in
   double.  Uses the single-bounce linearity: annual(v) = sum_h H_h(v) + sum_n w(v,n) sum_h H_h(n).
   voxels: optional subset (count n_sel) or NULL for all; out[i] is for voxels[i]. */
void orc_annual_fp64(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                     const float* wall_z, const float* slopes, const float* aspects,
                     const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                     const float* dhi, const uint8_t* svf, const int64_t* off, const int32_t* ids,
                     const int32_t* voxels, int n_sel, int nthreads, double* out);
/* Same for a voxel list with its lists by position (off_sub: n_sel + 1, ids_sub): for scenes
 * whose whole CSR does not fit host memory. */
void orc_annual_fp64_subset(const float* grid, int w, int h, int n_total, int n_top, const int* pix_parent,
                            const float* wall_z, const float* slopes, const float* aspects,
                            const float* nudged_az_alt, const int* sun_dense_idx, const float* dni,
                            const float* dhi, const uint8_t* svf, const int32_t* voxels, int64_t n_sel,
                            const int64_t* off_sub, const int32_t* ids_sub, int nthreads, double* out);

/* Per-hour FP64 sum without the linearity shortcut (tiny scenes; pins orc_annual_fp64). */
void orc_annual_fp64_direct(const float* grid, int w, int h, int n_total, int n_top,
                            const int* pix_parent, const float* wall_z, const float* slopes,
                            const float* aspects, const float* nudged_az_alt, const int* sun_dense_idx,
                            const float* dni, const float* dhi, const uint8_t* svf, const int64_t* off,
                            const int32_t* ids, double* out);

#ifdef __cplusplus
}
#endif
