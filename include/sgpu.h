/* sgpu.h -- C ABI of the MI355X-native SolarGPU pipeline (libsolargpu.so).
 *
 * Plain pointers and sizes only.  Every entry point returns an sgpu_status;
 * on failure sgpu_last_error() gives a message for the calling thread.  No
 * entry point calls exit().
 *
 * Reference interfaces these entry points replace (SAMPA-Project/SolarGPU):
 *   sgpu_uncompressed_irradiance  include/solar_gpu_api.h:37-42 (uncompressed_gpu_irradiance)
 *   sgpu_compressed_irradiance    include/solar_gpu_api.h:49-54 (compressed_gpu_irradiance)
 *   sgpu_sun_positions            include/sun_position.h:24-40 (SunPositionCalculator::calc)
 *   sgpu_sky_bins                 src/sky_bins.cpp:9-52 (compute_sky_bins)
 *   sgpu_detect_walls             src/voxelization.cpp:10-64 (detect_wall_voxels)
 *   sgpu_top_slope_aspect         src/voxelization.cpp:66-97 (compute_top_slope_aspect)
 *   sgpu_assemble_voxels          src/main.cpp:112-127 (voxel array assembly)
 *   sgpu_viewshed                 include/cuda/viewshed_stage.cuh:36-125 (run_viewshed_out_of_core)
 *                                 + the CSR consumer (src/cuda/gpu_irradiance_uncompressed.cu:79-95)
 *                                 + the ba-RLE consumer (src/cuda/gpu_irradiance_compressed.cu:81-100,
 *                                   src/barle_compression.cpp:189-207)
 *   sgpu_hour_loop                src/cuda/gpu_irradiance_uncompressed.cu:107-135 and
 *                                 src/cuda/gpu_irradiance_compressed.cu:116-144 (stage-2 hour loop)
 *   sgpu_hour_single              one iteration of that loop (src/cuda/kernels/solar_kernel.cu:54-77,
 *                                 src/cuda/kernels/reflection_kernels.cu:70-183)
 *   sgpu_barle_encode             src/barle_compression.cpp:156-207 (compress_all_neighbors), on the GPU
 * Entry points with no reference counterpart are marked "(new)".
 */
#ifndef SGPU_H
#define SGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SGPU_HOURS_PER_YEAR 8760

typedef enum sgpu_status {
    SGPU_OK = 0,
    SGPU_ERR_ARG = 1,     /* invalid argument / inconsistent sizes */
    SGPU_ERR_HIP = 2,     /* HIP runtime failure */
    SGPU_ERR_OOM = 3,     /* device or host allocation failed */
    SGPU_ERR_STATE = 4,   /* call out of order (e.g. hour loop before viewshed) */
    SGPU_ERR_IO = 5,      /* file could not be read/written */
    SGPU_ERR_FORMAT = 6   /* malformed input file or stream */
} sgpu_status;

/* Neighbour-list storage formats (bit flags for sgpu_viewshed). */
#define SGPU_FMT_CSR 1u
#define SGPU_FMT_BARLE 2u

typedef struct sgpu_ctx sgpu_ctx;

/* Run statistics of a context (new). Times are HIP-event milliseconds on the context stream. */
typedef struct sgpu_info {
    int32_t grid_width, grid_height;
    int64_t num_total, num_top;
    int32_t active_bins;        /* sun-visited upper-hemisphere bins */
    int32_t svf_bytes_per_voxel;
    int32_t sunlit_hours;       /* hours with nudged altitude >= 1 deg */
    int32_t viewshed_chunks;
    int64_t chunk_voxels;
    int64_t nnz;                /* CSR neighbour ids */
    int64_t barle_bytes;        /* ba-RLE stream bytes */
    double t_trace_ms;          /* DDA trace kernels */
    double t_lists_ms;          /* dedupe/compaction/encode kernels */
    double t_viewshed_ms;       /* whole stage 1 */
    double t_hours_ms;          /* whole stage 2 (last sgpu_hour_loop) */
    int32_t hours_computed;     /* last sgpu_hour_loop */
    int32_t hour_passes;        /* kernel passes of the last sgpu_hour_loop */
    int32_t hours_per_pass;
    int32_t timed_launches;     /* reflection launches timed individually (sgpu_set_kernel_timing) */
    double t_refl_kernel_ms;    /* sum of those launch times */
    double t_solar_kernel_ms;   /* sum of solar-top launch times */
    int64_t dda_steps;          /* only when step counting is enabled */
    /* the step-table trace of the last sgpu_viewshed (new): rays left to the pooled walk after
       their table's steps, voxels whose cells are near a binade end (all their rays by the pooled
       walk), chunks traced by tables, and chunks traced again by the pooled walk alone because the
       pending rays overflowed their buffer */
    int64_t table_pending_rays;
    int64_t table_irregular_voxels;
    int32_t table_chunks;
    int32_t table_fallbacks;
} sgpu_info;

const char* sgpu_version(void);
const char* sgpu_last_error(void);

/* ---- host-side contract (CPU) ---- */
int sgpu_sun_positions(float lon, float lat, int tz, float* az_alt_out /* 2*8760 */);
int sgpu_sky_bins(const float* az_alt /* 2*8760 */, float* nudged_out /* 2*8760 */,
                  int32_t* full_to_dense /* 72*18 */, int32_t* sun_dense_idx /* 8760 */,
                  int32_t* active_bins, int32_t* svf_bytes_per_voxel);
int sgpu_directions(float* cdir /* 3*2592 */, int32_t* is_upper /* 2592 */); /* (new) host Cdir table */
int sgpu_detect_walls(const float* heights, int w, int h, float voxel_size, int32_t* parent,
                      float* z, float* aspect, int64_t cap, int64_t* count);
int sgpu_top_slope_aspect(const float* heights, int w, int h, double scale_x, double scale_y,
                          float* slope, float* aspect);
/* Full voxel arrays from a heightmap: n_total/n_top out; arrays sized by the caller with
   capacity cap (call once with cap = 0 to get n_total). */
int sgpu_assemble_voxels(const float* heights, int w, int h, double scale_x, double scale_y,
                         int64_t cap, int64_t* n_total, int32_t* pixel_parent, float* slopes,
                         float* aspects, float* wall_z);
/* GeoTIFF heightmap I/O without libtiff (geotiff_io.cpp:50-178, heightmap.cpp:7-24).
   Open parses the file; dims/scale come from it (the pixel-scale tag is required, as in
   heightmap.cpp:11-14); read copies the raster; write emits a float32 GeoTIFF carrying the
   opened file's geo tags and compression (one strip, no predictor). */
typedef struct sgpu_geotiff sgpu_geotiff;
int sgpu_geotiff_open(const char* path, sgpu_geotiff** out);
int sgpu_geotiff_info(const sgpu_geotiff* g, int* width, int* height, double* scale_x, double* scale_y,
                      int* compression);
int sgpu_geotiff_read(const sgpu_geotiff* g, float* out /* width*height */);
int sgpu_geotiff_write(const sgpu_geotiff* meta, const char* path, const float* data, int width, int height);
int sgpu_geotiff_close(sgpu_geotiff* g);
/* Whitespace-separated TMY series (main.cpp:30-40); fails on fewer than n values. */
int sgpu_read_tmy(const char* path, float* out, int n);
/* (gx, gy, z, irradiance) float records after (n_total, n_top) int32 (main.cpp:44-67). */
int sgpu_write_voxel_dump(const char* path, int n_total, int n_top, int grid_width, const int32_t* pixel_parent,
                          const float* grid, const float* wall_z, const float* irradiance);

/* Seeded synthetic inputs (new). */
int sgpu_synth_dsm(uint64_t seed, int w, int h, float cell_m, float building_density, int trees,
                   float* heights_out);
int sgpu_synth_weather(uint64_t seed, const float* az_alt, float* dni_out, float* dhi_out);

/* ---- device pipeline ---- */
int sgpu_create(int device, sgpu_ctx** out);
int sgpu_destroy(sgpu_ctx* ctx);
int sgpu_get_info(sgpu_ctx* ctx, sgpu_info* out);

/* Set the scene from a heightmap, voxelized on the GPU: the wall voxels and slopes/aspects of
   src/voxelization.cpp:10-97 (detect_wall_voxels, compute_top_slope_aspect) and the assembly of
   src/main.cpp:112-127 (tops in raster order, then walls; voxel size = scale_x; max_height = the
   heightmap's maximum), in one call with the arrays left on the device; the processing orders are
   built there too.  Parents and wall heights equal sgpu_assemble_voxels' bit for bit; slopes and
   aspects agree to the device atan2f's rounding (3 ulp). */
int sgpu_set_scene_heights(sgpu_ctx* ctx, const float* heights, int grid_width, int grid_height,
                           double scale_x, double scale_y);
/* The context's processing orders (num_total entries each; either may be NULL): the reflection
   pass's (Hilbert-ordered 16-cell tiles, each top voxel followed by its walls) and the DDA's lane
   order (blocks of 512 ids sorted by parent-cell height, voxel height).  No reference counterpart;
   for tests and tooling. */
int sgpu_get_orders(sgpu_ctx* ctx, int32_t* order, int32_t* lane_order);
/* The context's voxel arrays (num_total entries; wall_z num_total - num_top); any may be NULL. */
int sgpu_get_voxels(sgpu_ctx* ctx, int32_t* pixel_parent, float* slopes, float* aspects, float* wall_z);

/* Persisted neighbour lists (no reference counterpart: the reference recomputes the viewshed on
   every run).  sgpu_save_lists writes the SVF bitmap and the device-resident lists of `formats`
   (CSR and/or ba-RLE, int64 offsets) to `path`; sgpu_load_lists reads them into a context whose
   scene (dimensions and a checksum of heights, parents, wall heights and height limit) and sky
   (sun bins) match the file's, in place of sgpu_viewshed.  Mismatches return SGPU_ERR_ARG,
   damaged files SGPU_ERR_FORMAT. */
int sgpu_save_lists(sgpu_ctx* ctx, const char* path, unsigned formats);
int sgpu_load_lists(sgpu_ctx* ctx, const char* path);

/* Upload the scene.  wall_z may be NULL when num_total == num_top. */
int sgpu_set_scene(sgpu_ctx* ctx, const float* grid, int grid_width, int grid_height, int64_t num_total,
                   int64_t num_top, const int32_t* pixel_parent, const float* wall_z, float max_height,
                   const float* slopes, const float* aspects);
/* Sun positions (2*8760 floats, az/alt degrees, NOT nudged) and TMY DNI/DHI (8760 each). */
int sgpu_set_sky(sgpu_ctx* ctx, const float* az_alt, const float* dni, const float* dhi);
/* Override the viewshed chunk (voxels per out-of-core chunk); 0 = from free memory. (new) */
int sgpu_set_chunk_voxels(sgpu_ctx* ctx, int64_t chunk_voxels);
/* Where sgpu_viewshed leaves the neighbour lists (new).  HOST: every chunk's lists go to
   pinned host memory as soon as they are built (the device holds one chunk at a time), and
   sgpu_hour_loop streams them back tile by tile, double-buffered with hipMemcpyAsync on a
   second stream overlapping the reflection kernel.  Results are identical to DEVICE. */
#define SGPU_LISTS_DEVICE 0
#define SGPU_LISTS_HOST 1
int sgpu_set_list_residency(sgpu_ctx* ctx, int residency);
/* Stage 1: trace every voxel x 2592 directions; builds the SVF bitmap and the requested formats. */
int sgpu_viewshed(sgpu_ctx* ctx, unsigned formats);
int sgpu_get_svf(sgpu_ctx* ctx, uint8_t* out /* svf_bytes_per_voxel * num_total */);
int sgpu_get_csr(sgpu_ctx* ctx, int64_t* offsets /* num_total+1 */, int32_t* ids /* nnz */);
int sgpu_get_barle(sgpu_ctx* ctx, int64_t* offsets /* num_total+1 */, uint8_t* bytes);
/* CSR lists of voxels [v0, v0+count): offsets rebased to 0 (count+1), ids up to cap (new). */
int sgpu_get_csr_range(sgpu_ctx* ctx, int64_t v0, int64_t count, int64_t* offsets, int32_t* ids, int64_t cap);
/* ba-RLE streams of voxels [v0, v0+count): offsets rebased to 0 (count+1), bytes up to cap (new). */
int sgpu_get_barle_range(sgpu_ctx* ctx, int64_t v0, int64_t count, int64_t* offsets, uint8_t* bytes, int64_t cap);
/* Raw first-hit ids (id+1, 0 = none) of voxels [v0, v0+count) x 2592 directions (new; tests). */
int sgpu_trace_raw(sgpu_ctx* ctx, int64_t v0, int64_t count, int32_t* raw_out);
/* Supply lists/SVF from elsewhere (tests, a persisted cache) instead of tracing. (new) */
int sgpu_set_svf(sgpu_ctx* ctx, const uint8_t* svf);
int sgpu_set_csr(sgpu_ctx* ctx, const int64_t* offsets, const int32_t* ids);
int sgpu_set_barle(sgpu_ctx* ctx, const int64_t* offsets, const uint8_t* bytes);
/* GPU ba-RLE encoder from the context's CSR lists (device to device). */
int sgpu_barle_encode(sgpu_ctx* ctx);

/* Stage 2: the annual loop over sunlit hours with format SGPU_FMT_CSR or SGPU_FMT_BARLE.
   hours_per_pass (1, 2, 4, ..., 256) hours share one pass over the neighbour stream; 0 = autotuned.
   hour_stride > 1 computes every hour_stride-th sunlit hour only (bounded samples).
   Annual sums are FP32 Kahan-compensated. */
int sgpu_hour_loop(sgpu_ctx* ctx, unsigned format, int hours_per_pass, int hour_stride);
/* Allocation-free hour loop.  No counterpart in the reference: it is for callers that cannot
   allocate device memory while they run (code under HIP graph capture, tight serving loops,
   harnesses that forbid hipMalloc).  The full sunlit-hour loop of sgpu_hour_loop (same kernels,
   same results bit for bit) over the context's device-resident lists of `format`; the annual
   sums go to irradiance_out, a device buffer of num_total floats owned by the caller.  The
   per-pass scratch (num_top x hours_per_pass floats of hourly rows + num_total floats of Kahan
   compensation, each rounded up to 256 bytes, + 32 KiB of reflection queue counters) lives
   in a static device array of 128 MiB per
   GPU, held from the first context on that GPU to the end of the process; a scene that needs
   more is refused with SGPU_ERR_ARG.  hours_per_pass 0 = the tuned value, halved until it fits.
   Calls no hipMalloc/hipFree; enqueued on the context stream (sgpu_sync waits).  Calls that
   overlap on one GPU share its array and race. */
int sgpu_hour_loop_noalloc(sgpu_ctx* ctx, unsigned format, int hours_per_pass, float* irradiance_out);
/* Record the hour loop as one hipGraph and replay it (new). */
int sgpu_hour_loop_graph(sgpu_ctx* ctx, unsigned format, int hours_per_pass, int replays);
int sgpu_get_irradiance(sgpu_ctx* ctx, float* out /* num_total */);
/* One hour (index day*24+hour) through the solar + reflection kernels; host outputs. */
int sgpu_hour_single(sgpu_ctx* ctx, unsigned format, int hour_index, float* hourly_out,
                     float* reflected_out);
/* Time each reflection / solar launch with HIP events on the context stream (new). */
int sgpu_set_kernel_timing(sgpu_ctx* ctx, int enable);
/* Diagnostic (new; tests): the hour passes may use at most `floats` of the hourly irradiance
   buffer (0 = all of it).  A pass whose n_top x hours_per_pass rows do not fit is refused with
   SGPU_ERR_STATE before its kernels launch. */
int sgpu_set_hourly_limit(sgpu_ctx* ctx, int64_t floats);
/* Count DDA steps in the trace kernel (slower; new). */
int sgpu_set_step_counting(sgpu_ctx* ctx, int enable);
/* Trace whole-scene viewsheds by step tables (enable = 1) or by the pooled walk alone (0, the
   default: on the 2048^2 bench scene the tables run 3-7% behind it); the lists are the same
   either way.  pending_cap > 0 caps the pending rays per chunk below the buffer (tests
   of the overflow path); 0 = the buffer (chunk x 2592 / 8 rays).  (new; the reference traces
   every ray on its own, viewshed_kernels.cu:7-59) */
int sgpu_set_trace_tables(sgpu_ctx* ctx, int enable, int64_t pending_cap);
/* Shape of the step-table trace (new): directions per workgroup x steps per table, one of
   4 x 255 (the default), 4 x 510, 8 x 255, 8 x 126, 4 x 126, 16 x 126, 8 x 63.  Larger tables leave
   fewer rays to the pooled walk; smaller ones take less LDS.  skip = 1 (the default) lets a ray
   jump 9 steps at once when no cell among them can stop it as a hit (quadrant max maps of the
   heightmap); the results are the same either way.  Step counting walks every step. */
int sgpu_set_trace_table_shape(sgpu_ctx* ctx, int dirs, int steps, int skip);
/* Diagnostic (new; tests): receiver pairs (receiver_queue 2) whose kept neighbour lists hold more
   than `entries` entries together are walked receiver by receiver (0, the default: only pairs
   whose lists overflow the pair scratch, 2 x 240 entries in LDS + 4096 per wave); a small limit
   forces that path. */
int sgpu_set_pair_limit(sgpu_ctx* ctx, int entries);
/* Synchronise the context stream. */
int sgpu_sync(sgpu_ctx* ctx);

/* Autotuner (new).  Launch shapes of the hot kernels, measured on the context's own scene:
   the DDA tile (directions per workgroup of 256 voxels), the hours per stage-2 pass, the
   waves per reflection workgroup (LDS footprint / occupancy), the order in which the
   eight XCDs walk the voxels and how the resident waves share receivers (cache sharing).  Results persist in a
   text cache ($SGPU_AUTOTUNE_CACHE, else ~/.cache/solargpu/autotune.txt) keyed by device,
   format and scene size class; a cached key is applied without measuring.  Stage-2 shapes
   are tuned only when the lists of `format` exist.  Every shape gives identical results. */
typedef struct sgpu_tuning {
    int32_t dirs_per_group;      /* 16, 32 or 48 */
    int32_t hours_per_pass;      /* 1, 2, 4, ..., 256 */
    int32_t waves_per_group;     /* 2 or 4 */
    int32_t xcd_band;            /* reflection workgroups per XCD band: 0 (one contiguous eighth
                                    of the voxels per XCD) or 16..4096 (XCDs side by side) */
    int32_t group_blocks;        /* reflection workgroups whose waves interleave their receivers
                                    over a tile-major voxel order: 0 (each wave 64 consecutive
                                    voxels) or 2..1024 */
    int32_t from_cache;
    double rays_per_s;           /* best measured trace rate (0 if from cache) */
    double voxel_hours_per_s;    /* best measured stage-2 rate (0 if from cache / not tuned) */
    int32_t receiver_queue;      /* 1: the reflection waves take receivers one by one from queues per
                                    XCD in the tile-major order (XCD bands of xcd_band workgroups'
                                    worth of receivers; group_blocks then caps the resident
                                    workgroups per CU, 0 = as many as fit); 0: static assignment.
                                    Both give bit-identical annual sums.  2: queued pairs of
                                    consecutive receivers, each walked by one wave that gathers
                                    every neighbour row of the pair once (ba-RLE lists, 16 or more
                                    hours per pass; otherwise as 1); the sums differ from 1 only in
                                    the order of the additions. */
} sgpu_tuning;
int sgpu_autotune(sgpu_ctx* ctx, unsigned format, sgpu_tuning* out);
int sgpu_set_tuning(sgpu_ctx* ctx, const sgpu_tuning* t);
int sgpu_get_tuning(sgpu_ctx* ctx, sgpu_tuning* out);
/* Plain device copy of `bytes` (float4 loads/stores), `reps` launches; *gbps = read+write
   bytes / mean launch time.  Calibrates HBM counters and the measured copy bandwidth (new). */
int sgpu_copy_bandwidth(sgpu_ctx* ctx, int64_t bytes, int reps, double* gbps);
/* VALU issue rate of the device in wave64 instructions/s, measured with a kernel of known
   instruction count that fills every SIMD (new; the peak the DDA's VALU roofline uses):
   mix 0 = FP32 FMAs only, mix 1 = the DDA's own classes (integer add, compare, select). */
int sgpu_valu_rate(sgpu_ctx* ctx, int mix, double* wave_insts_per_s);
/* L2 read bandwidth in GB/s: every wave streams a `footprint`-byte buffer (resident in each XCD's
   L2, beyond a CU's L1) as 1 KB rows of 16-byte lane loads, `reps` launches (new; the peak of
   the reflection pass's row gathers).  footprint: a multiple of 4 KiB in 64 KiB .. 64 MiB. */
int sgpu_l2_bandwidth(sgpu_ctx* ctx, int64_t footprint, int reps, double* gbps);
/* L1 gather rate in tag lookups/s (the unit of the TCP_TOTAL_CACHE_ACCESSES counter): every
   lane of every wave loads 4-byte values from its own 128-byte line of a `footprint`-byte buffer,
   `reps` launches (new; the second peak of the DDA, whose pooled lanes gather one height each
   from scattered cells).  footprint: a power of two in 4 KiB .. 1 GiB (16 KiB stays in L1). */
int sgpu_l1_gather_rate(sgpu_ctx* ctx, int64_t footprint, int reps, double* accesses_per_s);

/* ---- per-voxel surrogate MLP (new; no reference counterpart) ----
   64 inputs (53 features + zero padding) -> 4 x 256 GELU -> 2 outputs; BF16 weights and
   activations, FP32 accumulation on MFMA; Adam.  Parameters as one fp32 vector: per layer
   W[out][in] then b[out]. */
typedef struct sgpu_mlp sgpu_mlp;
const char* sgpu_mlp_last_error(void);
int sgpu_mlp_create(int device, uint64_t seed, sgpu_mlp** out);
int sgpu_mlp_destroy(sgpu_mlp* mlp);
int sgpu_mlp_param_count(void);
int sgpu_mlp_get_params(sgpu_mlp* mlp, float* params);
int sgpu_mlp_set_params(sgpu_mlp* mlp, const float* params);
int sgpu_mlp_set_loss_weights(sgpu_mlp* mlp, float w0, float w1);
/* Inference kernel (new): rows = 32 or 64 runs the fused kernel with that many voxels per
   workgroup (activations kept on chip through all layers; the default is 64), rows = 0 the
   layer-by-layer GEMMs.  Both give bit-identical outputs. */
int sgpu_mlp_set_inference(sgpu_mlp* mlp, int rows);
/* X: n x 64 bf16 (host), Y: n x 2 (host). */
int sgpu_mlp_forward(sgpu_mlp* mlp, const uint16_t* X, int64_t n, float* Y);
/* X, Y device pointers; *ms = kernel time (HIP events). */
int sgpu_mlp_forward_device(sgpu_mlp* mlp, const uint16_t* X, int64_t n, float* Y, double* ms);
int sgpu_mlp_train(sgpu_mlp* mlp, const uint16_t* X, const float* T, int64_t n, int epochs, int batch, float lr,
                   uint64_t seed, double* epoch_losses);
/* Gradient of the mean weighted squared error over rows [0, B) (B multiple of 64). */
int sgpu_mlp_gradient(sgpu_mlp* mlp, const uint16_t* X, const float* T, int B, float* grad, double* loss);

/* ---- surrogate path (new; no reference counterpart) ----
   The reflection pass replaced by the MLP where its predicted error of the annual total is
   under gate->tau; elsewhere the exact plain result (2592-direction trace, ba-RLE lists,
   per-hour reflection loop).  Direct and diffuse are always exact. */
typedef struct sgpu_gate {
    float tau;               /* admissible predicted relative error of the annual total */
    float refl_rel_err;      /* calibrated relative error of the predicted reflected share */
    float diff_miss_weight;  /* weight of the diffuse-share miss (out-of-distribution signal) */
    float share_abs_err;     /* calibrated absolute error of the reflected share */
} sgpu_gate;

typedef struct sgpu_surrogate_stats {
    int64_t fallback;        /* voxels sent to the exact path */
    double t_svf_ms, t_dd_ms, t_features_ms, t_mlp_ms, t_fallback_ms, t_total_ms;
} sgpu_surrogate_stats;

/* Sun-bin SVF (identical bits to sgpu_viewshed's), exact annual direct + diffuse and the
   feature matrix, on the device.  lat (site latitude, degrees) and cell_m (DSM cell size,
   metres) are features; sgpu_surrogate_run reuses them. */
int sgpu_surrogate_prepare(sgpu_ctx* ctx, float lat, float cell_m, sgpu_surrogate_stats* stats);
/* Download features (N x 64 bf16), annual direct+diffuse and diffuse (any may be NULL). */
int sgpu_surrogate_get(sgpu_ctx* ctx, uint16_t* X, float* dd, float* diff);
/* Annual reflected irradiance of the plain method (needs the neighbour lists). */
int sgpu_surrogate_labels(sgpu_ctx* ctx, unsigned format, float* reflected);
/* The whole surrogate path; out: annual totals (N), fallback: per-voxel flag (N, optional).
   The gated voxels get the exact plain result from their own 2592-direction lists (ba-RLE):
   hours_per_pass > 0 runs the sunlit-hour loop on them, K hours per pass;
   hours_per_pass == SGPU_FALLBACK_ANNUAL walks each list once with the neighbours' annual
   direct + diffuse totals in place of one hour's values (the reflection is linear in them
   and its weights do not depend on the hour, so the annual sums are the same).
   error_mlp (may be NULL): a network of the same shape whose first output is log10 of the
   predicted relative error of each voxel's surrogate total; voxels whose predicted error
   exceeds gate->tau go to the exact path (the gate's other fields are then unused).  Without
   it the gate scores each voxel from the surrogate's own outputs. */
#define SGPU_FALLBACK_ANNUAL (-1)
int sgpu_surrogate_run(sgpu_ctx* ctx, sgpu_mlp* mlp, sgpu_mlp* error_mlp, const sgpu_gate* gate, int hours_per_pass,
                       float* out,
                       int32_t* fallback, sgpu_surrogate_stats* stats);

/* ---- batched tiles (new; no reference counterpart: the reference runs one DSM per call,
   solar_gpu_api.h:37-54) ----
   Many DSM tiles of different sizes and sites in one object: stage 1 is one persistent DDA
   launch per voxel chunk over all tiles (workgroups take (256-voxel block, 32-direction)
   items from a global counter) plus one list launch per chunk; stage 2 is one solar and one
   reflection launch per pass of K hours covering every tile, the whole sunlit-hour loop
   recorded as one hipGraph.  Each tile's lists, SVF bits and annual sums are the ones a
   context of its own computes (bit-identical). */
typedef struct sgpu_batch sgpu_batch;
typedef struct sgpu_batch_info {
    int32_t tiles;
    int32_t chunks;              /* stage-1 voxel chunks */
    int64_t voxels;              /* all tiles */
    int64_t nnz;                 /* neighbour ids, all tiles */
    int64_t barle_bytes;
    int32_t hours_per_pass, passes;
    int64_t voxel_hours;         /* sum over tiles of voxels x sunlit hours */
    double t_viewshed_ms, t_trace_ms, t_lists_ms, t_encode_ms, t_hours_ms;
} sgpu_batch_info;
int sgpu_batch_create(int device, sgpu_batch** out);
int sgpu_batch_destroy(sgpu_batch* b);
/* Append one tile: the arrays of sgpu_set_scene plus its sky (sgpu_set_sky); index = its number. */
int sgpu_batch_add_tile(sgpu_batch* b, const float* grid, int w, int h, int64_t num_total, int64_t num_top,
                        const int32_t* pixel_parent, const float* wall_z, float max_height, const float* slopes,
                        const float* aspects, const float* az_alt, const float* dni, const float* dhi,
                        int32_t* index);
/* Stage 1 for every tile: SVF bits, CSR lists and ba-RLE streams. */
int sgpu_batch_viewshed(sgpu_batch* b);
/* Stage 2 for every tile (format CSR or BARLE, K hours per pass, 2 or 4 waves per reflection
   workgroup); graph != 0 records the loop as one hipGraph and launches that. */
int sgpu_batch_hour_loop(sgpu_batch* b, unsigned format, int hours_per_pass, int waves, int graph);
/* Reflection-pass receiver assignment of the batched hour loop (new): 1 = waves take receivers one
   by one from per-XCD queues over every tile's positions; 0 (the default) = static blocks of the
   concatenated tiles; both give bit-identical annual sums.  2 = waves take pairs of consecutive
   receivers of a tile (ba-RLE lists, 16 or more hours per pass; otherwise as 1), each tile's sums
   equal to those of a context of its own with receiver_queue 2. */
int sgpu_batch_set_receiver_queue(sgpu_batch* b, int on);
int sgpu_batch_get_info(sgpu_batch* b, sgpu_batch_info* out);
/* Per tile: sizes, then the reference-layout products (offsets rebased to the tile). */
int sgpu_batch_tile_sizes(sgpu_batch* b, int32_t tile, int64_t* num_total, int64_t* nnz, int64_t* barle_bytes,
                          int32_t* svf_bytes);
int sgpu_batch_get_svf(sgpu_batch* b, int32_t tile, uint8_t* svf);
int sgpu_batch_get_csr(sgpu_batch* b, int32_t tile, int64_t* offsets, int32_t* ids);
int sgpu_batch_get_barle(sgpu_batch* b, int32_t tile, int64_t* offsets, uint8_t* bytes);
int sgpu_batch_get_irradiance(sgpu_batch* b, int32_t tile, float* out);

/* ---- drop-in whole pipelines (solar_gpu_api.h:37-54) ----
   sun_az_alt: 2*8760 floats; dni/dhi: 8760 floats each. */
int sgpu_uncompressed_irradiance(const float* grid_data, int grid_width, int grid_height, int num_total_voxels,
                                 int num_top_voxels, const int32_t* pixel_parent, const float* wall_z,
                                 float max_height, const float* sun_az_alt, const float* dni, const float* dhi,
                                 const float* slopes, const float* aspects, float* irradiance_out);
int sgpu_compressed_irradiance(const float* grid_data, int grid_width, int grid_height, int num_total_voxels,
                               int num_top_voxels, const int32_t* pixel_parent, const float* wall_z,
                               float max_height, const float* sun_az_alt, const float* dni, const float* dhi,
                               const float* slopes, const float* aspects, float* irradiance_out);

#ifdef __cplusplus
}
#endif
#endif /* SGPU_H */
