/* The per-context device state behind the C ABI and the helpers shared by
 * the ABI translation units (sgpu_capi.hip, sgpu_surrogate.hip). */
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "sgpu_internal.h"

namespace sgpu {

/* Message of the last failing C-ABI call on this thread. */
std::string& last_error();

template <typename F>
int guarded(F&& fn) {
    try {
        fn();
        return SGPU_OK;
    } catch (const Failure& f) {
        last_error() = f.what();
        return f.code;
    } catch (const std::bad_alloc&) {
        last_error() = "host allocation failed";
        return SGPU_ERR_OOM;
    } catch (const std::exception& e) {
        last_error() = e.what();
        return SGPU_ERR_HIP;
    }
}

inline void need(bool ok, int code, const char* msg) {
    if (!ok) fail(code, "%s", msg);
}

template <typename T>
T* dalloc(size_t n) {
    void* p = nullptr;
    if (n == 0) n = 1;
    SGPU_HIP(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T*>(p);
}

template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    EventPair() {
        SGPU_HIP(hipEventCreate(&a));
        SGPU_HIP(hipEventCreate(&b));
    }
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    EventPair(EventPair&& o) noexcept : a(o.a), b(o.b) { o.a = o.b = nullptr; }
    float ms() const {
        float t = 0.0f;
        SGPU_HIP(hipEventElapsedTime(&t, a, b));
        return t;
    }
};

/* Neighbour lists kept in pinned host memory, one piece per viewshed chunk (voxels
   [v0, v0+count), chunk-local int64 offsets), streamed to the GPU tile by tile in stage 2. */
struct HostPiece {
    int64_t v0 = 0, count = 0;
    int64_t* off = nullptr;   // count + 1, pinned
    uint8_t* data = nullptr;  // int32 ids (CSR) or stream bytes (ba-RLE), pinned
    int64_t bytes = 0;
};
struct HostLists {
    std::vector<HostPiece> csr, barle;
    int64_t max_count = 0, max_bytes_csr = 0, max_bytes_barle = 0;
    int64_t total_csr = 0, total_barle = 0;
    static void drop(std::vector<HostPiece>& v) {
        for (HostPiece& p : v) {
            if (p.off) (void)hipHostFree(p.off);
            if (p.data) (void)hipHostFree(p.data);
        }
        v.clear();
    }
    void release() {
        drop(csr);
        drop(barle);
        max_count = max_bytes_csr = max_bytes_barle = total_csr = total_barle = 0;
    }
};

/* Surrogate-path buffers of a context (sgpu_surrogate.hip). */
struct SurrogateState {
    uint16_t* X = nullptr;       // features [N][64] bf16
    float* dd = nullptr;         // annual direct + diffuse
    float* diff = nullptr;       // annual diffuse
    uint16_t* sun_hits = nullptr;
    int16_t* sun_dirs = nullptr;
    int2* hz = nullptr;          // horizon march results of a chunk of voxels
    int64_t hz_cap = 0;          // voxels it holds
    bool ready = false;
    void release() {
        dfree(hz);
        hz_cap = 0;
        dfree(X);
        dfree(dd);
        dfree(diff);
        dfree(sun_hits);
        dfree(sun_dirs);
        ready = false;
    }
};

}  // namespace sgpu

struct sgpu_ctx {
    int device = 0;
    hipStream_t stream = nullptr;

    int W = 0, H = 0;
    int64_t n_total = 0, n_top = 0;
    float maxz = 0.0f;
    float* grid = nullptr;
    int32_t* pix_parent = nullptr;
    float* wall_z = nullptr;
    float* slopes = nullptr;
    float* aspects = nullptr;
    int32_t* order = nullptr;    // reflection-pass processing order (Hilbert-ordered cell tiles of kOrderTile, walls after their cell)
    int32_t* trace_perm = nullptr;   // DDA lane order (blocks of kTraceOrderBlock ids sorted by height)
    int32_t* class_perm = nullptr;   // step-table lane order (blocks of kClassBlock ids sorted by traversal class)
    std::vector<sgpu::ClassRun> class_runs;   // its runs of one class
    bool trace_tables = false;       // sgpu_set_trace_tables: whole-scene viewsheds by the step-table trace
    int64_t pend_cap_override = 0;   // pending-ray capacity per chunk (0: chunk x kDirs / 8)
    int table_dirs = 4, table_steps = 255;   // step-table shape (sgpu_set_trace_table_shape)
    bool table_skip = true;                  // skip windows no cell of which can be hit (max maps)
    float* max_maps = nullptr;               // the four quadrant max maps of the heightmap
    bool scene = false;

    bool sky_set = false;
    sgpu::SkyTables sky;
    std::vector<sgpu::HourConst> hours;
    sgpu::DirConst* dirs = nullptr;
    float lat = 0.0f;
    float cell_m = 1.0f;     // DSM cell size of the scene (surrogate feature)

    uint8_t* svf = nullptr;
    bool svf_ok = false;
    int64_t* csr_off = nullptr;
    int32_t* csr_ids = nullptr;
    int64_t nnz = -1;
    int64_t* br_off = nullptr;
    uint8_t* br_bytes = nullptr;
    int64_t br_size = -1;

    float* sum = nullptr;
    float* comp = nullptr;
    float* hourly_top = nullptr;
    size_t hourly_top_elems = 0;
    int64_t hourly_limit = 0;          // sgpu_set_hourly_limit: floats the passes may use (0 = all)
    sgpu::HourConst* hours_d = nullptr;
    size_t hours_d_cap = 0;
    sgpu::HourConst* hours_all_d = nullptr;   // every sunlit hour of the sky, uploaded by sgpu_set_sky
    size_t hours_all_n = 0;
    void* noalloc_scratch = nullptr;          // this device's allocation-free scratch (sgpu_noalloc.hip)
    bool irr_ok = false;

    /* Viewshed workspace for chunks of up to ws_chunk voxels (raw first-hit table, per-voxel
       counts and offsets, scan scratch), kept between viewsheds of the scene: a 2^22-voxel
       chunk's table is 43.5 GB, whose allocation is not repeated per call. */
    int32_t* ws_raw = nullptr;
    int32_t* ws_counts = nullptr;
    int64_t* ws_off = nullptr;
    int64_t* ws_scratch = nullptr;
    int64_t ws_chunk = 0;
    size_t ws_bytes = 0;
    /* the step-table trace's pending rays (8 bytes each), their counter and a saved step count,
       and the classes' resume states */
    uint2* ws_pend = nullptr;
    int64_t ws_pend_cap = 0;
    unsigned long long* ws_pend_count = nullptr;   // [0] pending rays, [1] saved step count
    sgpu::TableResume* ws_resume = nullptr;
    int4* ws_rec = nullptr;   // voxel records of a chunk in the class order
    int64_t ws_rec_cap = 0;

    int64_t chunk_override = 0;
    bool kernel_timing = false;
    bool count_steps = false;
    sgpu_info info{};
    sgpu::SurrogateState sur;
    /* trace group, hours/pass, waves, XCD band, group; from_cache, rates, receiver queue: the
       shape the autotuner picks on the 2048^2 scene, until a scene of its own is tuned */
    sgpu_tuning tune{32, 128, 4, 32, 0, 0, 0.0, 0.0, 1};
    unsigned* reflect_queue = nullptr;                   // kXcds counters of the queued reflection
    int2* grp_scratch = nullptr;                         // kGroupCap entries per wave, paired reflection
    float* grp_own = nullptr;                            // own irradiance per position over a pass
    int64_t grp_waves = 0;                               // waves grp_scratch covers
    int group_limit = 0;                                 // kept entries walked as a pair at most (0: any)
    int residency = SGPU_LISTS_DEVICE;          // where sgpu_viewshed leaves the lists
    sgpu::HostLists host;                       // the lists when host-resident
    hipStream_t copy_stream = nullptr;          // H2D tile copies of the streamed hour loop

    void set_device() const { SGPU_HIP(hipSetDevice(device)); }

    void drop_products() {
        sgpu::dfree(svf);
        sgpu::dfree(csr_off);
        sgpu::dfree(csr_ids);
        sgpu::dfree(br_off);
        sgpu::dfree(br_bytes);
        svf_ok = false;
        nnz = -1;
        br_size = -1;
        irr_ok = false;
        sur.release();
        host.release();
    }
    void drop_workspace() {
        sgpu::dfree(ws_raw);
        sgpu::dfree(ws_counts);
        sgpu::dfree(ws_off);
        sgpu::dfree(ws_scratch);
        sgpu::dfree(ws_pend);
        sgpu::dfree(ws_pend_count);
        sgpu::dfree(ws_resume);
        sgpu::dfree(ws_rec);
        ws_rec_cap = 0;
        ws_pend_cap = 0;
        ws_chunk = 0;
        ws_bytes = 0;
    }
    void drop_scene() {
        drop_workspace();
        sgpu::dfree(grid);
        sgpu::dfree(pix_parent);
        sgpu::dfree(wall_z);
        sgpu::dfree(slopes);
        sgpu::dfree(aspects);
        sgpu::dfree(order);
        sgpu::dfree(trace_perm);
        sgpu::dfree(class_perm);
        sgpu::dfree(max_maps);
        class_runs.clear();
        sgpu::dfree(sum);
        sgpu::dfree(comp);
        sgpu::dfree(hourly_top);
        sgpu::dfree(reflect_queue);
        sgpu::dfree(grp_scratch);
        sgpu::dfree(grp_own);
        grp_waves = 0;
        hourly_top_elems = 0;
        scene = false;
    }
    ~sgpu_ctx() {
        if (device >= 0) (void)hipSetDevice(device);
        drop_products();
        drop_scene();
        sgpu::dfree(dirs);
        sgpu::dfree(hours_d);
        sgpu::dfree(hours_all_d);
        if (stream) (void)hipStreamDestroy(stream);
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
    }
};

namespace sgpu {

/* Shared pieces of the pipeline driver (sgpu_capi.hip) used by the surrogate path and
   the autotuner. */
HourArgs stage2_args(sgpu_ctx* c, unsigned fmt);
void ensure_accumulators(sgpu_ctx* c, int K);
void ensure_group_scratch(sgpu_ctx* c);
void set_group_args(const sgpu_ctx* c, HourArgs& a, int queue);
void ensure_hours_d(sgpu_ctx* c, const std::vector<HourConst>& hs);
void zero_floats(float* p, int64_t n, hipStream_t s);
/* The current device's address of the allocation-free hour loop's scratch array, looked up
   once per device (sgpu_noalloc.hip); called when a context is created. */
void* noalloc_scratch_setup(int device);
/* Viewshed of a voxel subset (ids on the device) into compact CSR lists or, with barle,
   into ba-RLE streams only (encoded chunk by chunk), indexed by subset position. */
struct SubsetLists {
    int64_t* csr_off = nullptr;
    int32_t* csr_ids = nullptr;
    int64_t nnz = 0;
    int64_t* br_off = nullptr;
    uint8_t* br_bytes = nullptr;
    int64_t br_size = 0;
    void release() {
        dfree(csr_off);
        dfree(csr_ids);
        dfree(br_off);
        dfree(br_bytes);
    }
};
SubsetLists subset_viewshed(sgpu_ctx* c, const int32_t* subset, int64_t n, bool barle);
/* The plain per-hour annual loop for a voxel subset; annual totals per position (device). */
void subset_hour_loop(sgpu_ctx* c, const int32_t* subset, int64_t n, const SubsetLists& L, unsigned fmt, int K,
                      float* out);
/* Annual reflected irradiance per voxel from the annual direct + diffuse of the top cells
   (device arrays), by single-bounce linearity. */
void annual_reflected_linear(sgpu_ctx* c, unsigned fmt, const float* dd, float* out_dev);
/* Apply the autotuner's cached launch shapes for this scene class, if any (no measuring). */
bool apply_cached_tuning(sgpu_ctx* c, unsigned fmt);
/* ba-RLE streams of N CSR lists on a stream (one encoding pass into per-voxel slots, or a size
   pass and a write pass; int64 scan); the caller owns *br_off (N + 1) and *br_bytes (*size
   bytes).  scratch: idle device memory the slots may use (the viewshed's raw table once its
   lists are out), else they are allocated when they fit. */
void encode_barle_on(hipStream_t s, const int64_t* off, const int32_t* ids, int64_t N, int64_t** br_off,
                     uint8_t** br_bytes, int64_t* size,
                     uint8_t* scratch = nullptr, size_t scratch_bytes = 0);
/* Same walk over a subset's own lists (positions of `subset`); out: n reflected sums. */
void subset_reflected_linear(sgpu_ctx* c, const int32_t* subset, int64_t n, const SubsetLists& L, unsigned fmt,
                             const float* dd, float* out);

}  // namespace sgpu
