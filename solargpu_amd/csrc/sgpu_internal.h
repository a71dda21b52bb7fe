/* Internal declarations of libsolargpu: host tables, kernel parameter blocks
 * and launcher entry points shared by the host and HIP translation units. */
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "sgpu.h"

namespace sgpu {

constexpr int kHours = SGPU_HOURS_PER_YEAR;
constexpr int kAzBins = 72;            // 360 / 5 deg  (sky_bins.cpp:13)
constexpr int kElBins = 18;            // 90 / 5 deg   (sky_bins.cpp:14)
constexpr int kDirs = kAzBins * kElBins * 2;  // 2592 ray directions (viewshed_stage.cuh:46)
constexpr float kBinDeg = 5.0f;
constexpr float kDegToRad = 0.0174532925f;    // pipeline_math_constants.h:15
constexpr int kMaxHoursPerPass = 256;

/* ---- errors: thrown inside the library, turned into status codes at the C ABI ---- */
struct Failure : std::runtime_error {
    int code;
    Failure(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};
[[noreturn]] void fail(int code, const char* fmt, ...);
void hip_check(hipError_t e, const char* what, const char* file, int line);
#define SGPU_HIP(expr) ::sgpu::hip_check((expr), #expr, __FILE__, __LINE__)

/* ---- host-side contract (host_sky.cpp, host_voxel.cpp, host_synth.cpp) ---- */
void sun_positions(float lon, float lat, int tz, float* az_alt);

struct SkyTables {
    std::vector<float> nudged;          // 2*8760, az/alt - 1e-8f (sky_bins.cpp:17-25)
    std::vector<int32_t> full_to_dense; // 72*18
    std::vector<int32_t> sun_dense;     // 8760
    int active = 0;
    int svf_bytes = 0;                  // ceil(active / 8)
};
SkyTables sky_bins(const float* az_alt);

/* One ray direction, hour-invariant constants of the DDA (viewshed_kernels.cu:27-35). */
struct alignas(16) DirConst {
    float cx, cy, cz;
    float ax, ay;   // fmaxf(fabsf(c), 1e-8f)
    float rx, ry;   // 1.0f / a
    int32_t info;   // bit 0: upper hemisphere; bits 8..: dense SVF bin + 1 (0 = none)
};
void directions(const SkyTables& sky, DirConst* out);   // kDirs entries, reference order

/* Per sunlit-hour constants of the solar kernel (solar_kernel.cu:8-50, 54-77). */
struct alignas(16) HourConst {
    float cos_z, sin_z;                 // cos zenith, sin(acos(cos_z))
    float cos_az, sin_az;               // of the sun azimuth (radians)
    float dni_eff, dhi;                 // fmaxf(DNI, 1e-3f), DHI
    float inv_b, f1, f2;                // 1 / max(cos 85deg, cos_z), Perez F1, F2
    int32_t dense;                      // sun bin (dense SVF index) or -1
    int32_t index;                      // day*24 + hour
    int32_t eps_bin;
};
std::vector<HourConst> hour_constants(const SkyTables& sky, const float* dni, const float* dhi);

struct Walls {
    std::vector<int32_t> parent;
    std::vector<float> z, aspect;
};
Walls detect_walls(const float* heights, int w, int h, float voxel);
void top_slope_aspect(const float* heights, int w, int h, double sx, double sy, float* slope, float* aspect);

void synth_dsm(uint64_t seed, int w, int h, float cell_m, float density, int trees, float* out);
/* Reflection-pass processing order: cells in tiles of tile x tile (tiles and cells in raster
   order), each top voxel followed by the wall voxels it parents. */
std::vector<int32_t> processing_order(int w, int h, int64_t n_total, const int32_t* pix_parent, int tile);
/* Side of the square cell tiles of the reflection pass's processing order.  Simulated over the
   2048^2 bench scene's lists (LRU of 4096 one-KB rows per XCD, 896 receivers in flight), tiles
   of 8 or 16 cells miss 17% less than tiles of 32 and 26% less than tiles of 64. */
constexpr int kOrderTile = 16;
/* The tiles (x, y) of a tw x th tile grid along a Hilbert curve (side: the next power of two),
   so that consecutive tiles are neighbours: the order of processing_order's tiles. */
std::vector<std::pair<int, int>> hilbert_tiles(int tw, int th);
/* DDA processing order: within every block of `block` consecutive voxel ids, the ids sorted
   by (height of the parent cell, voxel height), so a wave's 64 rays of one direction have
   similar lengths.  perm[i] is the voxel traced by lane i; blocks start at id 0. */
std::vector<int32_t> trace_order(int64_t n_total, const float* grid, const int32_t* pix_parent, const float* wall_z,
                                 int64_t n_top, int block);
constexpr int kTraceOrderBlock = 512;
/* Directions per item of the persistent batched trace (a divisor of the 2592). */
constexpr int kPersistentDirs = 32;
void synth_weather(uint64_t seed, const float* az_alt, float* dni, float* dhi);

/* ---- device kernels (k_viewshed.hip, k_barle.hip, k_hourly.hip) ---- */
/* Grids the DDA addresses by 32-bit byte offsets, with room for its column / row wrap-around
   bounds tests (a step or two past the edge must not wrap into range): the supported grids
   (up to 32766 x 32766 cells; voxel counts are below 2^31 anyway). */
__device__ __host__ inline bool grid_off32(unsigned long long w, unsigned long long h) {
    return w * (h + 2ull) <= (1ull << 30);
}
/* Rows of a supported grid.  The reflection pass splits a cell id into (x, y) with a float
   reciprocal of the width and one fix-up step (cell_xy, k_hourly.hip): the estimate's relative
   error is below 3 * 2^-24, so it is off by at most one row while the row index is below 2^22. */
constexpr unsigned kMaxGridRows = 1u << 22;
__device__ __host__ inline bool grid_supported(unsigned long long w, unsigned long long h) {
    return grid_off32(w, h) && h <= kMaxGridRows;
}

struct TraceArgs {
    const float* grid;
    int W, H;
    int64_t n_top;
    const int32_t* pix_parent;
    const float* wall_z;
    float maxz;
    const DirConst* dirs;
    int64_t v0, count;      // voxel range of this chunk (positions in `subset` when given)
    int32_t* raw;           // count x kDirs first-hit ids (+1), 0 = none
    unsigned long long* steps;  // optional DDA step counter
    const int32_t* subset;  // optional voxel id list
    int group;              // directions per workgroup: 16, 32 or 48 (0 = 32)
    /* optional lane order (no subset; v0 and the chunk a multiple of kTraceOrderBlock): lane i
       traces voxel perm[v0 + i] into raw row perm[v0 + i] - v0 */
    const int32_t* perm;
};
void launch_trace(const TraceArgs& a, hipStream_t s);

/* ---- the step-table trace (k_trace_table.hip) ----
 * A cell coordinate g is in binade e when 2^e <= g < 2^(e+1).  When g also lies two or more
 * from both ends of its binade (and e >= 3), g + 2c rounds to g + (2c rounded to the binade's
 * spacing) for every |c| <= 1, so the ray set-up from g (start cell - g, first-edge distances)
 * is the same for every such g: every cell of one (x binade, y binade) class walks the same
 * step sequence along a direction.  Class key (ex - 3) * 16 + (ey - 3); kClassIrregular for
 * the cells near a binade end or below 8. */
constexpr int kClassIrregular = 255;
constexpr int kClassBlock = 65536;   // the class-sorted lane order permutes within blocks of this many ids
inline int class_of_coord(unsigned g) {
    if (g < 8) return -1;
    const int e = 31 - __builtin_clz(g);
    const unsigned lo = 1u << e, hi = (2u << e) - 1u;
    return (g - lo >= 2 && hi - g >= 2 && e <= 17) ? e : -1;
}
inline int class_key(int gx, int gy) {
    const int ex = class_of_coord((unsigned)gx), ey = class_of_coord((unsigned)gy);
    return (ex < 0 || ey < 0) ? kClassIrregular : (ex - 3) * 16 + (ey - 3);
}
/* A run of one class in the class-sorted lane order: positions [pos, pos + count). */
struct ClassRun {
    int64_t pos;
    int32_t count, key;
};
/* Lane order sorted by class within blocks of kClassBlock ids (stable: raster order inside a
   class), and its runs. */
void class_order(int64_t n_total, int W, const int32_t* pix_parent, std::vector<int32_t>& perm,
                 std::vector<ClassRun>& runs);
/* An item of the table trace: chunk positions [pos, pos + count) of one regular class, at most
   kTableItemVoxels (a multiple of the table workgroup's 256 lanes). */
constexpr int kTableItemVoxels = 2048;
struct TableItem {
    int32_t pos, count, key, pad;
};
/* The resume state of a class and direction after its table's steps. */
struct TableResume {
    float tmx, tmy;
    int32_t mx, my;   // x and y steps taken
};
/* table shapes built: directions per workgroup x steps per table (steps a multiple of the walk's
   3-step unroll, below 2^16): 8 x 255, 4 x 255, 4 x 510, 8 x 126, 4 x 126, 16 x 126, 8 x 63 */
constexpr int kTableClasses = 256;
struct TableArgs {
    const float* grid;
    int W, H;
    int64_t n_top;
    const int32_t* pix_parent;
    const float* wall_z;
    float maxz;
    const DirConst* dirs;
    int64_t v0;                  // chunk start (a multiple of kClassBlock)
    const int4* rec;             // per chunk position: (gx | gy << 16, bits of the base z, raw row)
    const TableItem* items;
    int32_t* raw;                // the chunk's first-hit table
    unsigned long long* steps;   // optional step counter
    unsigned long long* pend_count;
    uint2* pend;                 // (raw row, direction | key << 12) of the rays that outlive their table
    int64_t pend_cap;
    TableResume* resume;         // kTableClasses x kDirs
    const float* qmax;           // the four quadrant max maps (map_cells each), or null: no skipping
    int64_t map_cells;
};
void launch_trace_table(const TableArgs& a, int n_items, int dirs, int steps, hipStream_t s);
/* Cells of a quadrant max-map window a side: map q at cell (x, y) holds the highest height of the
   cells (x + sx a, y + sy b), a, b in [0, kMaxMapWindow), sx = +1 when bit 0 of q is set (else -1),
   sy likewise with bit 1. */
constexpr int kMaxMapWindow = 10;
/* The four maps (4 W H floats) from the heightmap; tmp holds 2 W H floats. */
void launch_max_maps(const float* grid, int W, int H, float* tmp, float* qmax, hipStream_t s);
/* The records of chunk positions [0, n) of the class-sorted order (voxel perm[v0 + p]). */
void launch_table_records(const int32_t* perm, int64_t v0, int64_t n, const int32_t* pix_parent, const float* grid,
                          const float* wall_z, int64_t n_top, int W, int4* rec, hipStream_t s);
/* The rest of the chunk's rays by the pooled walk: every direction of the irregular positions,
   then the pending rays from their resume states. */
struct PendingArgs {
    const float* grid;
    int W, H;
    int64_t n_top;
    const int32_t* pix_parent;
    const float* wall_z;
    float maxz;
    const DirConst* dirs;
    int64_t v0;
    const int32_t* perm;
    const int32_t* irregular;    // chunk positions of the irregular voxels
    int64_t n_irregular;
    const uint2* pend;
    int64_t n_pend;
    const TableResume* resume;
    int32_t* raw;
    unsigned long long* steps;
};
void launch_trace_pending(const PendingArgs& a, hipStream_t s);

struct ListArgs {
    int32_t* raw;           // pass 1 leaves each voxel's kept ids at the front of its row
    int64_t v0, count, n_total;
    const DirConst* dirs;
    int svf_bytes;
    uint8_t* svf;           // word-major SVF bitmap (voxel_types.h:28-38); NULL = not written
    int32_t* counts;        // pass 1 output
    const int64_t* offsets; // pass 2 input (chunk-local, count+1)
    int32_t* ids;           // pass 2 output
};
void launch_lists(const ListArgs& a, bool write_pass, hipStream_t s);

/* Batched viewshed over many tiles: tiles' voxels are concatenated into one index space
   (tile t owns [vbase, vbase + n_total)); each tile keeps its own heightmap, sky (dense SVF
   bins in its DirConst table) and SVF bitmap. */
struct TileDesc {
    const float* grid;
    int W, H;
    int64_t n_top, n_total;
    const int32_t* pix_parent;
    const float* wall_z;
    float maxz;
    const DirConst* dirs;
    uint8_t* svf;
    int svf_bytes;
    int64_t vbase;
};
/* One persistent trace launch over a chunk of the batch: slice i is a run of one tile's voxels
   (TraceArgs with the tile's scene, directions and lane order; v0 a multiple of kTraceOrderBlock,
   raw pointing at the slice's first raw row), block k of the chunk is block block_local[k] of
   slice block_slice[k]. */
struct BatchTraceArgs {
    const TraceArgs* slices;
    const int32_t* block_slice;
    const int32_t* block_local;
    unsigned long long* next;   // work-item counter (zeroed before the launch)
    int64_t items;              // blocks x (kDirs / kPersistentDirs) direction groups
};
void launch_trace_persistent(const BatchTraceArgs& a, int workgroups, hipStream_t s);
struct BatchListArgs {
    const TileDesc* tiles;
    int ntiles;
    int32_t* raw;
    int64_t v0, count;
    int32_t* counts;
    const int64_t* offsets;
    int32_t* ids;
};
void launch_lists_batch(const BatchListArgs& a, bool write_pass, hipStream_t s);

/* Voxel arrays built on the device from a heightmap (k_voxelize.hip): num_total entries each,
   wall_z num_total - num_top (at least one).  The caller owns them (hipFree). */
struct DeviceVoxels {
    int32_t* pix_parent = nullptr;
    float* slopes = nullptr;
    float* aspects = nullptr;
    float* wall_z = nullptr;
    int64_t* wall_off = nullptr;   // num_top + 1: cell c's walls are ids num_top + wall_off[c] ..
};
/* The voxels of the w x h heightmap heights_dev (device) with pixel scales sx, sy (the voxel size
   is sx); returns num_total. */
int64_t voxelize_device(hipStream_t s, const float* heights_dev, int w, int h, double sx, double sy,
                        DeviceVoxels* out);
/* processing_order and trace_order of a scene voxelized by voxelize_device, built on the device
   (identical to the host functions): order and perm get num_total entries (caller-allocated). */
void orders_device(hipStream_t s, const float* heights_dev, int w, int h, const DeviceVoxels& v, int64_t n_total,
                   int tile, int block, int32_t* order, int32_t* perm);

/* Exclusive scan of n int32 counts into n+1 int64 offsets (offsets[n] = total). */
void launch_scan_counts(const int32_t* counts, int64_t n, int64_t* offsets, int64_t* scratch_blocks,
                        hipStream_t s);
size_t scan_scratch_elems(int64_t n);

struct BarleArgs {
    const int64_t* csr_off;
    const int32_t* csr_ids;
    int64_t n;
    int32_t* sizes;         // size pass output
    const int64_t* out_off; // write pass input
    uint8_t* out;           // write pass output
    int32_t* long_ids;      // queues of voxels with more than 512 ids (n entries; filled by the size pass):
                            // up to 1024 ids from the front, more from the back
    unsigned* long_count;   // their lengths, 2 entries (zeroed by the size pass)
    /* One pass instead of size + write: voxel v's stream goes to its slot at byte
       6 csr_off[v] + 8 v of `out` (at most 5.375 n + 3 bytes for n ids), the write pass stores
       `sizes` and queues the long lists itself; launch_barle_compact then packs the slots. */
    bool slots;
};
/* Packs one-pass slots (BarleArgs::slots) into the stream at the scanned offsets. */
void launch_barle_compact(const uint8_t* slots, const int64_t* csr_off, const int32_t* sizes, const int64_t* out_off,
                          int64_t n, uint8_t* out, hipStream_t s);
void launch_barle(const BarleArgs& a, bool write_pass, hipStream_t s);

struct HourArgs {
    const float* grid;
    int W, H;
    int64_t n_total, n_top;
    const int32_t* pix_parent;
    const float* wall_z;
    const float* slopes;
    const float* aspects;
    const uint8_t* svf;
    const HourConst* hours;    // this pass's hours (nh entries)
    int nh;
    float* hourly_top;         // n_top x K, cell-major (row of cell c: row_slot(c, row_swz))
    int64_t hourly_top_cap;    // floats available at hourly_top (the launchers check the rows x K)
    int row_swz;               // 0: row c holds cell c; S > 0: cell c's row is c ^ ((c >> S) & 1023)
    const int64_t* nb_off;     // CSR offsets or ba-RLE byte offsets
    const int32_t* nb_ids;
    const uint8_t* nb_bytes;
    int64_t nb_len;            // bytes in nb_bytes (ba-RLE): the stream window never reads past it
    float* sum;                // Kahan annual sum
    float* comp;               // Kahan compensation
    float* dbg_hourly;         // optional single-hour outputs
    float* dbg_reflected;
    /* Positions p in [0, n_sub) (all n_total voxels when n_sub == 0) index nb_off, sum, comp
       and dbg_*; the voxel of position p is subset[p] if given, else v_base + p. */
    const int32_t* subset;
    int64_t n_sub;
    int64_t v_base;
    int waves;                 // waves per k_reflect workgroup: 2 or 4 (0 = 4)
    int band;                  // k_reflect blocks per XCD band (0 = one contiguous eighth per XCD)
    /* Receiver assignment.  group <= 1: wave w of logical block b takes the 64 consecutive
       positions from (b * waves + w) * 64.  group > 1: the waves of `group` consecutive blocks
       (S = group * waves of them) share 64 * S positions, wave i taking every S-th from i, so
       at any moment the resident waves work on S adjacent positions.  order (full scene only):
       position p is voxel order[p], and voxel ids index nb_off, sum, comp and dbg_*. */
    const int32_t* order;
    int group;
    /* Queued assignment (full-scene order mode, no subset): when queue is set, a persistent
       grid takes the receivers one by one from kQueuesPerXcd counters per XCD (kQueueWords
       unsigned, zeroed by the launcher), XCD x walking bands x, x + 8, ... of queue_band
       positions (0 = one contiguous eighth per XCD). */
    unsigned* queue;
    int64_t queue_band;
    int queue_blocks_per_cu;   // resident workgroups per CU of the queued grid (0 = as many as fit)
    /* Receiver pairs (queued, ba-RLE, K >= 16; grp_size == 2): the queue hands out pairs of
       consecutive positions, each walked by one wave that gathers every kept neighbour row of the
       pair once.  Kept entries beyond the first 240 of a list go to wave w's grp_cap entries of
       grp_scratch, from w * grp_cap; the buffer covers grp_waves waves.  grp_own: each position's
       own irradiance over the pass (by list / sum index, n_total floats), written by the launcher's
       first kernel.  grp_limit > 0: pairs with more kept entries are walked receiver by receiver.
       grp_size <= 1 or no scratch: receivers one by one. */
    int2* grp_scratch;
    float* grp_own;
    int grp_cap;
    int grp_size;
    int grp_limit;
    int64_t grp_waves;
};
/* Entries per wave of the pair scratch (8 bytes each), and the most waves a CU holds at once. */
constexpr int kGroupCap = 4096;
constexpr int kMaxWavesPerCu = 32;
/* Rows of the hourly buffer for n_top cells: swizzled rows stay inside aligned blocks of 1024. */
inline int64_t hourly_rows(int64_t n_top, int row_swz) { return row_swz > 0 ? (n_top + 1023) / 1024 * 1024 : n_top; }
constexpr int kQueuesPerXcd = 16;
constexpr int kQueueStride = 64;                       // unsigned per counter: a 256-byte line each
constexpr int kQueueWords = 8 * kQueuesPerXcd * kQueueStride;   // 8 XCDs: 32 KB
void launch_solar_top(const HourArgs& a, int K, hipStream_t s);

/* Batched tiles in stage 2: one k_solar_top and one k_reflect launch per pass cover every
   tile.  Tile t's arguments address its own scene, sky and lists (positions = its voxel ids,
   n_sub = its voxel count, nb_off/sum/comp pointing at its slice of the batch arrays); its
   reflection blocks start at block0, its solar blocks at top_block0. */
struct BatchHour {
    HourArgs a;          // hours: the tile's first sunlit-hour constant; nh unused
    int nsel;            // the tile's sunlit hours
    int64_t block0, top_block0;
    int64_t pos0;        // the tile's first position in the batch's voxel space (queued reflection)
    int64_t own_block0;  // receiver pairs: the tile's first block of the own-irradiance pass
    int64_t pslot0;      // receiver pairs: the tile's first pair slot (its positions 2k, 2k + 1 = slot pslot0 + k)
};
/* Receiver pairs of the batched hour loop (ba-RLE, K >= 16): the queue hands out pair slots of every
   tile (nslots in all), each walked by one wave with kGroupCap entries of `scratch` (covering
   `waves` waves); each tile's own irradiance (a.grp_own) is written by own_blocks blocks first. */
struct BatchPairs {
    int2* scratch;
    int64_t waves;
    int64_t own_blocks;
    int64_t nslots;
};
/* queue == NULL: static assignment (blocks of the concatenated tiles); else the queued kernel over
   the n_all positions of the batch (kQueueWords counters, zeroed by the launcher), or, with pairs,
   the paired kernel over the pair slots. */
void launch_hour_pass_batch(const BatchHour* d, int ntiles, int64_t top_blocks, int64_t blocks, int K,
                            unsigned fmt, int waves, int pass, unsigned* queue, int64_t n_all,
                            const BatchPairs* pairs, hipStream_t s);
void launch_reflect_accum(const HourArgs& a, int K, unsigned format, hipStream_t s);
bool supported_hours_per_pass(int K);

}  // namespace sgpu
