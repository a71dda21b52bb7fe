/* C ABI of the batched-tile path (include/sgpu.h, "batched tiles").
 *
 * The tiles' voxels are concatenated into one index space (tile t owns
 * [vbase_t, vbase_t + N_t)); every tile keeps its own heightmap, sky tables,
 * direction table (its dense SVF bins) and SVF bitmap.
 *   stage 1  per voxel chunk of the concatenated space: one persistent DDA
 *            launch over all tiles, the list count pass (SVF bits), an int64
 *            scan, the list write pass; then one ba-RLE encode of all lists.
 *            Lists are stored batch-wide (global offsets, tile-local cell ids).
 *   stage 2  per pass of K hours: one k_solar_top launch and one k_reflect
 *            launch over every tile (tile found per workgroup); the loop is
 *            recorded once as a hipGraph.
 * A tile's lists, SVF bits and annual sums equal those of a context of its own:
 * the per-voxel arithmetic is the same code on the same inputs. */
#include <cstring>
#include <vector>

#include "sgpu_ctx.h"
#include "sgpu_internal.h"

using namespace sgpu;

namespace {

__global__ void k_rebase(const int64_t* src, int64_t n, int64_t base, int64_t* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + base;
}

struct BatchTile {
    int W = 0, H = 0;
    int64_t n_total = 0, n_top = 0;
    float maxz = 0.0f;
    int64_t vbase = 0, top_base = 0;
    SkyTables sky;
    std::vector<HourConst> hours;
    float* grid = nullptr;
    int32_t* pix_parent = nullptr;
    float* wall_z = nullptr;
    float* slopes = nullptr;
    float* aspects = nullptr;
    DirConst* dirs = nullptr;
    HourConst* hours_d = nullptr;
    uint8_t* svf = nullptr;
    int32_t* order = nullptr;   // reflection-pass processing order (Hilbert-ordered cell tiles, walls after their cell)
    int32_t* perm = nullptr;    // DDA lane order (blocks of kTraceOrderBlock ids sorted by height)

    void release() {
        dfree(order);
        dfree(perm);
        dfree(grid);
        dfree(pix_parent);
        dfree(wall_z);
        dfree(slopes);
        dfree(aspects);
        dfree(dirs);
        dfree(hours_d);
        dfree(svf);
    }
};

}  // namespace

struct sgpu_batch {
    int device = 0;
    int cus = 0;
    hipStream_t stream = nullptr;
    std::vector<BatchTile> tiles;
    int64_t n_all = 0, top_all = 0;
    bool lists = false;
    int64_t* csr_off = nullptr;   // n_all + 1, batch-wide
    int32_t* csr_ids = nullptr;
    int64_t nnz = 0;
    int64_t* br_off = nullptr;
    uint8_t* br_bytes = nullptr;
    int64_t br_size = 0;
    float* sum = nullptr;
    float* comp = nullptr;
    float* hourly_top = nullptr;
    size_t hourly_top_elems = 0;
    unsigned* queue = nullptr;   // receiver-queue counters of the queued reflection (kQueueWords)
    int queue_mode = 0;          // 0: static blocks (faster than 1 on 64 tiles at K = 256: 6.45 vs 6.59 s);
                                 // 1: queued over all tiles' positions; 2: queued receiver pairs
    int2* pair_scratch = nullptr;   // receiver pairs: kGroupCap entries per wave
    int64_t pair_waves = 0;
    float* own = nullptr;           // receiver pairs: each voxel's own irradiance of the pass (n_all)
    bool irr_ok = false;
    sgpu_batch_info info{};

    void set_device() const { SGPU_HIP(hipSetDevice(device)); }
    void drop_lists() {
        dfree(csr_off);
        dfree(csr_ids);
        dfree(br_off);
        dfree(br_bytes);
        lists = false;
        irr_ok = false;
    }
    ~sgpu_batch() {
        (void)hipSetDevice(device);
        drop_lists();
        dfree(sum);
        dfree(comp);
        dfree(hourly_top);
        dfree(queue);
        dfree(pair_scratch);
        dfree(own);
        for (auto& t : tiles) t.release();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

void add_tile(sgpu_batch* b, const float* grid, int w, int h, int64_t n_total, int64_t n_top, const int32_t* pp,
              const float* wall_z, float maxz, const float* slopes, const float* aspects, const float* az_alt,
              const float* dni, const float* dhi, int32_t* index) {
    need(grid && pp && slopes && aspects && az_alt && dni && dhi, SGPU_ERR_ARG, "null argument");
    need(w > 0 && h > 0 && grid_supported((unsigned)w, (unsigned)h), SGPU_ERR_ARG,
         "grid size out of range (width x (height + 2) must not exceed 2^30 cells, height 2^22 rows)");
    need(n_top == (int64_t)w * h, SGPU_ERR_ARG, "num_top_voxels must equal grid_width * grid_height");
    need(n_total >= n_top && n_total < (int64_t)1 << 31, SGPU_ERR_ARG, "num_total_voxels out of range");
    need(n_total == n_top || wall_z != nullptr, SGPU_ERR_ARG, "wall_z required when there are wall voxels");
    need(b->n_all + n_total < (int64_t)1 << 31, SGPU_ERR_ARG, "batch exceeds 2^31 voxels");
    for (int64_t i = 0; i < n_total; ++i)
        if (pp[i] < 0 || pp[i] >= n_top) fail(SGPU_ERR_ARG, "pixel_parent[%lld] out of range", (long long)i);
    b->drop_lists();
    BatchTile t;
    t.W = w;
    t.H = h;
    t.n_total = n_total;
    t.n_top = n_top;
    t.maxz = maxz;
    t.vbase = b->n_all;
    t.top_base = b->top_all;
    t.sky = sky_bins(az_alt);
    need(t.sky.svf_bytes <= kAzBins * kElBins / 8 + 1, SGPU_ERR_ARG, "sky has too many sun bins");
    t.hours = hour_constants(t.sky, dni, dhi);
    try {
        const int64_t nw = n_total - n_top;
        t.grid = dalloc<float>(n_top);
        t.pix_parent = dalloc<int32_t>(n_total);
        t.wall_z = dalloc<float>(std::max<int64_t>(nw, 1));
        t.slopes = dalloc<float>(n_total);
        t.aspects = dalloc<float>(n_total);
        t.dirs = dalloc<DirConst>(kDirs);
        t.hours_d = dalloc<HourConst>(std::max<size_t>(t.hours.size(), 1));
        t.svf = dalloc<uint8_t>((size_t)std::max(t.sky.svf_bytes, 1) * n_total);
        t.order = dalloc<int32_t>(n_total);
        std::vector<DirConst> d(kDirs);
        directions(t.sky, d.data());
        hipStream_t s = b->stream;
        SGPU_HIP(hipMemcpyAsync(t.grid, grid, n_top * sizeof(float), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(t.pix_parent, pp, n_total * sizeof(int32_t), hipMemcpyHostToDevice, s));
        if (nw > 0) SGPU_HIP(hipMemcpyAsync(t.wall_z, wall_z, nw * sizeof(float), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(t.slopes, slopes, n_total * sizeof(float), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(t.aspects, aspects, n_total * sizeof(float), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipMemcpyAsync(t.dirs, d.data(), kDirs * sizeof(DirConst), hipMemcpyHostToDevice, s));
        if (!t.hours.empty())
            SGPU_HIP(hipMemcpyAsync(t.hours_d, t.hours.data(), t.hours.size() * sizeof(HourConst),
                                    hipMemcpyHostToDevice, s));
        const std::vector<int32_t> order = processing_order(w, h, n_total, pp, kOrderTile);
        SGPU_HIP(hipMemcpyAsync(t.order, order.data(), n_total * sizeof(int32_t), hipMemcpyHostToDevice, s));
        t.perm = dalloc<int32_t>(n_total);
        const std::vector<int32_t> perm = trace_order(n_total, grid, pp, wall_z, n_top, kTraceOrderBlock);
        SGPU_HIP(hipMemcpyAsync(t.perm, perm.data(), n_total * sizeof(int32_t), hipMemcpyHostToDevice, s));
        SGPU_HIP(hipStreamSynchronize(s));
    } catch (...) {
        t.release();
        throw;
    }
    b->n_all += n_total;
    b->top_all += n_top;
    b->tiles.push_back(std::move(t));
    if (index) *index = (int32_t)b->tiles.size() - 1;
}

int64_t batch_chunk(sgpu_batch* b) {
    size_t free_b = 0, total_b = 0;
    SGPU_HIP(hipMemGetInfo(&free_b, &total_b));
    const int64_t by_mem = (int64_t)(0.4 * (double)free_b / ((double)kDirs * sizeof(int32_t)));
    return std::max<int64_t>(kTraceOrderBlock, std::min<int64_t>({b->n_all, by_mem, (int64_t)4 << 20}));
}

/* A chunk of the batch's voxel space [c0, c1) and its slices: runs (tile, first tile-local voxel,
   count) of one tile each, cut at multiples of kTraceOrderBlock within a tile so that every lane
   order block lies in one slice. */
struct Slice {
    int tile;
    int64_t v0, count;
};
struct ChunkPlan {
    int64_t c0 = 0, c1 = 0;
    std::vector<Slice> slices;
};
std::vector<ChunkPlan> plan_chunks(const sgpu_batch* b, int64_t cap) {
    std::vector<ChunkPlan> plan(1);
    for (int t = 0; t < (int)b->tiles.size(); ++t) {
        const BatchTile& tile = b->tiles[t];
        for (int64_t v = 0; v < tile.n_total;) {
            ChunkPlan& c = plan.back();
            const int64_t space = cap - (c.c1 - c.c0);
            int64_t take = std::min(tile.n_total - v, space);
            if (take < tile.n_total - v) take -= take % kTraceOrderBlock;
            if (take <= 0) {   // the chunk is full: the next one starts here
                ChunkPlan nc;
                nc.c0 = nc.c1 = c.c1;
                plan.push_back(nc);
                continue;
            }
            c.slices.push_back({t, v, take});
            c.c1 += take;
            v += take;
        }
    }
    return plan;
}

void viewshed(sgpu_batch* b) {
    need(!b->tiles.empty(), SGPU_ERR_STATE, "no tiles");
    b->drop_lists();
    hipStream_t s = b->stream;
    const int nt = (int)b->tiles.size();
    std::vector<TileDesc> td(nt);
    for (int i = 0; i < nt; ++i) {
        const BatchTile& t = b->tiles[i];
        td[i] = TileDesc{t.grid, t.W, t.H, t.n_top, t.n_total, t.pix_parent, t.wall_z, t.maxz, t.dirs, t.svf,
                         t.sky.svf_bytes, t.vbase};
    }
    const int64_t N = b->n_all;
    const int64_t chunk = batch_chunk(b);
    const std::vector<ChunkPlan> plan = plan_chunks(b, chunk);
    size_t max_slices = 0, max_blocks = 0;
    for (const ChunkPlan& c : plan) {
        size_t blocks = 0;
        for (const Slice& sl : c.slices) blocks += (size_t)((sl.count + 255) / 256);
        max_slices = std::max(max_slices, c.slices.size());
        max_blocks = std::max(max_blocks, blocks);
    }
    TileDesc* tiles_d = nullptr;
    int32_t* raw = nullptr;
    int32_t* counts = nullptr;
    int64_t* off = nullptr;
    int64_t* scratch = nullptr;
    unsigned long long* next = nullptr;
    TraceArgs* slices_d = nullptr;
    int32_t* blocks_d = nullptr;   // block_slice then block_local
    std::vector<std::pair<int32_t*, int64_t>> pieces;
    EventPair all, et, el, ee;
    b->info = sgpu_batch_info{};
    try {
        tiles_d = dalloc<TileDesc>(nt);
        SGPU_HIP(hipMemcpyAsync(tiles_d, td.data(), nt * sizeof(TileDesc), hipMemcpyHostToDevice, s));
        raw = dalloc<int32_t>((size_t)chunk * kDirs);
        counts = dalloc<int32_t>(chunk);
        off = dalloc<int64_t>(chunk + 1);
        scratch = dalloc<int64_t>(scan_scratch_elems(chunk));
        next = dalloc<unsigned long long>(1);
        slices_d = dalloc<TraceArgs>(std::max<size_t>(max_slices, 1));
        blocks_d = dalloc<int32_t>(2 * std::max<size_t>(max_blocks, 1));
        b->csr_off = dalloc<int64_t>(N + 1);
        SGPU_HIP(hipEventRecord(all.a, s));
        int64_t base = 0;
        for (const ChunkPlan& cp : plan) {
            const int64_t v0 = cp.c0, cnt = cp.c1 - cp.c0;
            if (cnt <= 0) continue;
            std::vector<TraceArgs> sl;
            std::vector<int32_t> bs, bl;
            for (const Slice& x : cp.slices) {
                const BatchTile& t = b->tiles[x.tile];
                TraceArgs a{};
                a.grid = t.grid;
                a.W = t.W;
                a.H = t.H;
                a.n_top = t.n_top;
                a.pix_parent = t.pix_parent;
                a.wall_z = t.wall_z;
                a.maxz = t.maxz;
                a.dirs = t.dirs;
                a.v0 = x.v0;
                a.count = x.count;
                a.raw = raw + (t.vbase + x.v0 - v0) * kDirs;
                a.group = kPersistentDirs;
                a.perm = t.perm;
                for (int64_t k = 0; k < (x.count + 255) / 256; ++k) {
                    bs.push_back((int32_t)sl.size());
                    bl.push_back((int32_t)k);
                }
                sl.push_back(a);
            }
            const int64_t nblocks = (int64_t)bs.size();
            bs.insert(bs.end(), bl.begin(), bl.end());
            SGPU_HIP(hipMemcpyAsync(slices_d, sl.data(), sl.size() * sizeof(TraceArgs), hipMemcpyHostToDevice, s));
            SGPU_HIP(hipMemcpyAsync(blocks_d, bs.data(), bs.size() * sizeof(int32_t), hipMemcpyHostToDevice, s));
            const int64_t items = nblocks * (kDirs / kPersistentDirs);
            SGPU_HIP(hipMemsetAsync(next, 0, sizeof(unsigned long long), s));
            BatchTraceArgs ta{slices_d, blocks_d, blocks_d + nblocks, next, items};
            SGPU_HIP(hipEventRecord(et.a, s));
            launch_trace_persistent(ta, (int)std::min<int64_t>(items, (int64_t)b->cus * 8), s);
            SGPU_HIP(hipEventRecord(et.b, s));
            BatchListArgs la{tiles_d, nt, raw, v0, cnt, counts, off, nullptr};
            SGPU_HIP(hipEventRecord(el.a, s));
            launch_lists_batch(la, false, s);
            launch_scan_counts(counts, cnt, off, scratch, s);
            int64_t piece = 0;
            SGPU_HIP(hipMemcpyAsync(&piece, off + cnt, sizeof piece, hipMemcpyDeviceToHost, s));
            SGPU_HIP(hipStreamSynchronize(s));
            int32_t* ids = dalloc<int32_t>((size_t)std::max<int64_t>(piece, 1));
            pieces.push_back({ids, piece});
            la.ids = ids;
            launch_lists_batch(la, true, s);
            hipLaunchKernelGGL(k_rebase, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, s, off, cnt, base,
                               b->csr_off + v0);
            SGPU_HIP(hipGetLastError());
            SGPU_HIP(hipEventRecord(el.b, s));
            SGPU_HIP(hipStreamSynchronize(s));
            b->info.t_trace_ms += et.ms();
            b->info.t_lists_ms += el.ms();
            ++b->info.chunks;
            base += piece;
        }
        SGPU_HIP(hipMemcpyAsync(b->csr_off + N, &base, sizeof base, hipMemcpyHostToDevice, s));
        if (pieces.size() == 1) {
            b->csr_ids = pieces[0].first;
            pieces.clear();
        } else {
            b->csr_ids = dalloc<int32_t>((size_t)std::max<int64_t>(base, 1));
            int64_t pos = 0;
            for (auto& p : pieces) {
                if (p.second)
                    SGPU_HIP(hipMemcpyAsync(b->csr_ids + pos, p.first, p.second * sizeof(int32_t),
                                            hipMemcpyDeviceToDevice, s));
                pos += p.second;
            }
            SGPU_HIP(hipStreamSynchronize(s));
            for (auto& p : pieces) dfree(p.first);
            pieces.clear();
        }
        b->nnz = base;
        SGPU_HIP(hipEventRecord(ee.a, s));
        encode_barle_on(s, b->csr_off, b->csr_ids, N, &b->br_off, &b->br_bytes, &b->br_size);
        SGPU_HIP(hipEventRecord(ee.b, s));
        SGPU_HIP(hipEventRecord(all.b, s));
        SGPU_HIP(hipStreamSynchronize(s));
        b->info.t_encode_ms = ee.ms();
        b->info.t_viewshed_ms = all.ms();
    } catch (...) {
        for (auto& p : pieces) dfree(p.first);
        dfree(tiles_d);
        dfree(raw);
        dfree(counts);
        dfree(off);
        dfree(scratch);
        dfree(next);
        dfree(slices_d);
        dfree(blocks_d);
        b->drop_lists();
        throw;
    }
    dfree(tiles_d);
    dfree(raw);
    dfree(counts);
    dfree(off);
    dfree(scratch);
    dfree(next);
    dfree(slices_d);
    dfree(blocks_d);
    b->lists = true;
    b->info.tiles = nt;
    b->info.voxels = N;
    b->info.nnz = b->nnz;
    b->info.barle_bytes = b->br_size;
}

/* receiver group of the batched reflection pass (workgroups whose waves interleave receivers) */
constexpr int kBatchGroup = 32;

void hour_loop(sgpu_batch* b, unsigned fmt, int K, int waves, bool graph) {
    need(b->lists, SGPU_ERR_STATE, "run sgpu_batch_viewshed first");
    need(fmt == SGPU_FMT_CSR || fmt == SGPU_FMT_BARLE, SGPU_ERR_ARG, "format must be CSR or BARLE");
    need(supported_hours_per_pass(K), SGPU_ERR_ARG, "hours_per_pass must be 1, 2, 4, 8, 16, 32, 64, 128 or 256");
    need(waves == 2 || waves == 4, SGPU_ERR_ARG, "waves must be 2 or 4");
    hipStream_t s = b->stream;
    const int nt = (int)b->tiles.size();
    if (!b->sum) b->sum = dalloc<float>(b->n_all);
    if (!b->comp) b->comp = dalloc<float>(b->n_all);
    if (!b->queue) b->queue = dalloc<unsigned>(kQueueWords);
    const size_t elems = (size_t)b->top_all * K;
    if (elems > b->hourly_top_elems) {
        dfree(b->hourly_top);
        b->hourly_top = dalloc<float>(elems);
        b->hourly_top_elems = elems;
    }
    /* receiver pairs: ba-RLE lists, K >= 16 */
    const bool pairs = b->queue_mode == 2 && fmt == SGPU_FMT_BARLE && K >= 16;
    if (pairs && !b->pair_scratch) {
        int cus = 0;
        SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, b->device));
        b->pair_waves = (int64_t)std::max(cus, 1) * kMaxWavesPerCu;
        b->pair_scratch = dalloc<int2>((size_t)b->pair_waves * kGroupCap);
    }
    if (pairs && !b->own) b->own = dalloc<float>((size_t)std::max<int64_t>(b->n_all, 1));
    std::vector<BatchHour> d(nt);
    int64_t blocks = 0, top_blocks = 0, own_blocks = 0, nslots = 0;
    int passes = 0;
    int64_t vh = 0;
    for (int i = 0; i < nt; ++i) {
        const BatchTile& t = b->tiles[i];
        HourArgs a{};
        a.grid = t.grid;
        a.W = t.W;
        a.H = t.H;
        a.n_total = t.n_total;
        a.n_top = t.n_top;
        a.pix_parent = t.pix_parent;
        a.wall_z = t.wall_z;
        a.slopes = t.slopes;
        a.aspects = t.aspects;
        a.svf = t.svf;
        a.hours = t.hours_d;
        a.hourly_top = b->hourly_top + t.top_base * K;
        a.hourly_top_cap = (int64_t)b->hourly_top_elems - t.top_base * K;
        need((int64_t)t.n_top * K <= a.hourly_top_cap, SGPU_ERR_STATE, "batch hourly buffer too small");
        if (fmt == SGPU_FMT_CSR) {
            a.nb_off = b->csr_off + t.vbase;
            a.nb_ids = b->csr_ids;
        } else {
            a.nb_off = b->br_off + t.vbase;
            a.nb_bytes = b->br_bytes;
            a.nb_len = b->br_size;
        }
        a.sum = b->sum + t.vbase;
        a.comp = b->comp + t.vbase;
        a.n_sub = t.n_total;   // positions = the tile's voxel ids
        a.v_base = 0;
        a.waves = waves;
        /* the single-scene receiver assignment within every tile: tile-major order (voxel ids
           index nb_off, sum and comp) and receiver groups, each tile's blocks whole groups */
        a.order = t.order;
        a.group = kBatchGroup;
        if (pairs) a.grp_own = b->own + t.vbase;   // by the tile's voxel ids
        d[i].a = a;
        d[i].nsel = (int)t.hours.size();
        d[i].block0 = blocks;
        d[i].pos0 = t.vbase;
        d[i].top_block0 = top_blocks;
        d[i].own_block0 = own_blocks;
        d[i].pslot0 = nslots;
        own_blocks += (t.n_total + 255) / 256;
        nslots += (t.n_total + 1) / 2;
        blocks += ((t.n_total + 64 * waves - 1) / (64 * waves) + kBatchGroup - 1) / kBatchGroup * kBatchGroup;
        top_blocks += (t.n_top + 255) / 256;
        passes = std::max(passes, (d[i].nsel + K - 1) / K);
        vh += t.n_total * (int64_t)t.hours.size();
    }
    BatchHour* dd = dalloc<BatchHour>(nt);
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    try {
        SGPU_HIP(hipMemcpyAsync(dd, d.data(), nt * sizeof(BatchHour), hipMemcpyHostToDevice, s));
        zero_floats(b->sum, b->n_all, s);
        zero_floats(b->comp, b->n_all, s);
        SGPU_HIP(hipStreamSynchronize(s));
        const BatchPairs bp{b->pair_scratch, b->pair_waves, own_blocks, nslots};
        auto record = [&] {
            for (int p = 0; p < passes; ++p)
                launch_hour_pass_batch(dd, nt, top_blocks, blocks, K, fmt, waves, p,
                                       b->queue_mode != 0 ? b->queue : nullptr, b->n_all, pairs ? &bp : nullptr, s);
        };
        EventPair ev;
        if (graph) {
            SGPU_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            try {
                record();
            } catch (...) {
                (void)hipStreamEndCapture(s, &g);
                throw;
            }
            SGPU_HIP(hipStreamEndCapture(s, &g));
            SGPU_HIP(hipGraphInstantiate(&ge, g, nullptr, nullptr, 0));
            SGPU_HIP(hipEventRecord(ev.a, s));
            SGPU_HIP(hipGraphLaunch(ge, s));
            SGPU_HIP(hipEventRecord(ev.b, s));
        } else {
            SGPU_HIP(hipEventRecord(ev.a, s));
            record();
            SGPU_HIP(hipEventRecord(ev.b, s));
        }
        SGPU_HIP(hipStreamSynchronize(s));
        b->info.t_hours_ms = ev.ms();
    } catch (...) {
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        dfree(dd);
        throw;
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    dfree(dd);
    b->info.hours_per_pass = K;
    b->info.passes = passes;
    b->info.voxel_hours = vh;
    b->irr_ok = true;
}

const BatchTile& tile_of(sgpu_batch* b, int32_t t) {
    need(t >= 0 && t < (int32_t)b->tiles.size(), SGPU_ERR_ARG, "tile index out of range");
    return b->tiles[t];
}

/* The tile's offsets rebased to its first list entry; returns that first entry. */
int64_t rebased_offsets(sgpu_batch* b, const BatchTile& t, const int64_t* off_d, int64_t* out) {
    SGPU_HIP(hipMemcpyAsync(out, off_d + t.vbase, (t.n_total + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                            b->stream));
    SGPU_HIP(hipStreamSynchronize(b->stream));
    const int64_t first = out[0];
    for (int64_t i = 0; i <= t.n_total; ++i) out[i] -= first;
    return first;
}

}  // namespace

extern "C" {

int sgpu_batch_create(int device, sgpu_batch** out) {
    return guarded([&] {
        need(out != nullptr, SGPU_ERR_ARG, "null output");
        int n = 0;
        SGPU_HIP(hipGetDeviceCount(&n));
        need(device >= 0 && device < n, SGPU_ERR_ARG, "no such device");
        auto* b = new sgpu_batch;
        b->device = device;
        try {
            b->set_device();
            hipDeviceProp_t p;
            SGPU_HIP(hipGetDeviceProperties(&p, device));
            b->cus = p.multiProcessorCount;
            SGPU_HIP(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
        } catch (...) {
            delete b;
            throw;
        }
        *out = b;
    });
}

int sgpu_batch_destroy(sgpu_batch* b) {
    return guarded([&] { delete b; });
}

int sgpu_batch_add_tile(sgpu_batch* b, const float* grid, int w, int h, int64_t num_total, int64_t num_top,
                        const int32_t* pixel_parent, const float* wall_z, float max_height, const float* slopes,
                        const float* aspects, const float* az_alt, const float* dni, const float* dhi,
                        int32_t* index) {
    return guarded([&] {
        need(b != nullptr, SGPU_ERR_ARG, "null batch");
        b->set_device();
        add_tile(b, grid, w, h, num_total, num_top, pixel_parent, wall_z, max_height, slopes, aspects, az_alt, dni,
                 dhi, index);
    });
}

int sgpu_batch_viewshed(sgpu_batch* b) {
    return guarded([&] {
        need(b != nullptr, SGPU_ERR_ARG, "null batch");
        b->set_device();
        viewshed(b);
    });
}

int sgpu_batch_hour_loop(sgpu_batch* b, unsigned format, int hours_per_pass, int waves, int graph) {
    return guarded([&] {
        need(b != nullptr, SGPU_ERR_ARG, "null batch");
        b->set_device();
        hour_loop(b, format, hours_per_pass, waves, graph != 0);
    });
}

int sgpu_batch_set_receiver_queue(sgpu_batch* b, int on) {
    return guarded([&] {
        need(b != nullptr, SGPU_ERR_ARG, "null batch");
        need(on >= 0 && on <= 2, SGPU_ERR_ARG, "on must be 0, 1 or 2");
        b->queue_mode = on;
    });
}

int sgpu_batch_get_info(sgpu_batch* b, sgpu_batch_info* out) {
    return guarded([&] {
        need(b && out, SGPU_ERR_ARG, "null argument");
        *out = b->info;
        out->tiles = (int32_t)b->tiles.size();
        out->voxels = b->n_all;
    });
}

int sgpu_batch_tile_sizes(sgpu_batch* b, int32_t tile, int64_t* num_total, int64_t* nnz, int64_t* barle_bytes,
                          int32_t* svf_bytes) {
    return guarded([&] {
        need(b != nullptr, SGPU_ERR_ARG, "null batch");
        b->set_device();
        const BatchTile& t = tile_of(b, tile);
        if (num_total) *num_total = t.n_total;
        if (svf_bytes) *svf_bytes = t.sky.svf_bytes;
        if (nnz || barle_bytes) {
            need(b->lists, SGPU_ERR_STATE, "run sgpu_batch_viewshed first");
            int64_t c[2], r[2];
            SGPU_HIP(hipMemcpyAsync(c, b->csr_off + t.vbase, sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
            SGPU_HIP(hipMemcpyAsync(&c[1], b->csr_off + t.vbase + t.n_total, sizeof(int64_t), hipMemcpyDeviceToHost,
                                    b->stream));
            SGPU_HIP(hipMemcpyAsync(r, b->br_off + t.vbase, sizeof(int64_t), hipMemcpyDeviceToHost, b->stream));
            SGPU_HIP(hipMemcpyAsync(&r[1], b->br_off + t.vbase + t.n_total, sizeof(int64_t), hipMemcpyDeviceToHost,
                                    b->stream));
            SGPU_HIP(hipStreamSynchronize(b->stream));
            if (nnz) *nnz = c[1] - c[0];
            if (barle_bytes) *barle_bytes = r[1] - r[0];
        }
    });
}

int sgpu_batch_get_svf(sgpu_batch* b, int32_t tile, uint8_t* svf) {
    return guarded([&] {
        need(b && svf, SGPU_ERR_ARG, "null argument");
        need(b->lists, SGPU_ERR_STATE, "run sgpu_batch_viewshed first");
        b->set_device();
        const BatchTile& t = tile_of(b, tile);
        SGPU_HIP(hipMemcpyAsync(svf, t.svf, (size_t)t.sky.svf_bytes * t.n_total, hipMemcpyDeviceToHost, b->stream));
        SGPU_HIP(hipStreamSynchronize(b->stream));
    });
}

int sgpu_batch_get_csr(sgpu_batch* b, int32_t tile, int64_t* offsets, int32_t* ids) {
    return guarded([&] {
        need(b && offsets && ids, SGPU_ERR_ARG, "null argument");
        need(b->lists, SGPU_ERR_STATE, "run sgpu_batch_viewshed first");
        b->set_device();
        const BatchTile& t = tile_of(b, tile);
        const int64_t first = rebased_offsets(b, t, b->csr_off, offsets);
        if (offsets[t.n_total] > 0)
            SGPU_HIP(hipMemcpyAsync(ids, b->csr_ids + first, offsets[t.n_total] * sizeof(int32_t),
                                    hipMemcpyDeviceToHost, b->stream));
        SGPU_HIP(hipStreamSynchronize(b->stream));
    });
}

int sgpu_batch_get_barle(sgpu_batch* b, int32_t tile, int64_t* offsets, uint8_t* bytes) {
    return guarded([&] {
        need(b && offsets && bytes, SGPU_ERR_ARG, "null argument");
        need(b->lists, SGPU_ERR_STATE, "run sgpu_batch_viewshed first");
        b->set_device();
        const BatchTile& t = tile_of(b, tile);
        const int64_t first = rebased_offsets(b, t, b->br_off, offsets);
        if (offsets[t.n_total] > 0)
            SGPU_HIP(hipMemcpyAsync(bytes, b->br_bytes + first, offsets[t.n_total], hipMemcpyDeviceToHost, b->stream));
        SGPU_HIP(hipStreamSynchronize(b->stream));
    });
}

int sgpu_batch_get_irradiance(sgpu_batch* b, int32_t tile, float* out) {
    return guarded([&] {
        need(b && out, SGPU_ERR_ARG, "null argument");
        need(b->irr_ok, SGPU_ERR_STATE, "run sgpu_batch_hour_loop first");
        b->set_device();
        const BatchTile& t = tile_of(b, tile);
        SGPU_HIP(hipMemcpyAsync(out, b->sum + t.vbase, t.n_total * sizeof(float), hipMemcpyDeviceToHost, b->stream));
        SGPU_HIP(hipStreamSynchronize(b->stream));
    });
}

}  // extern "C"
