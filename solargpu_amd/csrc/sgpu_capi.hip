/* C ABI of libsolargpu (include/sgpu.h): per-context device state, the
 * out-of-core viewshed driver, the GPU ba-RLE encoder driver and the
 * sunlit-hour loop.  All work runs on the context's own HIP stream; every
 * failure becomes a status code and a thread-local message. */
#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "sgpu_ctx.h"
#include "sgpu_internal.h"

using namespace sgpu;

namespace sgpu {
std::string& last_error() {
    thread_local std::string msg;
    return msg;
}
}  // namespace sgpu

namespace {

__global__ void k_add_base(const int64_t* src, int64_t n, int64_t base, int64_t* dst) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i] + base;
}

__global__ void k_fill_zero(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

/* Grid-stride float4 copy: the calibration kernel for HBM counters. */
__global__ __launch_bounds__(256) void k_copy_f4(const float4* __restrict__ src, float4* __restrict__ dst,
                                                  int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

/* L2 read-bandwidth probe: every wave streams the whole footprint (n float4, a few MB: resident
 * in each XCD's L2, far beyond a CU's L1) from its own rotated start, 16 bytes per lane and four
 * loads in flight, so one load instruction reads 1 KB of consecutive lines -- the access of one
 * 256-hour row of the reflection pass.  The sum keeps the loads live. */
__global__ __launch_bounds__(256) void k_l2_read(const float4* __restrict__ src, int64_t n, float* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    const int64_t wave = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
    const int64_t rows = n / 64;                         // 1 KB rows
    int64_t r = (wave * 97) % rows;                       // rotated start
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    for (int64_t k = 0; k + 4 <= rows; k += 4) {
        float4 x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            int64_t rr = r + u;
            if (rr >= rows) rr -= rows;
            x[u] = src[rr * 64 + lane];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            acc.x += x[u].x;
            acc.y += x[u].y;
            acc.z += x[u].z;
            acc.w += x[u].w;
        }
        r += 4;
        if (r >= rows) r -= rows;
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = acc.x + acc.y + acc.z + acc.w;
}

/* L1 gather-rate probe: every lane loads 4-byte values from its own 128-byte line of an
   L1-resident buffer (a power-of-two float count `mask` + 1), 8 independent loads per loop turn,
   addresses independent of the data: each wave instruction is 64 L1 tag lookups (the unit of
   TCP_TOTAL_CACHE_ACCESSES), the access pattern of the DDA's pooled height gathers. */
constexpr int kGatherTurns = 64;
__global__ __launch_bounds__(256) void k_gather_probe(const float* __restrict__ buf, unsigned mask, float* out) {
    const unsigned lane = threadIdx.x & 63, wave = blockIdx.x * 4u + (threadIdx.x >> 6);
    const unsigned base = ((wave * 2654435761u + lane * 40503u * 32u) & mask) & ~31u;
    float acc = 0.0f;
    for (int i = 0; i < kGatherTurns; ++i) {
#pragma unroll
        for (int u = 0; u < 8; ++u) acc += buf[(base + (unsigned)(i * 8 + u) * 7919u * 32u) & mask];
    }
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = acc;
}

/* VALU issue-rate probe: eight independent FMA chains per lane, 16 x 8 = 128 FMAs per loop turn
 * (the loop's own counter and branch are scalar instructions), so every wave issues 128 VALU
 * instructions per turn: the calibration kernel for the VALU-bound DDA. */
constexpr int kProbeFmaPerTurn = 128;
__global__ __launch_bounds__(256) void k_valu_probe(float* __restrict__ out, int turns, float m, float a) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) x[k] = (float)(threadIdx.x + k);
    for (int t = 0; t < turns; ++t) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) x[k] = fmaf(x[k], m, a);
        }
    }
    float sum = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += x[k];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

/* The DDA's instruction classes (integer adds, compares into the lane mask, selects): per chain
 * and step y = (x += a) < t ? x : y, which compiles to one v_add_u32, one v_cmp_gt_u32 and one
 * v_cndmask_b32 (the gfx950 assembly of this build has exactly 384 VALU instructions in the loop
 * body: eight chains, 16 steps); the loop's own counter and branch are scalar. */
constexpr int kProbeMixPerTurn = 8 * 16 * 3;
__global__ __launch_bounds__(256) void k_valu_probe_mix(unsigned* __restrict__ out, int turns, unsigned a,
                                                        unsigned t) {
    unsigned x[8], y[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        x[k] = threadIdx.x + (unsigned)k;
        y[k] = threadIdx.x ^ (unsigned)(k * 977);
    }
    for (int r = 0; r < turns; ++r) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                x[k] += a;
                y[k] = (x[k] < t) ? x[k] : y[k];
            }
        }
    }
    unsigned sum = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) sum += x[k] ^ y[k];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = sum;
}

/* Compares into SGPR lane masks, OR-ed by the scalar unit (the DDA's exit tests): per chain and
 * step x += a (v_add_u32) and m |= (x < t) (v_cmp_gt_u32_e64 into an SGPR pair, s_or_b64), with
 * every chain's mask live, so the compares cannot share VCC; 256 VALU instructions per turn. */
constexpr int kProbeCmpPerTurn = 8 * 16 * 2;
__global__ __launch_bounds__(256) void k_valu_probe_cmp(unsigned* __restrict__ out, int turns, unsigned a,
                                                        unsigned t) {
    unsigned x[8];
    uint64_t m[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        x[k] = threadIdx.x * 2654435761u + (unsigned)k;
        m[k] = 0;
    }
    for (int r = 0; r < turns; ++r) {
#pragma unroll
        for (int u = 0; u < 16; ++u) {
#pragma unroll
            for (int k = 0; k < 8; ++k) {
                x[k] += a;
                m[k] |= __ballot(x[k] < t);
            }
        }
    }
    uint64_t any = 0;
#pragma unroll
    for (int k = 0; k < 8; ++k) any ^= m[k];
    out[(int64_t)blockIdx.x * 256 + threadIdx.x] = x[0] ^ (unsigned)any ^ (unsigned)(any >> 32);
}

int64_t choose_chunk(const sgpu_ctx* c, int64_t n) {
    if (n <= 0) return 1;
    if (c->chunk_override > 0) return std::min<int64_t>(c->chunk_override, n);
    /* Raw first-hit table + list counters per voxel; keep 40% of free memory as the
       ceiling (the reference's budget, gpu_memory_budget.cuh:19-35), capped at 2^31
       table entries per chunk. */
    size_t free_b = 0, total_b = 0;
    SGPU_HIP(hipMemGetInfo(&free_b, &total_b));
    free_b += c->ws_bytes;   // the kept workspace is reused or replaced
    /* + the step-table trace's pending rays (up to one in eight, 8 bytes each) and voxel records */
    const double per_voxel = (double)kDirs * 5.0 + 48.0;
    int64_t chunk = (int64_t)(0.4 * (double)free_b / per_voxel);
    chunk = std::min<int64_t>(chunk, (int64_t)1 << 20 << 2);
    /* chunks start on lane-order blocks (kClassBlock is a multiple of kTraceOrderBlock) */
    if (chunk > kClassBlock) chunk -= chunk % kClassBlock;
    else if (chunk > kTraceOrderBlock) chunk -= chunk % kTraceOrderBlock;
    chunk = std::max<int64_t>(chunk, 1);
    return std::min<int64_t>(chunk, n);
}

}  // namespace

namespace sgpu {

void zero_floats(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
    SGPU_HIP(hipGetLastError());
}

void ensure_hours_d(sgpu_ctx* c, const std::vector<HourConst>& hs) {
    if (hs.size() > c->hours_d_cap) {
        dfree(c->hours_d);
        c->hours_d = dalloc<HourConst>(hs.size());
        c->hours_d_cap = hs.size();
    }
    if (!hs.empty())
        SGPU_HIP(hipMemcpyAsync(c->hours_d, hs.data(), hs.size() * sizeof(HourConst), hipMemcpyHostToDevice,
                                c->stream));
}

/* Swizzle shift of the hourly rows for this scene: the bit width of the grid width, so that a
   cell's grid row (cell id / W, up to rounding of W to a power of two) feeds the low bits of its row
   index.  At 2048^2 / K = 256 the reflection pass's L2 hit rate goes from 0.77 to 0.80 and the
   year's hour loop from 2.17 to 2.00 s (same box); at K = 128 it makes no difference. */
int row_swizzle(const sgpu_ctx* c) {
    int s = 0;
    while (s < 30 && ((int64_t)1 << s) < (int64_t)c->W) ++s;
    return s;
}

/* The grouped reflection's per-wave scratch: kGroupCap entries for each wave the device can hold
   at once (kMaxWavesPerCu per CU; 256 CUs: 268 MB). */
void ensure_group_scratch(sgpu_ctx* c) {
    if (c->grp_scratch) return;
    int dev = 0, cus = 0;
    SGPU_HIP(hipGetDevice(&dev));
    SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int64_t waves = (int64_t)std::max(cus, 1) * kMaxWavesPerCu;
    c->grp_scratch = dalloc<int2>((size_t)waves * kGroupCap);
    c->grp_waves = waves;
    c->grp_own = dalloc<float>((size_t)std::max<int64_t>(c->n_total, 1));
}

void set_group_args(const sgpu_ctx* c, HourArgs& a, int queue) {
    a.grp_size = 0;
    a.grp_scratch = nullptr;
    a.grp_own = nullptr;
    a.grp_limit = 0;
    if (queue == 2 && c->grp_scratch) {
        a.grp_size = 2;
        a.grp_scratch = c->grp_scratch;
        a.grp_own = c->grp_own;
        a.grp_cap = kGroupCap;
        a.grp_limit = c->group_limit;
        a.grp_waves = c->grp_waves;
    }
}

void ensure_accumulators(sgpu_ctx* c, int K) {
    if (!c->sum) c->sum = dalloc<float>(c->n_total);
    if (!c->comp) c->comp = dalloc<float>(c->n_total);
    if (!c->reflect_queue) c->reflect_queue = dalloc<unsigned>(kQueueWords);
    if (c->tune.receiver_queue >= 2) ensure_group_scratch(c);
    const size_t need_elems = (size_t)hourly_rows(c->n_top, row_swizzle(c)) * K;
    if (need_elems > c->hourly_top_elems) {
        dfree(c->hourly_top);
        c->hourly_top = dalloc<float>(need_elems);
        c->hourly_top_elems = need_elems;
    }
}

}  // namespace sgpu

namespace {

HourArgs hour_args(sgpu_ctx* c, unsigned fmt) {
    HourArgs a{};
    a.grid = c->grid;
    a.W = c->W;
    a.H = c->H;
    a.n_total = c->n_total;
    a.n_top = c->n_top;
    a.pix_parent = c->pix_parent;
    a.wall_z = c->wall_z;
    a.slopes = c->slopes;
    a.aspects = c->aspects;
    a.svf = c->svf;
    a.hourly_top = c->hourly_top;
    a.hourly_top_cap = (int64_t)c->hourly_top_elems;
    a.row_swz = row_swizzle(c);
    if (c->hourly_limit > 0) a.hourly_top_cap = std::min<int64_t>(a.hourly_top_cap, c->hourly_limit);
    a.sum = c->sum;
    a.comp = c->comp;
    a.waves = c->tune.waves_per_group;
    a.band = c->tune.xcd_band;
    a.order = c->order;   // subset and streamed callers clear it
    a.group = c->tune.group_blocks;
    if (c->tune.receiver_queue && c->reflect_queue) {
        a.queue = c->reflect_queue;
        a.queue_band = (int64_t)c->tune.xcd_band * 64 * (c->tune.waves_per_group == 2 ? 2 : 4);
        a.queue_blocks_per_cu = c->tune.group_blocks;   // group_blocks in queue mode: workgroups per CU
        set_group_args(c, a, c->tune.receiver_queue);
    }
    if (fmt == SGPU_FMT_CSR) {
        a.nb_off = c->csr_off;
        a.nb_ids = c->csr_ids;
    } else {
        a.nb_off = c->br_off;
        a.nb_bytes = c->br_bytes;
        a.nb_len = c->br_size;
    }
    return a;
}

void check_stage2_ready(sgpu_ctx* c, unsigned fmt) {
    need(c->scene, SGPU_ERR_STATE, "scene not set");
    need(c->sky_set, SGPU_ERR_STATE, "sky not set");
    need(c->svf_ok, SGPU_ERR_STATE, "SVF bitmap missing (run sgpu_viewshed or sgpu_set_svf)");
    if (fmt == SGPU_FMT_CSR) need(c->nnz >= 0, SGPU_ERR_STATE, "CSR lists missing");
    else if (fmt == SGPU_FMT_BARLE) need(c->br_size >= 0, SGPU_ERR_STATE, "ba-RLE streams missing");
    else fail(SGPU_ERR_ARG, "format must be SGPU_FMT_CSR or SGPU_FMT_BARLE");
}

int autotuned_hours_per_pass(const sgpu_ctx* c, unsigned fmt) {
    (void)fmt;
    return c->tune.hours_per_pass;
}

/* Selected sunlit hours and the passes over them. */
std::vector<HourConst> select_hours(const sgpu_ctx* c, int stride) {
    std::vector<HourConst> hs;
    for (size_t i = 0; i < c->hours.size(); i += (size_t)std::max(stride, 1)) hs.push_back(c->hours[i]);
    return hs;
}

bool lists_on_host(const sgpu_ctx* c, unsigned fmt) {
    return (fmt == SGPU_FMT_CSR) ? (c->nnz >= 0 && !c->csr_off) : (c->br_size >= 0 && !c->br_off);
}

/* Host-resident lists: per pass, the solar term for every top cell, then the voxel tiles
   (one per viewshed chunk) copied by hipMemcpyAsync on the copy stream into one of two
   device buffers while the reflection kernel works on the other. */
void run_hour_passes_streamed(sgpu_ctx* c, unsigned fmt, int K, int nsel, std::vector<EventPair>* ev_refl,
                              std::vector<EventPair>* ev_solar) {
    const std::vector<HostPiece>& pieces = (fmt == SGPU_FMT_CSR) ? c->host.csr : c->host.barle;
    const int64_t max_bytes = (fmt == SGPU_FMT_CSR) ? c->host.max_bytes_csr : c->host.max_bytes_barle;
    if (!c->copy_stream) SGPU_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
    int64_t* doff[2] = {nullptr, nullptr};
    uint8_t* ddata[2] = {nullptr, nullptr};
    hipEvent_t ready[2] = {nullptr, nullptr}, freed[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        for (int b = 0; b < 2; ++b) {
            dfree(doff[b]);
            dfree(ddata[b]);
            if (ready[b]) (void)hipEventDestroy(ready[b]);
            if (freed[b]) (void)hipEventDestroy(freed[b]);
        }
    };
    try {
        for (int b = 0; b < 2; ++b) {
            doff[b] = dalloc<int64_t>(c->host.max_count + 1);
            ddata[b] = dalloc<uint8_t>((size_t)std::max<int64_t>(max_bytes, 1));
            SGPU_HIP(hipEventCreateWithFlags(&ready[b], hipEventDisableTiming));
            SGPU_HIP(hipEventCreateWithFlags(&freed[b], hipEventDisableTiming));
            SGPU_HIP(hipEventRecord(freed[b], c->stream));
        }
        HourArgs a = hour_args(c, fmt);
        int pass = 0, slot = 0;
        for (int h0 = 0; h0 < nsel; h0 += K, ++pass) {
            a.hours = c->hours_d + h0;
            a.nh = std::min(K, nsel - h0);
            if (ev_solar) SGPU_HIP(hipEventRecord((*ev_solar)[pass].a, c->stream));
            launch_solar_top(a, K, c->stream);
            if (ev_solar) SGPU_HIP(hipEventRecord((*ev_solar)[pass].b, c->stream));
            if (ev_refl) SGPU_HIP(hipEventRecord((*ev_refl)[pass].a, c->stream));
            for (const HostPiece& p : pieces) {
                const int b = slot;
                slot ^= 1;
                SGPU_HIP(hipStreamWaitEvent(c->copy_stream, freed[b], 0));
                SGPU_HIP(hipMemcpyAsync(doff[b], p.off, (size_t)(p.count + 1) * 8, hipMemcpyHostToDevice,
                                        c->copy_stream));
                if (p.bytes)
                    SGPU_HIP(hipMemcpyAsync(ddata[b], p.data, (size_t)p.bytes, hipMemcpyHostToDevice, c->copy_stream));
                SGPU_HIP(hipEventRecord(ready[b], c->copy_stream));
                SGPU_HIP(hipStreamWaitEvent(c->stream, ready[b], 0));
                HourArgs t = a;
                t.v_base = p.v0;
                t.order = nullptr;
                t.n_sub = p.count;
                t.nb_off = doff[b];
                t.nb_ids = reinterpret_cast<const int32_t*>(ddata[b]);
                t.nb_bytes = ddata[b];
                t.nb_len = p.bytes;
                t.sum = c->sum + p.v0;
                t.comp = c->comp + p.v0;
                launch_reflect_accum(t, K, fmt, c->stream);
                SGPU_HIP(hipEventRecord(freed[b], c->stream));
            }
            if (ev_refl) SGPU_HIP(hipEventRecord((*ev_refl)[pass].b, c->stream));
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
        SGPU_HIP(hipStreamSynchronize(c->copy_stream));
    } catch (...) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamSynchronize(c->copy_stream);
        cleanup();
        throw;
    }
    cleanup();
}

void run_hour_passes(sgpu_ctx* c, unsigned fmt, int K, int nsel, std::vector<EventPair>* ev_refl,
                     std::vector<EventPair>* ev_solar) {
    if (lists_on_host(c, fmt)) {
        run_hour_passes_streamed(c, fmt, K, nsel, ev_refl, ev_solar);
        return;
    }
    HourArgs a = hour_args(c, fmt);
    int pass = 0;
    for (int h0 = 0; h0 < nsel; h0 += K, ++pass) {
        a.hours = c->hours_d + h0;
        a.nh = std::min(K, nsel - h0);
        if (ev_solar) SGPU_HIP(hipEventRecord((*ev_solar)[pass].a, c->stream));
        launch_solar_top(a, K, c->stream);
        if (ev_solar) SGPU_HIP(hipEventRecord((*ev_solar)[pass].b, c->stream));
        if (ev_refl) SGPU_HIP(hipEventRecord((*ev_refl)[pass].a, c->stream));
        launch_reflect_accum(a, K, fmt, c->stream);
        if (ev_refl) SGPU_HIP(hipEventRecord((*ev_refl)[pass].b, c->stream));
    }
}

void zero_accumulators(sgpu_ctx* c) {
    zero_floats(c->sum, c->n_total, c->stream);
    zero_floats(c->comp, c->n_total, c->stream);
}

/* Lists of N voxels (voxel ids 0..N-1, or subset[0..N) when given) in out-of-core
 * chunks: trace -> count pass (+ SVF bytes when svf) -> int64 scan -> write pass.
 * Each chunk's compact ids land in their own buffer; one chunk keeps it, several are
 * concatenated.  Offsets are int64, indexed by position. */
struct ListsOut {
    int64_t* off = nullptr;
    int32_t* ids = nullptr;
    int64_t nnz = 0;
    int64_t* br_off = nullptr;   // barle_only builds
    uint8_t* br_bytes = nullptr;
    int64_t br_size = 0;
};

void encode_barle(sgpu_ctx* c, const int64_t* off, const int32_t* ids, int64_t N, int64_t** br_off,
                  uint8_t** br_bytes, int64_t* size);

void* pinned(size_t bytes) {
    void* p = nullptr;
    SGPU_HIP(hipHostMalloc(&p, std::max<size_t>(bytes, 1), hipHostMallocDefault));
    return p;
}

/* Copy one chunk's lists (chunk-local device offsets and ids) into pinned host pieces. */
void host_take(sgpu_ctx* c, bool csr, bool barle, int64_t v0, int64_t cnt, const int64_t* off_d,
               const int32_t* ids_d, int64_t nnz) {
    HostLists& H = c->host;
    H.max_count = std::max(H.max_count, cnt);
    if (csr) {
        H.csr.push_back(HostPiece{v0, cnt, nullptr, nullptr, nnz * 4});
        HostPiece& p = H.csr.back();
        p.off = static_cast<int64_t*>(pinned((size_t)(cnt + 1) * 8));
        p.data = static_cast<uint8_t*>(pinned((size_t)p.bytes));
        SGPU_HIP(hipMemcpyAsync(p.off, off_d, (size_t)(cnt + 1) * 8, hipMemcpyDeviceToHost, c->stream));
        if (p.bytes) SGPU_HIP(hipMemcpyAsync(p.data, ids_d, (size_t)p.bytes, hipMemcpyDeviceToHost, c->stream));
        H.max_bytes_csr = std::max(H.max_bytes_csr, p.bytes);
        H.total_csr += p.bytes;
    }
    if (barle) {
        int64_t* boff = nullptr;
        uint8_t* bytes = nullptr;
        int64_t size = 0;
        encode_barle(c, off_d, ids_d, cnt, &boff, &bytes, &size);
        try {
            H.barle.push_back(HostPiece{v0, cnt, nullptr, nullptr, size});
            HostPiece& p = H.barle.back();
            p.off = static_cast<int64_t*>(pinned((size_t)(cnt + 1) * 8));
            p.data = static_cast<uint8_t*>(pinned((size_t)size));
            SGPU_HIP(hipMemcpyAsync(p.off, boff, (size_t)(cnt + 1) * 8, hipMemcpyDeviceToHost, c->stream));
            if (size) SGPU_HIP(hipMemcpyAsync(p.data, bytes, (size_t)size, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            H.max_bytes_barle = std::max(H.max_bytes_barle, size);
            H.total_barle += size;
        } catch (...) {
            dfree(boff);
            dfree(bytes);
            throw;
        }
        dfree(boff);
        dfree(bytes);
    }
    SGPU_HIP(hipStreamSynchronize(c->stream));
}

/* The step-table trace's class-sorted lane order and its runs, built on the host from the
   scene's parent cells the first time a viewshed uses them. */
void ensure_class_order(sgpu_ctx* c) {
    if (!c->max_maps) {   // the quadrant max maps of the heightmap (skip tests)
        const int64_t cells = (int64_t)c->W * c->H;
        float* tmp = dalloc<float>((size_t)(2 * cells));
        try {
            c->max_maps = dalloc<float>((size_t)(4 * cells));
            launch_max_maps(c->grid, c->W, c->H, tmp, c->max_maps, c->stream);
            SGPU_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            dfree(tmp);
            throw;
        }
        dfree(tmp);
    }
    if (c->class_perm) return;
    std::vector<int32_t> pp((size_t)c->n_total), perm;
    SGPU_HIP(hipMemcpyAsync(pp.data(), c->pix_parent, (size_t)c->n_total * sizeof(int32_t), hipMemcpyDeviceToHost,
                            c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    class_order(c->n_total, c->W, pp.data(), perm, c->class_runs);
    c->class_perm = dalloc<int32_t>(c->n_total);
    SGPU_HIP(hipMemcpyAsync(c->class_perm, perm.data(), c->n_total * sizeof(int32_t), hipMemcpyHostToDevice,
                            c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
}

/* The first-hit table of chunk [v0, v0 + cnt) by the step-table trace: the regular classes' items,
   then the pooled walk over the irregular voxels and the rays that outlived their table.  If the
   pending rays overflow their buffer, the chunk is traced again by the pooled walk alone (the
   step count restored first).  The chunk starts on a class block. */
void trace_chunk_tables(sgpu_ctx* c, int64_t v0, int64_t cnt, int32_t* raw, unsigned long long* steps) {
    std::vector<TableItem> items;
    std::vector<int32_t> irr;
    const auto lo = std::lower_bound(c->class_runs.begin(), c->class_runs.end(), v0,
                                     [](const ClassRun& r, int64_t p) { return r.pos < p; });
    for (auto r = lo; r != c->class_runs.end() && r->pos < v0 + cnt; ++r) {
        const int32_t p0 = (int32_t)(r->pos - v0);
        if (r->key == kClassIrregular) {
            for (int32_t i = 0; i < r->count; ++i) irr.push_back(p0 + i);
        } else {
            for (int32_t o = 0; o < r->count; o += kTableItemVoxels)
                items.push_back(TableItem{p0 + o, std::min(kTableItemVoxels, r->count - o), r->key, 0});
        }
    }
    need(items.size() <= 65535, SGPU_ERR_ARG, "too many step-table items in one chunk");
    TableItem* items_d = dalloc<TableItem>(items.size());
    int32_t* irr_d = dalloc<int32_t>(irr.size());
    try {
        if (!items.empty())
            SGPU_HIP(hipMemcpyAsync(items_d, items.data(), items.size() * sizeof(TableItem), hipMemcpyHostToDevice,
                                    c->stream));
        if (!irr.empty())
            SGPU_HIP(hipMemcpyAsync(irr_d, irr.data(), irr.size() * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipMemsetAsync(c->ws_pend_count, 0, sizeof(unsigned long long), c->stream));
        if (steps)
            SGPU_HIP(hipMemcpyAsync(c->ws_pend_count + 1, steps, sizeof(unsigned long long), hipMemcpyDeviceToDevice,
                                    c->stream));
        const int64_t cap = c->pend_cap_override > 0 ? std::min(c->pend_cap_override, c->ws_pend_cap) : c->ws_pend_cap;
        launch_table_records(c->class_perm, v0, cnt, c->pix_parent, c->grid, c->wall_z, c->n_top, c->W, c->ws_rec,
                             c->stream);
        TableArgs ta{c->grid, c->W,     c->H,   c->n_top,           c->pix_parent, c->wall_z, c->maxz, c->dirs,
                     v0,      c->ws_rec, items_d, raw, steps, c->ws_pend_count, c->ws_pend, cap, c->ws_resume,
                     c->table_skip ? c->max_maps : nullptr, (int64_t)c->W * c->H};
        launch_trace_table(ta, (int)items.size(), c->table_dirs, c->table_steps, c->stream);
        unsigned long long npend = 0;
        SGPU_HIP(hipMemcpyAsync(&npend, c->ws_pend_count, sizeof npend, hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->info.table_chunks += 1;
        c->info.table_irregular_voxels += (int64_t)irr.size();
        if (npend > (unsigned long long)cap) {
            c->info.table_fallbacks += 1;
            if (steps)
                SGPU_HIP(hipMemcpyAsync(steps, c->ws_pend_count + 1, sizeof(unsigned long long),
                                        hipMemcpyDeviceToDevice, c->stream));
            TraceArgs t{c->grid, c->W, c->H, c->n_top, c->pix_parent, c->wall_z, c->maxz, c->dirs, v0, cnt, raw,
                        steps, nullptr, c->tune.dirs_per_group, c->trace_perm};
            launch_trace(t, c->stream);
        } else {
            c->info.table_pending_rays += (int64_t)npend;
            PendingArgs pa{c->grid, c->W,       c->H,       c->n_top,        c->pix_parent,
                           c->wall_z, c->maxz,  c->dirs,    v0,              c->class_perm,
                           irr_d,   (int64_t)irr.size(), c->ws_pend, (int64_t)npend, c->ws_resume,
                           raw,     steps};
            launch_trace_pending(pa, c->stream);
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        dfree(items_d);
        dfree(irr_d);
        throw;
    }
    dfree(items_d);
    dfree(irr_d);
}

/* host_csr / host_barle: send every chunk's lists to pinned host memory instead of keeping
   them on the device (then out.off / out.ids stay null).  barle_only: encode every chunk's
   lists to ba-RLE at once and keep only the streams (out.br_*), so the whole CSR never
   sits on the device (scenes whose CSR exceeds HBM). */
ListsOut build_lists(sgpu_ctx* c, const int32_t* subset, int64_t N, uint8_t* svf, unsigned long long* steps,
                     bool host_csr = false, bool host_barle = false, bool barle_only = false) {
    const bool to_host = host_csr || host_barle;
    const int64_t chunk = choose_chunk(c, N);
    c->info.chunk_voxels = chunk;
    c->info.viewshed_chunks = (int32_t)((N + chunk - 1) / chunk);
    EventPair ev_trace, ev_lists;
    ListsOut out;
    std::vector<std::pair<uint8_t*, int64_t>> bpieces;
    int64_t bbase = 0;
    if (!to_host && !barle_only) out.off = dalloc<int64_t>(N + 1);
    if (barle_only) out.br_off = dalloc<int64_t>(N + 1);
    int32_t* raw = nullptr;
    int32_t* counts = nullptr;
    int64_t* off = nullptr;
    int64_t* scratch = nullptr;
    std::vector<std::pair<int32_t*, int64_t>> pieces;
    int64_t base = 0;
    try {
        if (c->ws_chunk < chunk) {
            c->drop_workspace();
            c->ws_raw = dalloc<int32_t>((size_t)chunk * kDirs);
            c->ws_counts = dalloc<int32_t>(chunk);
            c->ws_off = dalloc<int64_t>(chunk + 1);
            c->ws_scratch = dalloc<int64_t>(scan_scratch_elems(chunk));
            c->ws_chunk = chunk;
            c->ws_bytes = (size_t)chunk * kDirs * 4 + (size_t)chunk * 4 + (size_t)(chunk + 1) * 8 +
                          (size_t)scan_scratch_elems(chunk) * 8;
        }
        raw = c->ws_raw;
        counts = c->ws_counts;
        off = c->ws_off;
        scratch = c->ws_scratch;
        /* the height-sorted lane order permutes within blocks of kTraceOrderBlock ids from 0:
           usable when every chunk starts on a block (whole scene, chunk a multiple of it) */
        const int32_t* perm = (!subset && (chunk >= N || chunk % kTraceOrderBlock == 0)) ? c->trace_perm : nullptr;
        /* the step-table trace likewise needs chunks on class blocks */
        const bool tables = c->trace_tables && !subset && N == c->n_total &&
                            (chunk >= N || chunk % kClassBlock == 0) && c->W < 65536 && c->H < 65536;
        if (tables) {
            ensure_class_order(c);
            const int64_t need_cap = std::max<int64_t>(chunk * kDirs / 8, 1);
            if (c->ws_pend_cap < need_cap) {
                dfree(c->ws_pend);
                c->ws_pend_cap = 0;
                c->ws_pend = dalloc<uint2>((size_t)need_cap);
                c->ws_pend_cap = need_cap;
                c->ws_bytes += (size_t)need_cap * sizeof(uint2);
            }
            if (c->ws_rec_cap < chunk) {
                dfree(c->ws_rec);
                c->ws_rec_cap = 0;
                c->ws_rec = dalloc<int4>((size_t)chunk);
                c->ws_rec_cap = chunk;
                c->ws_bytes += (size_t)chunk * sizeof(int4);
            }
            if (!c->ws_pend_count) c->ws_pend_count = dalloc<unsigned long long>(2);
            if (!c->ws_resume) c->ws_resume = dalloc<TableResume>((size_t)kTableClasses * kDirs);
        }
        for (int64_t v0 = 0; v0 < N; v0 += chunk) {
            const int64_t cnt = std::min(chunk, N - v0);
            SGPU_HIP(hipEventRecord(ev_trace.a, c->stream));
            if (tables) {
                trace_chunk_tables(c, v0, cnt, raw, steps);
            } else {
                TraceArgs ta{c->grid, c->W, c->H, c->n_top, c->pix_parent, c->wall_z, c->maxz, c->dirs, v0, cnt, raw,
                             steps, subset, c->tune.dirs_per_group, perm};
                launch_trace(ta, c->stream);
            }
            SGPU_HIP(hipEventRecord(ev_trace.b, c->stream));
            ListArgs la{raw, v0, cnt, c->n_total, c->dirs, c->sky.svf_bytes, svf, counts, off, nullptr};
            SGPU_HIP(hipEventRecord(ev_lists.a, c->stream));
            launch_lists(la, false, c->stream);
            launch_scan_counts(counts, cnt, off, scratch, c->stream);
            int64_t piece = 0;
            SGPU_HIP(hipMemcpyAsync(&piece, off + cnt, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            int32_t* ids = dalloc<int32_t>((size_t)std::max<int64_t>(piece, 1));
            pieces.push_back({ids, piece});
            la.ids = ids;
            launch_lists(la, true, c->stream);
            if (!to_host && !barle_only) {
                const unsigned g = (unsigned)((cnt + 255) / 256);
                hipLaunchKernelGGL(k_add_base, dim3(g), dim3(256), 0, c->stream, off, cnt, base, out.off + v0);
                SGPU_HIP(hipGetLastError());
            }
            if (barle_only) {
                int64_t* boff = nullptr;
                uint8_t* bytes = nullptr;
                int64_t size = 0;
                encode_barle_on(c->stream, off, ids, cnt, &boff, &bytes, &size, reinterpret_cast<uint8_t*>(raw),
                                (size_t)chunk * kDirs * sizeof(int32_t));   // the chunk's lists are out of raw
                bpieces.push_back({bytes, size});
                const unsigned g = (unsigned)((cnt + 255) / 256);
                hipLaunchKernelGGL(k_add_base, dim3(g), dim3(256), 0, c->stream, boff, cnt, bbase, out.br_off + v0);
                SGPU_HIP(hipGetLastError());
                SGPU_HIP(hipStreamSynchronize(c->stream));
                dfree(boff);
                bbase += size;
            }
            SGPU_HIP(hipEventRecord(ev_lists.b, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            c->info.t_trace_ms += ev_trace.ms();
            c->info.t_lists_ms += ev_lists.ms();
            if (to_host) {
                host_take(c, host_csr, host_barle, v0, cnt, off, ids, piece);
            }
            if (to_host || barle_only) {
                dfree(pieces.back().first);
                pieces.pop_back();
            }
            base += piece;
        }
        if (barle_only) {
            out.nnz = base;
            c->drop_workspace();   // this mode is for scenes that stretch the memory
            SGPU_HIP(hipMemcpyAsync(out.br_off + N, &bbase, sizeof bbase, hipMemcpyHostToDevice, c->stream));
            if (bpieces.size() == 1) {
                out.br_bytes = bpieces[0].first;
                bpieces.clear();
            } else {
                out.br_bytes = dalloc<uint8_t>((size_t)std::max<int64_t>(bbase, 1));
                int64_t pos = 0;
                for (auto& p : bpieces) {
                    if (p.second)
                        SGPU_HIP(hipMemcpyAsync(out.br_bytes + pos, p.first, (size_t)p.second, hipMemcpyDeviceToDevice,
                                                c->stream));
                    pos += p.second;
                    SGPU_HIP(hipStreamSynchronize(c->stream));
                    dfree(p.first);
                }
                bpieces.clear();
            }
            SGPU_HIP(hipStreamSynchronize(c->stream));
            out.br_size = bbase;
            return out;
        }
        if (to_host) {
            out.nnz = base;
            c->drop_workspace();   // this mode is for scenes that stretch the memory
            return out;
        }
        SGPU_HIP(hipMemcpyAsync(out.off + N, &base, sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        if (pieces.size() == 1) {
            out.ids = pieces[0].first;
            pieces.clear();
        } else {
            out.ids = dalloc<int32_t>((size_t)std::max<int64_t>(base, 1));
            int64_t pos = 0;
            for (auto& p : pieces) {
                if (p.second)
                    SGPU_HIP(hipMemcpyAsync(out.ids + pos, p.first, p.second * sizeof(int32_t), hipMemcpyDeviceToDevice,
                                            c->stream));
                pos += p.second;
            }
            SGPU_HIP(hipStreamSynchronize(c->stream));
            for (auto& p : pieces) dfree(p.first);
            pieces.clear();
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        for (auto& p : pieces) dfree(p.first);
        for (auto& p : bpieces) dfree(p.first);
        c->drop_workspace();
        dfree(out.off);
        dfree(out.ids);
        dfree(out.br_off);
        dfree(out.br_bytes);
        throw;
    }
    out.nnz = base;
    return out;
}

void encode_barle(sgpu_ctx* c, const int64_t* off, const int32_t* ids, int64_t N, int64_t** br_off,
                  uint8_t** br_bytes, int64_t* size) {
    /* the viewshed workspace's raw table is idle once the lists are built */
    encode_barle_on(c->stream, off, ids, N, br_off, br_bytes, size, reinterpret_cast<uint8_t*>(c->ws_raw),
                    c->ws_raw ? (size_t)c->ws_chunk * kDirs * sizeof(int32_t) : 0);
}

}  // namespace

/* ba-RLE streams of N CSR lists: size pass -> int64 scan -> write pass. */
void sgpu::encode_barle_on(hipStream_t stream, const int64_t* off, const int32_t* ids, int64_t N, int64_t** br_off,
                           uint8_t** br_bytes, int64_t* size, uint8_t* scratch, size_t scratch_bytes) {
    BarleArgs ba{};
    ba.csr_off = off;
    ba.csr_ids = ids;
    ba.n = N;
    int32_t* sizes = dalloc<int32_t>(N);
    int64_t* scr = nullptr;
    int64_t* boff = nullptr;
    uint8_t* bytes = nullptr;
    uint8_t* slots = nullptr;
    int32_t* long_ids = nullptr;
    unsigned* long_count = nullptr;
    try {
        scr = dalloc<int64_t>(scan_scratch_elems(N));
        boff = dalloc<int64_t>(N + 1);
        long_ids = dalloc<int32_t>(N);
        long_count = dalloc<unsigned>(2);
        ba.long_ids = long_ids;
        ba.long_count = long_count;
        ba.sizes = sizes;
        /* one encoding pass into per-voxel slots (6 B per id + 8 per voxel) in the caller's
           scratch, or allocated when they fit a quarter of the free memory; else a size pass and a
           write pass (SGPU_BARLE_PASSES=2 forces the latter) */
        int64_t nnz = 0;
        SGPU_HIP(hipMemcpyAsync(&nnz, off + N, sizeof nnz, hipMemcpyDeviceToHost, stream));
        SGPU_HIP(hipStreamSynchronize(stream));
        const size_t slot_bytes = (size_t)(6 * nnz + 8 * N + 16);
        const char* passes = getenv("SGPU_BARLE_PASSES");
        const bool two_pass = passes && passes[0] == '2';
        if (!two_pass && scratch && slot_bytes <= scratch_bytes) {
            ba.out = scratch;
            ba.slots = true;
        } else if (!two_pass) {
            size_t free_b = 0, total_b = 0;
            SGPU_HIP(hipMemGetInfo(&free_b, &total_b));
            if (slot_bytes <= free_b / 4) {
                slots = dalloc<uint8_t>(slot_bytes);
                ba.out = slots;
                ba.slots = true;
            }
        }
        launch_barle(ba, ba.slots, stream);
        launch_scan_counts(sizes, N, boff, scr, stream);
        int64_t total = 0;
        SGPU_HIP(hipMemcpyAsync(&total, boff + N, sizeof total, hipMemcpyDeviceToHost, stream));
        SGPU_HIP(hipStreamSynchronize(stream));
        bytes = dalloc<uint8_t>((size_t)std::max<int64_t>(total, 1));
        if (ba.slots) {
            launch_barle_compact(ba.out, off, sizes, boff, N, bytes, stream);
        } else {
            ba.out_off = boff;
            ba.out = bytes;
            launch_barle(ba, true, stream);
        }
        SGPU_HIP(hipStreamSynchronize(stream));
        *size = total;
    } catch (...) {
        dfree(sizes);
        dfree(scr);
        dfree(boff);
        dfree(bytes);
        dfree(slots);
        dfree(long_ids);
        dfree(long_count);
        throw;
    }
    dfree(sizes);
    dfree(scr);
    dfree(slots);
    dfree(long_ids);
    dfree(long_count);
    *br_off = boff;
    *br_bytes = bytes;
}

namespace {

void do_viewshed(sgpu_ctx* c, unsigned formats) {
    need(c->scene, SGPU_ERR_STATE, "scene not set");
    need(c->sky_set, SGPU_ERR_STATE, "sky not set");
    need((formats & ~(SGPU_FMT_CSR | SGPU_FMT_BARLE)) == 0 && formats != 0, SGPU_ERR_ARG, "bad formats");
    c->drop_products();
    const int64_t N = c->n_total;
    c->info.t_trace_ms = c->info.t_lists_ms = 0.0;
    c->info.dda_steps = 0;
    c->info.table_pending_rays = c->info.table_irregular_voxels = 0;
    c->info.table_chunks = c->info.table_fallbacks = 0;
    EventPair all, ev_enc;
    SGPU_HIP(hipEventRecord(all.a, c->stream));
    c->svf = dalloc<uint8_t>((size_t)std::max(c->sky.svf_bytes, 1) * N);
    unsigned long long* steps = c->count_steps ? dalloc<unsigned long long>(1) : nullptr;
    if (steps) SGPU_HIP(hipMemsetAsync(steps, 0, sizeof(unsigned long long), c->stream));
    ListsOut L;
    const bool to_host = c->residency == SGPU_LISTS_HOST;
    const bool barle_only = !to_host && formats == SGPU_FMT_BARLE;
    try {
        L = build_lists(c, nullptr, N, c->svf, steps, to_host && (formats & SGPU_FMT_CSR),
                        to_host && (formats & SGPU_FMT_BARLE), barle_only);
        if (steps) {
            unsigned long long st = 0;
            SGPU_HIP(hipMemcpyAsync(&st, steps, sizeof st, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            c->info.dda_steps = (int64_t)st;
        }
    } catch (...) {
        dfree(steps);
        throw;
    }
    dfree(steps);
    c->svf_ok = true;
    c->info.nnz = L.nnz;
    if (to_host) {
        /* lists live in pinned host pieces; null device pointers mark them host-resident */
        c->nnz = (formats & SGPU_FMT_CSR) ? L.nnz : -1;
        c->br_size = (formats & SGPU_FMT_BARLE) ? c->host.total_barle : -1;
        c->info.barle_bytes = c->host.total_barle;
        SGPU_HIP(hipEventRecord(all.b, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->info.t_viewshed_ms = all.ms();
        return;
    }
    if (barle_only) {
        c->br_off = L.br_off;
        c->br_bytes = L.br_bytes;
        c->br_size = L.br_size;
        c->nnz = -1;   // no CSR kept
        c->info.barle_bytes = L.br_size;
        SGPU_HIP(hipEventRecord(all.b, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->info.t_viewshed_ms = all.ms();
        return;
    }
    c->csr_off = L.off;
    c->csr_ids = L.ids;
    c->nnz = L.nnz;
    if (formats & SGPU_FMT_BARLE) {
        SGPU_HIP(hipEventRecord(ev_enc.a, c->stream));
        encode_barle(c, c->csr_off, c->csr_ids, N, &c->br_off, &c->br_bytes, &c->br_size);
        SGPU_HIP(hipEventRecord(ev_enc.b, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->info.t_lists_ms += ev_enc.ms();
        c->info.barle_bytes = c->br_size;
    }
    SGPU_HIP(hipEventRecord(all.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    c->info.t_viewshed_ms = all.ms();
}

void do_barle_encode(sgpu_ctx* c) {
    need(c->nnz >= 0 && c->csr_off, SGPU_ERR_STATE, "device-resident CSR lists missing");
    dfree(c->br_off);
    dfree(c->br_bytes);
    c->br_size = -1;
    int64_t size = 0;
    encode_barle(c, c->csr_off, c->csr_ids, c->n_total, &c->br_off, &c->br_bytes, &size);
    c->br_size = size;
    c->info.barle_bytes = size;
}

void do_hour_loop(sgpu_ctx* c, unsigned fmt, int K, int stride) {
    check_stage2_ready(c, fmt);
    if (K == 0) K = autotuned_hours_per_pass(c, fmt);
    need(supported_hours_per_pass(K), SGPU_ERR_ARG, "hours_per_pass must be 1, 2, 4, 8, 16, 32, 64, 128 or 256");
    need(stride >= 1, SGPU_ERR_ARG, "hour_stride must be >= 1");
    ensure_accumulators(c, K);
    std::vector<HourConst> hs = select_hours(c, stride);
    ensure_hours_d(c, hs);
    const int nsel = (int)hs.size();
    const int passes = (nsel + K - 1) / K;
    std::vector<EventPair> ev_r, ev_s;
    if (c->kernel_timing) {
        ev_r.resize(passes);
        ev_s.resize(passes);
    }
    EventPair all;
    zero_accumulators(c);
    SGPU_HIP(hipEventRecord(all.a, c->stream));
    run_hour_passes(c, fmt, K, nsel, c->kernel_timing ? &ev_r : nullptr, c->kernel_timing ? &ev_s : nullptr);
    SGPU_HIP(hipEventRecord(all.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    c->info.t_hours_ms = all.ms();
    c->info.hours_computed = nsel;
    c->info.hour_passes = passes;
    c->info.hours_per_pass = K;
    c->info.timed_launches = c->kernel_timing ? passes : 0;
    c->info.t_refl_kernel_ms = 0.0;
    c->info.t_solar_kernel_ms = 0.0;
    for (auto& e : ev_r) c->info.t_refl_kernel_ms += e.ms();
    for (auto& e : ev_s) c->info.t_solar_kernel_ms += e.ms();
    c->irr_ok = true;
}

void do_hour_loop_graph(sgpu_ctx* c, unsigned fmt, int K, int replays) {
    check_stage2_ready(c, fmt);
    need(!lists_on_host(c, fmt), SGPU_ERR_STATE, "graph capture needs device-resident lists");
    if (K == 0) K = autotuned_hours_per_pass(c, fmt);
    need(supported_hours_per_pass(K), SGPU_ERR_ARG, "hours_per_pass must be 1, 2, 4, 8, 16, 32, 64, 128 or 256");
    need(replays >= 1, SGPU_ERR_ARG, "replays must be >= 1");
    ensure_accumulators(c, K);
    ensure_hours_d(c, c->hours);
    SGPU_HIP(hipStreamSynchronize(c->stream));
    const int nsel = (int)c->hours.size();
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    SGPU_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    try {
        zero_accumulators(c);
        run_hour_passes(c, fmt, K, nsel, nullptr, nullptr);
    } catch (...) {
        hipGraph_t g = nullptr;
        (void)hipStreamEndCapture(c->stream, &g);
        if (g) (void)hipGraphDestroy(g);
        throw;
    }
    SGPU_HIP(hipStreamEndCapture(c->stream, &graph));
    hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(graph);
        SGPU_HIP(e);
    }
    EventPair all;
    SGPU_HIP(hipEventRecord(all.a, c->stream));
    for (int r = 0; r < replays; ++r) SGPU_HIP(hipGraphLaunch(exec, c->stream));
    SGPU_HIP(hipEventRecord(all.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    (void)hipGraphExecDestroy(exec);
    (void)hipGraphDestroy(graph);
    c->info.t_hours_ms = all.ms() / replays;
    c->info.hours_computed = nsel;
    c->info.hour_passes = (nsel + K - 1) / K;
    c->info.hours_per_pass = K;
    c->info.timed_launches = 0;
    c->irr_ok = true;
}

}  // namespace

namespace sgpu {

HourArgs stage2_args(sgpu_ctx* c, unsigned fmt) { return hour_args(c, fmt); }

SubsetLists subset_viewshed(sgpu_ctx* c, const int32_t* subset, int64_t n, bool barle) {
    SubsetLists s;
    if (n <= 0) return s;
    if (barle) {   // ba-RLE only, encoded chunk by chunk
        ListsOut L = build_lists(c, subset, n, nullptr, nullptr, false, false, true);
        s.nnz = L.nnz;
        s.br_off = L.br_off;
        s.br_bytes = L.br_bytes;
        s.br_size = L.br_size;
        return s;
    }
    ListsOut L = build_lists(c, subset, n, nullptr, nullptr);
    s.csr_off = L.off;
    s.csr_ids = L.ids;
    s.nnz = L.nnz;
    return s;
}

void subset_hour_loop(sgpu_ctx* c, const int32_t* subset, int64_t n, const SubsetLists& L, unsigned fmt, int K,
                      float* out) {
    if (n <= 0) return;
    need(supported_hours_per_pass(K), SGPU_ERR_ARG, "bad hours per pass");
    ensure_accumulators(c, K);
    ensure_hours_d(c, c->hours);
    float* comp = dalloc<float>(n);
    try {
        zero_floats(out, n, c->stream);
        zero_floats(comp, n, c->stream);
        HourArgs a = hour_args(c, fmt);
        a.subset = subset;
        a.order = nullptr;
        a.n_sub = n;
        a.sum = out;
        a.comp = comp;
        if (fmt == SGPU_FMT_CSR) {
            a.nb_off = L.csr_off;
            a.nb_ids = L.csr_ids;
        } else {
            a.nb_off = L.br_off;
            a.nb_bytes = L.br_bytes;
            a.nb_len = L.br_size;
        }
        const int nsel = (int)c->hours.size();
        for (int h0 = 0; h0 < nsel; h0 += K) {
            a.hours = c->hours_d + h0;
            a.nh = std::min(K, nsel - h0);
            launch_solar_top(a, K, c->stream);
            launch_reflect_accum(a, K, fmt, c->stream);
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        dfree(comp);
        throw;
    }
    dfree(comp);
}

void subset_reflected_linear(sgpu_ctx* c, const int32_t* subset, int64_t n, const SubsetLists& L, unsigned fmt,
                             const float* dd, float* out) {
    if (n <= 0) return;
    ensure_hours_d(c, c->hours);
    HourArgs a = hour_args(c, fmt);
    a.subset = subset;
    a.order = nullptr;
    a.n_sub = n;
    if (fmt == SGPU_FMT_CSR) {
        a.nb_off = L.csr_off;
        a.nb_ids = L.csr_ids;
    } else {
        a.nb_off = L.br_off;
        a.nb_bytes = L.br_bytes;
        a.nb_len = L.br_size;
    }
    a.hourly_top = const_cast<float*>(dd);   // K = 1: the top voxels' annual totals
    a.hourly_top_cap = c->n_top;
    a.row_swz = 0;
    a.hours = c->hours_d;
    a.nh = 1;
    float* own = dalloc<float>(n);
    try {
        a.dbg_hourly = own;
        a.dbg_reflected = out;
        launch_reflect_accum(a, 1, fmt, c->stream);
        SGPU_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        dfree(own);
        throw;
    }
    dfree(own);
}

/* Annual reflected irradiance via single-bounce linearity: one reflection walk with the
   annual top-cell totals in place of one hour's values. */
void annual_reflected_linear(sgpu_ctx* c, unsigned fmt, const float* dd, float* out_dev) {
    check_stage2_ready(c, fmt);
    need(!lists_on_host(c, fmt), SGPU_ERR_STATE, "needs device-resident lists");
    ensure_hours_d(c, c->hours);
    HourArgs a = hour_args(c, fmt);
    a.hourly_top = const_cast<float*>(dd);   // cell-major with K = 1: the top voxels' annual totals
    a.hourly_top_cap = c->n_top;
    a.row_swz = 0;
    a.hours = c->hours_d;
    a.nh = 1;
    float* own = dalloc<float>(c->n_total);
    try {
        a.dbg_hourly = own;
        a.dbg_reflected = out_dev;
        launch_reflect_accum(a, 1, fmt, c->stream);
        SGPU_HIP(hipStreamSynchronize(c->stream));
    } catch (...) {
        dfree(own);
        throw;
    }
    dfree(own);
}

}  // namespace sgpu

extern "C" {

const char* sgpu_version(void) { return "solargpu-mi355x 0.1"; }
const char* sgpu_last_error(void) { return last_error().c_str(); }

int sgpu_sun_positions(float lon, float lat, int tz, float* out) {
    return guarded([&] {
        need(out != nullptr, SGPU_ERR_ARG, "null output");
        sun_positions(lon, lat, tz, out);
    });
}

int sgpu_sky_bins(const float* az_alt, float* nudged, int32_t* f2d, int32_t* sdi, int32_t* active,
                  int32_t* svf_bytes) {
    return guarded([&] {
        need(az_alt && nudged && f2d && sdi && active && svf_bytes, SGPU_ERR_ARG, "null argument");
        SkyTables t = sky_bins(az_alt);
        std::memcpy(nudged, t.nudged.data(), t.nudged.size() * sizeof(float));
        std::memcpy(f2d, t.full_to_dense.data(), t.full_to_dense.size() * sizeof(int32_t));
        std::memcpy(sdi, t.sun_dense.data(), t.sun_dense.size() * sizeof(int32_t));
        *active = t.active;
        *svf_bytes = t.svf_bytes;
    });
}

int sgpu_directions(float* cdir, int32_t* is_upper) {
    return guarded([&] {
        need(cdir && is_upper, SGPU_ERR_ARG, "null argument");
        SkyTables t;
        t.full_to_dense.assign(kAzBins * kElBins, -1);
        std::vector<DirConst> d(kDirs);
        directions(t, d.data());
        for (int i = 0; i < kDirs; ++i) {
            cdir[3 * i] = d[i].cx;
            cdir[3 * i + 1] = d[i].cy;
            cdir[3 * i + 2] = d[i].cz;
            is_upper[i] = d[i].info & 1;
        }
    });
}

int sgpu_detect_walls(const float* heights, int w, int h, float voxel, int32_t* parent, float* z, float* aspect,
                      int64_t cap, int64_t* count) {
    return guarded([&] {
        need(heights && count && w > 0 && h > 0, SGPU_ERR_ARG, "bad arguments");
        Walls wl = detect_walls(heights, w, h, voxel);
        *count = (int64_t)wl.parent.size();
        const int64_t n = std::min<int64_t>(cap, *count);
        if (n > 0) {
            need(parent && z && aspect, SGPU_ERR_ARG, "null output");
            std::memcpy(parent, wl.parent.data(), n * sizeof(int32_t));
            std::memcpy(z, wl.z.data(), n * sizeof(float));
            std::memcpy(aspect, wl.aspect.data(), n * sizeof(float));
        }
    });
}

int sgpu_top_slope_aspect(const float* heights, int w, int h, double sx, double sy, float* slope, float* aspect) {
    return guarded([&] {
        need(heights && slope && aspect && w > 0 && h > 0, SGPU_ERR_ARG, "bad arguments");
        top_slope_aspect(heights, w, h, sx, sy, slope, aspect);
    });
}

int sgpu_assemble_voxels(const float* heights, int w, int h, double sx, double sy, int64_t cap, int64_t* n_total,
                         int32_t* pp, float* slopes, float* aspects, float* wall_z) {
    return guarded([&] {
        need(heights && n_total && w > 0 && h > 0, SGPU_ERR_ARG, "bad arguments");
        /* main.cpp:112-127: tops in raster order, then walls; the voxel size is the x pixel scale. */
        Walls wl = detect_walls(heights, w, h, (float)sx);
        const int64_t ntop = (int64_t)w * h;
        const int64_t ntot = ntop + (int64_t)wl.parent.size();
        *n_total = ntot;
        if (cap == 0) return;
        need(cap >= ntot, SGPU_ERR_ARG, "capacity smaller than the voxel count");
        need(pp && slopes && aspects, SGPU_ERR_ARG, "null output");
        top_slope_aspect(heights, w, h, sx, sy, slopes, aspects);
        for (int64_t i = 0; i < ntop; ++i) pp[i] = (int32_t)i;
        for (size_t k = 0; k < wl.parent.size(); ++k) {
            pp[ntop + k] = wl.parent[k];
            slopes[ntop + k] = (float)3.14159265358979323846 * 0.5f;
            aspects[ntop + k] = wl.aspect[k];
            if (wall_z) wall_z[k] = wl.z[k];
        }
    });
}

int sgpu_synth_dsm(uint64_t seed, int w, int h, float cell_m, float density, int trees, float* out) {
    return guarded([&] {
        need(out && w > 0 && h > 0 && cell_m > 0.0f, SGPU_ERR_ARG, "bad arguments");
        synth_dsm(seed, w, h, cell_m, density, trees, out);
    });
}

int sgpu_synth_weather(uint64_t seed, const float* az_alt, float* dni, float* dhi) {
    return guarded([&] {
        need(az_alt && dni && dhi, SGPU_ERR_ARG, "null argument");
        synth_weather(seed, az_alt, dni, dhi);
    });
}

int sgpu_create(int device, sgpu_ctx** out) {
    return guarded([&] {
        need(out != nullptr, SGPU_ERR_ARG, "null output");
        *out = nullptr;
        int n = 0;
        SGPU_HIP(hipGetDeviceCount(&n));
        need(device >= 0 && device < n, SGPU_ERR_ARG, "no such device");
        std::unique_ptr<sgpu_ctx> c(new sgpu_ctx());
        c->device = device;
        c->set_device();
        SGPU_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->dirs = dalloc<DirConst>(kDirs);
        c->noalloc_scratch = noalloc_scratch_setup(device);
        *out = c.release();
    });
}

int sgpu_destroy(sgpu_ctx* ctx) {
    return guarded([&] { delete ctx; });
}

int sgpu_get_info(sgpu_ctx* c, sgpu_info* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        c->info.grid_width = c->W;
        c->info.grid_height = c->H;
        c->info.num_total = c->n_total;
        c->info.num_top = c->n_top;
        c->info.active_bins = c->sky.active;
        c->info.svf_bytes_per_voxel = c->sky.svf_bytes;
        c->info.sunlit_hours = (int32_t)c->hours.size();
        *out = c->info;
    });
}

namespace {

/* The reflection-pass processing order and the DDA lane order of the context's scene, built on
   the host from its heightmap, parents and wall heights, then uploaded. */
void upload_orders(sgpu_ctx* c, const float* grid, const int32_t* pp, const float* wall_z) {
    const std::vector<int32_t> order = processing_order(c->W, c->H, c->n_total, pp, kOrderTile);
    c->order = dalloc<int32_t>(c->n_total);
    SGPU_HIP(hipMemcpyAsync(c->order, order.data(), c->n_total * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
    const std::vector<int32_t> tperm = trace_order(c->n_total, grid, pp, wall_z, c->n_top, kTraceOrderBlock);
    c->trace_perm = dalloc<int32_t>(c->n_total);
    SGPU_HIP(hipMemcpyAsync(c->trace_perm, tperm.data(), c->n_total * sizeof(int32_t), hipMemcpyHostToDevice,
                            c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
}

}  // namespace

int sgpu_set_scene(sgpu_ctx* c, const float* grid, int w, int h, int64_t n_total, int64_t n_top, const int32_t* pp,
                   const float* wall_z, float maxz, const float* slopes, const float* aspects) {
    return guarded([&] {
        need(c && grid && pp && slopes && aspects, SGPU_ERR_ARG, "null argument");
        need(w > 0 && h > 0 && grid_supported((unsigned)w, (unsigned)h), SGPU_ERR_ARG,
             "grid size out of range (width x (height + 2) must not exceed 2^30 cells, height 2^22 rows)");
        need(n_top == (int64_t)w * h, SGPU_ERR_ARG, "num_top_voxels must equal grid_width * grid_height");
        need(n_total >= n_top && n_total < (int64_t)1 << 31, SGPU_ERR_ARG, "num_total_voxels out of range");
        need(n_total == n_top || wall_z != nullptr, SGPU_ERR_ARG, "wall_z required when there are wall voxels");
        for (int64_t i = 0; i < n_total; ++i)
            if (pp[i] < 0 || pp[i] >= n_top) fail(SGPU_ERR_ARG, "pixel_parent[%lld] out of range", (long long)i);
        c->set_device();
        c->drop_products();
        c->drop_scene();
        c->W = w;
        c->H = h;
        c->n_total = n_total;
        c->n_top = n_top;
        c->maxz = maxz;
        c->grid = dalloc<float>(n_top);
        c->pix_parent = dalloc<int32_t>(n_total);
        c->slopes = dalloc<float>(n_total);
        c->aspects = dalloc<float>(n_total);
        const int64_t nw = n_total - n_top;
        c->wall_z = dalloc<float>(std::max<int64_t>(nw, 1));
        SGPU_HIP(hipMemcpyAsync(c->grid, grid, n_top * sizeof(float), hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipMemcpyAsync(c->pix_parent, pp, n_total * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipMemcpyAsync(c->slopes, slopes, n_total * sizeof(float), hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipMemcpyAsync(c->aspects, aspects, n_total * sizeof(float), hipMemcpyHostToDevice, c->stream));
        if (nw > 0)
            SGPU_HIP(hipMemcpyAsync(c->wall_z, wall_z, nw * sizeof(float), hipMemcpyHostToDevice, c->stream));
        upload_orders(c, grid, pp, wall_z);
        c->scene = true;
    });
}

int sgpu_set_scene_heights(sgpu_ctx* c, const float* heights, int w, int h, double scale_x, double scale_y) {
    return guarded([&] {
        need(c && heights, SGPU_ERR_ARG, "null argument");
        need(w > 0 && h > 0 && grid_supported((unsigned)w, (unsigned)h), SGPU_ERR_ARG,
             "grid size out of range (width x (height + 2) must not exceed 2^30 cells, height 2^22 rows)");
        need(scale_x > 0.0 && scale_y > 0.0, SGPU_ERR_ARG, "pixel scales must be positive");
        c->set_device();
        c->drop_products();
        c->drop_scene();
        const int64_t n_top = (int64_t)w * h;
        float maxz = heights[0];   // the reference's max_height: the heightmap's maximum
        for (int64_t i = 0; i < n_top; ++i) {
            need(std::isfinite(heights[i]), SGPU_ERR_ARG, "heights must be finite");
            maxz = std::max(maxz, heights[i]);
        }
        c->grid = dalloc<float>(n_top);
        SGPU_HIP(hipMemcpyAsync(c->grid, heights, n_top * sizeof(float), hipMemcpyHostToDevice, c->stream));
        DeviceVoxels dv;
        int64_t n_total = 0;
        try {
            n_total = voxelize_device(c->stream, c->grid, w, h, scale_x, scale_y, &dv);
        } catch (...) {
            dfree(dv.pix_parent);
            dfree(dv.slopes);
            dfree(dv.aspects);
            dfree(dv.wall_z);
            dfree(dv.wall_off);
            throw;
        }
        c->pix_parent = dv.pix_parent;
        c->slopes = dv.slopes;
        c->aspects = dv.aspects;
        c->wall_z = dv.wall_z;
        c->W = w;
        c->H = h;
        c->n_total = n_total;
        c->n_top = n_top;
        c->maxz = maxz;
        /* the processing orders, on the device too */
        try {
            c->order = dalloc<int32_t>(n_total);
            c->trace_perm = dalloc<int32_t>(n_total);
            orders_device(c->stream, c->grid, w, h, dv, n_total, kOrderTile, kTraceOrderBlock, c->order, c->trace_perm);
        } catch (...) {
            dfree(dv.wall_off);
            throw;
        }
        dfree(dv.wall_off);
        c->scene = true;
    });
}

int sgpu_get_orders(sgpu_ctx* c, int32_t* order, int32_t* lane_order) {
    return guarded([&] {
        need(c, SGPU_ERR_ARG, "null context");
        need(c->scene, SGPU_ERR_STATE, "scene not set");
        c->set_device();
        const size_t n = (size_t)c->n_total;
        if (order) SGPU_HIP(hipMemcpyAsync(order, c->order, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        if (lane_order)
            SGPU_HIP(hipMemcpyAsync(lane_order, c->trace_perm, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_get_voxels(sgpu_ctx* c, int32_t* pixel_parent, float* slopes, float* aspects, float* wall_z) {
    return guarded([&] {
        need(c, SGPU_ERR_ARG, "null context");
        need(c->scene, SGPU_ERR_STATE, "scene not set");
        c->set_device();
        const size_t n = (size_t)c->n_total, nw = (size_t)(c->n_total - c->n_top);
        if (pixel_parent)
            SGPU_HIP(hipMemcpyAsync(pixel_parent, c->pix_parent, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        if (slopes) SGPU_HIP(hipMemcpyAsync(slopes, c->slopes, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (aspects) SGPU_HIP(hipMemcpyAsync(aspects, c->aspects, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        if (wall_z && nw)
            SGPU_HIP(hipMemcpyAsync(wall_z, c->wall_z, nw * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_set_sky(sgpu_ctx* c, const float* az_alt, const float* dni, const float* dhi) {
    return guarded([&] {
        need(c && az_alt && dni && dhi, SGPU_ERR_ARG, "null argument");
        c->set_device();
        c->sky = sky_bins(az_alt);
        c->hours = hour_constants(c->sky, dni, dhi);
        std::vector<DirConst> d(kDirs);
        directions(c->sky, d.data());
        SGPU_HIP(hipMemcpyAsync(c->dirs, d.data(), kDirs * sizeof(DirConst), hipMemcpyHostToDevice, c->stream));
        /* every sunlit hour, for the allocation-free hour loop (which may not upload) */
        SGPU_HIP(hipStreamSynchronize(c->stream));
        dfree(c->hours_all_d);
        c->hours_all_n = 0;
        c->hours_all_d = dalloc<HourConst>(std::max<size_t>(c->hours.size(), 1));
        if (!c->hours.empty())
            SGPU_HIP(hipMemcpyAsync(c->hours_all_d, c->hours.data(), c->hours.size() * sizeof(HourConst),
                                    hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->hours_all_n = c->hours.size();
        c->drop_products();
        c->sky_set = true;
    });
}

int sgpu_set_chunk_voxels(sgpu_ctx* c, int64_t chunk) {
    return guarded([&] {
        need(c && chunk >= 0, SGPU_ERR_ARG, "bad argument");
        c->chunk_override = chunk;
    });
}

int sgpu_viewshed(sgpu_ctx* c, unsigned formats) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        do_viewshed(c, formats);
    });
}

int sgpu_get_svf(sgpu_ctx* c, uint8_t* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        need(c->svf_ok, SGPU_ERR_STATE, "no SVF bitmap");
        c->set_device();
        const size_t n = (size_t)c->sky.svf_bytes * c->n_total;
        if (n) SGPU_HIP(hipMemcpyAsync(out, c->svf, n, hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

/* Host-resident pieces -> one offsets array (rebased) and one concatenated data array. */
static void gather_host(const std::vector<HostPiece>& pieces, int64_t n_total, int64_t* off, uint8_t* data,
                        size_t elem) {
    int64_t pos = 0;
    for (const HostPiece& p : pieces) {
        for (int64_t i = 0; i < p.count; ++i) off[p.v0 + i] = pos / (int64_t)elem + p.off[i];
        if (p.bytes) std::memcpy(data + pos, p.data, (size_t)p.bytes);
        pos += p.bytes;
    }
    off[n_total] = pos / (int64_t)elem;
}

int sgpu_set_list_residency(sgpu_ctx* c, int residency) {
    return guarded([&] {
        need(c && (residency == SGPU_LISTS_DEVICE || residency == SGPU_LISTS_HOST), SGPU_ERR_ARG, "bad residency");
        c->residency = residency;
    });
}

int sgpu_get_csr(sgpu_ctx* c, int64_t* off, int32_t* ids) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->nnz >= 0, SGPU_ERR_STATE, "no CSR lists");
        c->set_device();
        if (!c->csr_off) {
            need(ids != nullptr || c->nnz == 0, SGPU_ERR_ARG, "null ids");
            gather_host(c->host.csr, c->n_total, off, reinterpret_cast<uint8_t*>(ids), sizeof(int32_t));
            return;
        }
        SGPU_HIP(hipMemcpyAsync(off, c->csr_off, (c->n_total + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                                c->stream));
        if (c->nnz > 0) {
            need(ids != nullptr, SGPU_ERR_ARG, "null ids");
            SGPU_HIP(hipMemcpyAsync(ids, c->csr_ids, c->nnz * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_get_barle(sgpu_ctx* c, int64_t* off, uint8_t* bytes) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->br_size >= 0, SGPU_ERR_STATE, "no ba-RLE streams");
        c->set_device();
        if (!c->br_off) {
            need(bytes != nullptr || c->br_size == 0, SGPU_ERR_ARG, "null bytes");
            gather_host(c->host.barle, c->n_total, off, bytes, 1);
            return;
        }
        SGPU_HIP(hipMemcpyAsync(off, c->br_off, (c->n_total + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                                c->stream));
        if (c->br_size > 0) {
            need(bytes != nullptr, SGPU_ERR_ARG, "null bytes");
            SGPU_HIP(hipMemcpyAsync(bytes, c->br_bytes, c->br_size, hipMemcpyDeviceToHost, c->stream));
        }
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_get_csr_range(sgpu_ctx* c, int64_t v0, int64_t count, int64_t* off, int32_t* ids, int64_t cap) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->nnz >= 0 && c->csr_off, SGPU_ERR_STATE, "no device-resident CSR lists");
        need(v0 >= 0 && count >= 0 && v0 + count <= c->n_total, SGPU_ERR_ARG, "voxel range out of bounds");
        c->set_device();
        SGPU_HIP(hipMemcpyAsync(off, c->csr_off + v0, (count + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                                c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        const int64_t base = off[0], n = off[count] - base;
        need(n <= cap, SGPU_ERR_ARG, "ids capacity too small");
        for (int64_t i = 0; i <= count; ++i) off[i] -= base;
        if (n > 0) {
            need(ids != nullptr, SGPU_ERR_ARG, "null ids");
            SGPU_HIP(hipMemcpyAsync(ids, c->csr_ids + base, n * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
        }
    });
}

int sgpu_get_barle_range(sgpu_ctx* c, int64_t v0, int64_t count, int64_t* off, uint8_t* bytes, int64_t cap) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->br_size >= 0 && c->br_off, SGPU_ERR_STATE, "no device-resident ba-RLE streams");
        need(v0 >= 0 && count >= 0 && v0 + count <= c->n_total, SGPU_ERR_ARG, "voxel range out of bounds");
        c->set_device();
        SGPU_HIP(hipMemcpyAsync(off, c->br_off + v0, (count + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                                c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        const int64_t base = off[0], n = off[count] - base;
        need(n <= cap, SGPU_ERR_ARG, "bytes capacity too small");
        for (int64_t i = 0; i <= count; ++i) off[i] -= base;
        if (n > 0) {
            need(bytes != nullptr, SGPU_ERR_ARG, "null bytes");
            SGPU_HIP(hipMemcpyAsync(bytes, c->br_bytes + base, n, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
        }
    });
}

int sgpu_trace_raw(sgpu_ctx* c, int64_t v0, int64_t count, int32_t* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        need(c->scene && c->sky_set, SGPU_ERR_STATE, "scene/sky not set");
        need(v0 >= 0 && count >= 0 && v0 + count <= c->n_total, SGPU_ERR_ARG, "voxel range out of bounds");
        c->set_device();
        int32_t* raw = dalloc<int32_t>((size_t)count * kDirs);
        try {
            TraceArgs ta{c->grid, c->W, c->H, c->n_top, c->pix_parent, c->wall_z, c->maxz, c->dirs, v0, count,
                         raw, nullptr, nullptr, 0, nullptr};
            launch_trace(ta, c->stream);
            SGPU_HIP(hipMemcpyAsync(out, raw, (size_t)count * kDirs * sizeof(int32_t), hipMemcpyDeviceToHost,
                                    c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            dfree(raw);
            throw;
        }
        dfree(raw);
    });
}

int sgpu_set_svf(sgpu_ctx* c, const uint8_t* svf) {
    return guarded([&] {
        need(c && svf, SGPU_ERR_ARG, "null argument");
        need(c->scene && c->sky_set, SGPU_ERR_STATE, "scene/sky not set");
        c->set_device();
        dfree(c->svf);
        const size_t n = (size_t)c->sky.svf_bytes * c->n_total;
        c->svf = dalloc<uint8_t>(std::max<size_t>(n, 1));
        if (n) SGPU_HIP(hipMemcpyAsync(c->svf, svf, n, hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->svf_ok = true;
    });
}

int sgpu_set_csr(sgpu_ctx* c, const int64_t* off, const int32_t* ids) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->scene, SGPU_ERR_STATE, "scene not set");
        const int64_t N = c->n_total;
        need(off[0] == 0, SGPU_ERR_ARG, "offsets[0] must be 0");
        for (int64_t i = 0; i < N; ++i) {
            const int64_t len = off[i + 1] - off[i];
            if (len < 0 || len > kDirs) fail(SGPU_ERR_ARG, "list %lld has bad length %lld", (long long)i, (long long)len);
        }
        const int64_t nnz = off[N];
        for (int64_t k = 0; k < nnz; ++k)
            if (ids[k] <= 0) fail(SGPU_ERR_ARG, "ids[%lld] must be a 1-based cell id", (long long)k);
        c->set_device();
        dfree(c->csr_off);
        dfree(c->csr_ids);
        c->csr_off = dalloc<int64_t>(N + 1);
        c->csr_ids = dalloc<int32_t>((size_t)std::max<int64_t>(nnz, 1));
        SGPU_HIP(hipMemcpyAsync(c->csr_off, off, (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        if (nnz) SGPU_HIP(hipMemcpyAsync(c->csr_ids, ids, nnz * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->nnz = nnz;
        c->info.nnz = nnz;
    });
}

int sgpu_set_barle(sgpu_ctx* c, const int64_t* off, const uint8_t* bytes) {
    return guarded([&] {
        need(c && off, SGPU_ERR_ARG, "null argument");
        need(c->scene, SGPU_ERR_STATE, "scene not set");
        const int64_t N = c->n_total;
        need(off[0] == 0, SGPU_ERR_ARG, "offsets[0] must be 0");
        /* Walk every stream once on the host so the decoder never reads past a range. */
        for (int64_t v = 0; v < N; ++v) {
            const int64_t a = off[v], b = off[v + 1];
            if (b < a || b - a > (int64_t)1 << 20) fail(SGPU_ERR_FORMAT, "stream %lld has a bad range", (long long)v);
            int64_t i = a;
            while (i < b) {
                if (i + 3 > b) fail(SGPU_ERR_FORMAT, "stream %lld: truncated block header", (long long)v);
                const unsigned bits = ((unsigned)bytes[i] << 16) | ((unsigned)bytes[i + 1] << 8) | bytes[i + 2];
                i += 3;
                unsigned w[8];
                int nc = 0;
                for (int j = 0; j < 8; ++j) {
                    const unsigned cd = (bits >> (21 - 3 * j)) & 7u;
                    if (cd == 7u) break;
                    if (cd > 4u) fail(SGPU_ERR_FORMAT, "stream %lld: width code %u", (long long)v, cd);
                    w[nc++] = cd;
                }
                if (i + nc > b) fail(SGPU_ERR_FORMAT, "stream %lld: truncated lengths", (long long)v);
                int64_t payload = 0;
                for (int j = 0; j < nc; ++j) {
                    const int64_t len = bytes[i + j];
                    payload += (w[j] == 0) ? (len + 1) / 2 : len * (int64_t)w[j];
                }
                i += nc + payload;
                if (i > b) fail(SGPU_ERR_FORMAT, "stream %lld: truncated payload", (long long)v);
            }
        }
        const int64_t total = off[N];
        c->set_device();
        dfree(c->br_off);
        dfree(c->br_bytes);
        c->br_off = dalloc<int64_t>(N + 1);
        c->br_bytes = dalloc<uint8_t>((size_t)std::max<int64_t>(total, 1));
        SGPU_HIP(hipMemcpyAsync(c->br_off, off, (N + 1) * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
        if (total) SGPU_HIP(hipMemcpyAsync(c->br_bytes, bytes, total, hipMemcpyHostToDevice, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
        c->br_size = total;
        c->info.barle_bytes = total;
    });
}

int sgpu_barle_encode(sgpu_ctx* c) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        do_barle_encode(c);
    });
}

int sgpu_hour_loop(sgpu_ctx* c, unsigned fmt, int K, int stride) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        do_hour_loop(c, fmt, K, stride);
    });
}

int sgpu_hour_loop_graph(sgpu_ctx* c, unsigned fmt, int K, int replays) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        do_hour_loop_graph(c, fmt, K, replays);
    });
}

int sgpu_get_irradiance(sgpu_ctx* c, float* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        need(c->irr_ok, SGPU_ERR_STATE, "no irradiance computed");
        c->set_device();
        SGPU_HIP(hipMemcpyAsync(out, c->sum, c->n_total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_hour_single(sgpu_ctx* c, unsigned fmt, int hour_index, float* hourly_out, float* refl_out) {
    return guarded([&] {
        need(c && hourly_out && refl_out, SGPU_ERR_ARG, "null argument");
        check_stage2_ready(c, fmt);
        need(!lists_on_host(c, fmt), SGPU_ERR_STATE, "needs device-resident lists");
        c->set_device();
        const HourConst* hc = nullptr;
        for (const HourConst& h : c->hours)
            if (h.index == hour_index) hc = &h;
        need(hc != nullptr, SGPU_ERR_ARG, "hour is not a sunlit hour");
        ensure_accumulators(c, 1);
        ensure_hours_d(c, std::vector<HourConst>{*hc});
        float* dh = dalloc<float>(c->n_total);
        float* dr = dalloc<float>(c->n_total);
        try {
            HourArgs a = hour_args(c, fmt);
            a.hours = c->hours_d;
            a.nh = 1;
            a.dbg_hourly = dh;
            a.dbg_reflected = dr;
            launch_solar_top(a, 1, c->stream);
            launch_reflect_accum(a, 1, fmt, c->stream);
            SGPU_HIP(hipMemcpyAsync(hourly_out, dh, c->n_total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipMemcpyAsync(refl_out, dr, c->n_total * sizeof(float), hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            dfree(dh);
            dfree(dr);
            throw;
        }
        dfree(dh);
        dfree(dr);
    });
}

int sgpu_set_kernel_timing(sgpu_ctx* c, int enable) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->kernel_timing = enable != 0;
    });
}

int sgpu_set_hourly_limit(sgpu_ctx* c, int64_t floats) {
    return guarded([&] {
        need(c != nullptr && floats >= 0, SGPU_ERR_ARG, "bad argument");
        c->hourly_limit = floats;
    });
}

int sgpu_set_step_counting(sgpu_ctx* c, int enable) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->count_steps = enable != 0;
    });
}

int sgpu_set_trace_tables(sgpu_ctx* c, int enable, int64_t pending_cap) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        need(pending_cap >= 0, SGPU_ERR_ARG, "pending_cap must be >= 0");
        c->trace_tables = enable != 0;
        c->pend_cap_override = pending_cap;
    });
}

int sgpu_set_pair_limit(sgpu_ctx* c, int entries) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        need(entries >= 0, SGPU_ERR_ARG, "pair limit must be >= 0");
        c->group_limit = entries;
    });
}

int sgpu_set_trace_table_shape(sgpu_ctx* c, int dirs, int steps, int skip) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        const bool built = (dirs == 8 && (steps == 255 || steps == 126 || steps == 63)) ||
                           (dirs == 4 && (steps == 510 || steps == 255 || steps == 126)) || (dirs == 16 && steps == 126);
        need(built, SGPU_ERR_ARG, "step-table shape not built (8x255, 4x255, 4x510, 8x126, 4x126, 16x126, 8x63)");
        c->table_dirs = dirs;
        c->table_steps = steps;
        c->table_skip = skip != 0;
    });
}

int sgpu_sync(sgpu_ctx* c) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_copy_bandwidth(sgpu_ctx* c, int64_t bytes, int reps, double* gbps) {
    return guarded([&] {
        need(c && gbps && bytes >= 16 && reps >= 1, SGPU_ERR_ARG, "bad arguments");
        c->set_device();
        const int64_t n = bytes / 16;
        float4* src = dalloc<float4>(n);
        float4* dst = nullptr;
        try {
            dst = dalloc<float4>(n);
            SGPU_HIP(hipMemsetAsync(src, 0, n * 16, c->stream));
            const unsigned g = (unsigned)std::min<int64_t>((n + 255) / 256, 256 * 1024);
            hipLaunchKernelGGL(k_copy_f4, dim3(g), dim3(256), 0, c->stream, src, dst, n);  // warm-up
            EventPair ev;
            SGPU_HIP(hipEventRecord(ev.a, c->stream));
            for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_copy_f4, dim3(g), dim3(256), 0, c->stream, src, dst, n);
            SGPU_HIP(hipEventRecord(ev.b, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            *gbps = 2.0 * (double)n * 16.0 * reps / (ev.ms() * 1e-3) / 1e9;
        } catch (...) {
            dfree(src);
            dfree(dst);
            throw;
        }
        dfree(src);
        dfree(dst);
    });
}

int sgpu_l2_bandwidth(sgpu_ctx* c, int64_t footprint, int reps, double* gbps) {
    return guarded([&] {
        need(c && gbps && reps >= 1, SGPU_ERR_ARG, "bad arguments");
        need(footprint >= 64 * 1024 && footprint <= ((int64_t)64 << 20) && footprint % 4096 == 0, SGPU_ERR_ARG,
             "footprint must be a multiple of 4 KiB in 64 KiB .. 64 MiB");
        c->set_device();
        int cus = 0;
        SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        const int blocks = std::max(cus, 1) * 8;   // 32 waves per CU
        const int64_t n = footprint / 16;
        float4* src = dalloc<float4>(n);
        float* out = nullptr;
        try {
            out = dalloc<float>((size_t)blocks * 256);
            SGPU_HIP(hipMemsetAsync(src, 0, n * 16, c->stream));
            hipLaunchKernelGGL(k_l2_read, dim3(blocks), dim3(256), 0, c->stream, src, n, out);   // warm-up
            SGPU_HIP(hipGetLastError());
            EventPair ev;
            SGPU_HIP(hipEventRecord(ev.a, c->stream));
            for (int r = 0; r < reps; ++r) hipLaunchKernelGGL(k_l2_read, dim3(blocks), dim3(256), 0, c->stream, src, n, out);
            SGPU_HIP(hipEventRecord(ev.b, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            const int64_t rows = n / 64;
            const double bytes = (double)blocks * 4.0 * (double)(rows / 4 * 4) * 1024.0 * reps;
            *gbps = bytes / (ev.ms() * 1e-3) / 1e9;
        } catch (...) {
            dfree(src);
            dfree(out);
            throw;
        }
        dfree(src);
        dfree(out);
    });
}

int sgpu_l1_gather_rate(sgpu_ctx* c, int64_t footprint, int reps, double* accesses_per_s) {
    return guarded([&] {
        need(c && accesses_per_s && reps >= 1, SGPU_ERR_ARG, "bad arguments");
        need(footprint >= 4096 && footprint <= ((int64_t)1 << 30) && (footprint & (footprint - 1)) == 0, SGPU_ERR_ARG,
             "footprint must be a power of two in 4 KiB .. 1 GiB");
        c->set_device();
        int cus = 0;
        SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        const int blocks = std::max(cus, 1) * 8;   // 32 waves per CU
        const int64_t n = footprint / 4;
        float* buf = dalloc<float>((size_t)n);
        float* out = nullptr;
        try {
            out = dalloc<float>((size_t)blocks * 256);
            SGPU_HIP(hipMemsetAsync(buf, 0, (size_t)footprint, c->stream));
            const unsigned mask = (unsigned)(n - 1);
            hipLaunchKernelGGL(k_gather_probe, dim3(blocks), dim3(256), 0, c->stream, buf, mask, out);   // warm-up
            SGPU_HIP(hipGetLastError());
            EventPair ev;
            SGPU_HIP(hipEventRecord(ev.a, c->stream));
            for (int r = 0; r < reps; ++r)
                hipLaunchKernelGGL(k_gather_probe, dim3(blocks), dim3(256), 0, c->stream, buf, mask, out);
            SGPU_HIP(hipEventRecord(ev.b, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            const double accesses = (double)blocks * 4.0 * kGatherTurns * 8.0 * 64.0 * reps;
            *accesses_per_s = accesses / (ev.ms() * 1e-3);
        } catch (...) {
            dfree(buf);
            dfree(out);
            throw;
        }
        dfree(buf);
        dfree(out);
    });
}

int sgpu_valu_rate(sgpu_ctx* c, int mix, double* wave_insts_per_s) {
    return guarded([&] {
        need(c && wave_insts_per_s, SGPU_ERR_ARG, "null argument");
        need(mix >= 0 && mix <= 2, SGPU_ERR_ARG, "mix must be 0 (FMA), 1 (add/compare/select) or 2 (compare to SGPR)");
        c->set_device();
        int cus = 0;
        SGPU_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
        const int blocks = std::max(cus, 1) * 8;   // 32 waves per CU: every SIMD holds 8
        const int turns = 4000, reps = 5;
        float* out = dalloc<float>((size_t)blocks * 256);
        try {
            auto launch = [&] {
                if (mix == 0)
                    hipLaunchKernelGGL(k_valu_probe, dim3(blocks), dim3(256), 0, c->stream, out, turns, 0.999f, 1e-3f);
                else if (mix == 1)
                    hipLaunchKernelGGL(k_valu_probe_mix, dim3(blocks), dim3(256), 0, c->stream,
                                       reinterpret_cast<unsigned*>(out), turns, 7u, 0x9E3779B9u);
                else
                    hipLaunchKernelGGL(k_valu_probe_cmp, dim3(blocks), dim3(256), 0, c->stream,
                                       reinterpret_cast<unsigned*>(out), turns, 0x9E3779B9u, 0x7FFFFFFFu);
            };
            launch();
            SGPU_HIP(hipGetLastError());
            EventPair ev;
            SGPU_HIP(hipEventRecord(ev.a, c->stream));
            for (int r = 0; r < reps; ++r) launch();
            SGPU_HIP(hipEventRecord(ev.b, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            const double per_turn = mix == 0 ? kProbeFmaPerTurn : mix == 1 ? kProbeMixPerTurn : kProbeCmpPerTurn;
            const double insts = (double)blocks * 4.0 * turns * per_turn * reps;
            *wave_insts_per_s = insts / (ev.ms() * 1e-3);
        } catch (...) {
            dfree(out);
            throw;
        }
        dfree(out);
    });
}

static int whole_pipeline(unsigned fmt, const float* grid, int w, int h, int n_total, int n_top, const int32_t* pp,
                          const float* wall_z, float maxz, const float* az_alt, const float* dni, const float* dhi,
                          const float* slopes, const float* aspects, float* out) {
    sgpu_ctx* c = nullptr;
    int st = sgpu_create(0, &c);
    if (st) return st;
    st = sgpu_set_scene(c, grid, w, h, n_total, n_top, pp, wall_z, maxz, slopes, aspects);
    if (!st) st = sgpu_set_sky(c, az_alt, dni, dhi);
    if (!st) st = sgpu_viewshed(c, fmt);
    if (!st) st = guarded([&] { apply_cached_tuning(c, fmt); });   // shapes a past sgpu_autotune measured
    if (!st) st = sgpu_hour_loop(c, fmt, 0, 1);
    if (!st) st = sgpu_get_irradiance(c, out);
    std::string keep = last_error();
    sgpu_destroy(c);
    last_error() = keep;
    return st;
}

int sgpu_uncompressed_irradiance(const float* grid, int w, int h, int n_total, int n_top, const int32_t* pp,
                                 const float* wall_z, float maxz, const float* az_alt, const float* dni,
                                 const float* dhi, const float* slopes, const float* aspects, float* out) {
    return whole_pipeline(SGPU_FMT_CSR, grid, w, h, n_total, n_top, pp, wall_z, maxz, az_alt, dni, dhi, slopes,
                          aspects, out);
}

int sgpu_compressed_irradiance(const float* grid, int w, int h, int n_total, int n_top, const int32_t* pp,
                               const float* wall_z, float maxz, const float* az_alt, const float* dni,
                               const float* dhi, const float* slopes, const float* aspects, float* out) {
    return whole_pipeline(SGPU_FMT_BARLE, grid, w, h, n_total, n_top, pp, wall_z, maxz, az_alt, dni, dhi, slopes,
                          aspects, out);
}

}  // extern "C"
