/* Allocation-free annual hour loop (no reference counterpart).  For callers that cannot
 * allocate device memory while they run -- code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc -- sgpu_hour_loop_noalloc runs the same kernels
 * as sgpu_hour_loop (k_solar_top, k_reflect), with the per-pass hourly rows of the top
 * cells, the Kahan compensation and the reflection queue counters in a file-scope device
 * array instead of context buffers, and writes the annual sums into the caller's device
 * buffer.
 *
 * The array takes kScratchBytes of every GPU the library runs on until the process ends;
 * its address on a device is looked up once, when the first context of that device is
 * created, and kept per device.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two contexts on two streams, say) race.  What must start at zero
 * (the compensation, the queue counters) is cleared by fill kernels at the start of every
 * call or pass; the hourly rows are rewritten by every pass before they are read. */
#include <mutex>
#include <vector>

#include "sgpu_ctx.h"

namespace {
constexpr size_t kScratchBytes = (size_t)128 << 20;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace sgpu {
namespace {

std::mutex g_scratch_mu;
std::vector<void*> g_scratch;   // by device ordinal: that device's copy of `scratch`

__global__ __launch_bounds__(256) void k_clear(float* __restrict__ p, int64_t n) {
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256) p[i] = 0.0f;
}

void clear_floats(float* p, int64_t n, hipStream_t s) {
    if (n <= 0) return;
    const int64_t blocks = std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_clear, dim3((unsigned)blocks), dim3(256), 0, s, p, n);
    SGPU_HIP(hipGetLastError());
}

size_t round256(size_t b) { return (b + 255) & ~(size_t)255; }

/* hourly rows (n_top x K floats), then the compensation (n_total floats), then the queue
   counters of the reflection pass, each 256-byte aligned */
constexpr size_t kQueueBytes = (size_t)kQueueWords * sizeof(unsigned);
size_t scratch_bytes(int64_t n_top, int64_t n_total, int K) {
    return round256((size_t)n_top * K * sizeof(float)) + round256((size_t)n_total * sizeof(float)) + kQueueBytes;
}

}  // namespace

void* noalloc_scratch_setup(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if ((size_t)device < g_scratch.size() && g_scratch[device]) return g_scratch[device];
    void* p = nullptr;
    SGPU_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));   // the current device's copy
    need(p != nullptr, SGPU_ERR_HIP, "no device address for the allocation-free scratch array");
    if (g_scratch.size() <= (size_t)device) g_scratch.resize((size_t)device + 1, nullptr);
    g_scratch[device] = p;
    return p;
}

}  // namespace sgpu

using namespace sgpu;

int sgpu_hour_loop_noalloc(sgpu_ctx* c, unsigned fmt, int K, float* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        need(c->noalloc_scratch != nullptr, SGPU_ERR_STATE, "allocation-free scratch not set up");
        need(c->scene, SGPU_ERR_STATE, "scene not set");
        need(c->sky_set && c->hours_all_d, SGPU_ERR_STATE, "sky not set");
        need(c->svf_ok, SGPU_ERR_STATE, "SVF bitmap missing (run sgpu_viewshed or sgpu_set_svf)");
        if (fmt == SGPU_FMT_CSR) need(c->nnz >= 0 && c->csr_off, SGPU_ERR_STATE, "device-resident CSR lists missing");
        else if (fmt == SGPU_FMT_BARLE)
            need(c->br_size >= 0 && c->br_off, SGPU_ERR_STATE, "device-resident ba-RLE streams missing");
        else fail(SGPU_ERR_ARG, "format must be SGPU_FMT_CSR or SGPU_FMT_BARLE");
        if (K == 0) {   // the tuned hours per pass, halved until the scratch holds them
            K = c->tune.hours_per_pass;
            while (K > 1 && scratch_bytes(c->n_top, c->n_total, K) > kScratchBytes) K /= 2;
        }
        need(supported_hours_per_pass(K), SGPU_ERR_ARG, "hours_per_pass must be 1, 2, 4, 8, 16, 32, 64, 128 or 256");
        const size_t want = scratch_bytes(c->n_top, c->n_total, K);
        if (want > kScratchBytes)
            fail(SGPU_ERR_ARG,
                 "scene needs %zu bytes of scratch at %d hours per pass; the allocation-free hour loop has %zu "
                 "(use fewer hours per pass, or sgpu_hour_loop)",
                 want, K, kScratchBytes);
        c->set_device();
        uint8_t* base = static_cast<uint8_t*>(c->noalloc_scratch);
        HourArgs a = stage2_args(c, fmt);
        a.hourly_top = reinterpret_cast<float*>(base);
        a.hourly_top_cap = c->n_top * (int64_t)K;
        a.row_swz = 0;   // rows in cell order: the scratch holds exactly n_top x K floats
        a.comp = reinterpret_cast<float*>(base + round256((size_t)c->n_top * K * sizeof(float)));
        a.sum = out;
        a.queue = nullptr;
        a.grp_size = 0;   // receivers one by one: the grouped walk's scratch is a context buffer
        a.grp_scratch = nullptr;
        if (c->tune.receiver_queue) {   // the launcher zeroes the counters with a kernel
            a.queue = reinterpret_cast<unsigned*>(base + want - kQueueBytes);
            a.queue_band = (int64_t)c->tune.xcd_band * 64 * (c->tune.waves_per_group == 2 ? 2 : 4);
            a.queue_blocks_per_cu = c->tune.group_blocks;
        }
        clear_floats(a.sum, c->n_total, c->stream);
        clear_floats(a.comp, c->n_total, c->stream);
        const int nsel = (int)c->hours_all_n;
        for (int h0 = 0; h0 < nsel; h0 += K) {
            a.hours = c->hours_all_d + h0;
            a.nh = std::min(K, nsel - h0);
            launch_solar_top(a, K, c->stream);
            launch_reflect_accum(a, K, fmt, c->stream);
        }
    });
}
