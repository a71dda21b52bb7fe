/* Autotuner of the hot kernels' launch shapes (C ABI sgpu_autotune /
 * sgpu_set_tuning / sgpu_get_tuning).
 *
 *   DDA tile       directions per 256-voxel workgroup, G in {16, 32, 48}:
 *                  timed on a 2^17-voxel slice from the middle of the scene;
 *   stage-2 pass   hours per pass K in {64, 128, 256} x waves per reflection
 *                  workgroup in {2, 4} (LDS footprint and occupancy) x XCD
 *                  band in {0, 64, 256} (contiguous eighths or side-by-side
 *                  bands), and the queued receivers with bands {32, 256, 512}, then
 *                  for the best K: waves {2, 4} x receiver groups {8, 32, 64,
 *                  128} x band {0, 256}: three passes each on the live lists,
 *                  scratch accumulators.
 * Waves, bands, receiver groups and the receiver queue give bit-identical results; K only regroups
 * the hours of the Kahan-summed annual totals (same sums to FP32 rounding).  The winners are
 * appended to a text cache keyed by device, format and scene size class. */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <sys/stat.h>

#include "sgpu_ctx.h"
#include "sgpu_internal.h"

using namespace sgpu;

namespace {

std::string cache_path() {
    if (const char* p = std::getenv("SGPU_AUTOTUNE_CACHE")) return p;
    const char* home = std::getenv("HOME");
    std::string dir = std::string(home ? home : "/tmp") + "/.cache";
    ::mkdir(dir.c_str(), 0755);
    dir += "/solargpu";
    ::mkdir(dir.c_str(), 0755);
    return dir + "/autotune.txt";
}

std::string tune_key(sgpu_ctx* c, unsigned fmt) {
    hipDeviceProp_t p;
    SGPU_HIP(hipGetDeviceProperties(&p, c->device));
    const int nb = (int)std::lround(std::log2((double)std::max<int64_t>(c->n_total, 1)));
    const int64_t bytes = (fmt == SGPU_FMT_CSR) ? 4 * std::max<int64_t>(c->nnz, 0) : std::max<int64_t>(c->br_size, 0);
    const int bb = (int)std::lround(std::log2(1.0 + (double)bytes / (double)std::max<int64_t>(c->n_total, 1)));
    std::ostringstream k;
    k << p.gcnArchName << "/" << p.multiProcessorCount << "/fmt" << fmt << "/n2^" << nb << "/b2^" << bb;
    std::string s = k.str();
    for (char& ch : s)
        if (ch == ' ' || ch == ':') ch = '_';
    return s;
}

bool lookup(const std::string& key, sgpu_tuning& t) {
    std::ifstream f(cache_path());
    std::string line;
    bool found = false;
    while (std::getline(f, line)) {
        std::istringstream in(line);
        std::string k;
        sgpu_tuning r{};
        if (in >> k >> r.dirs_per_group >> r.hours_per_pass >> r.waves_per_group >> r.xcd_band >> r.group_blocks &&
            k == key) {
            if (!(in >> r.receiver_queue)) r.receiver_queue = 0;   // entries written before the queue existed
            t = r;
            found = true;   // the last entry for a key wins
        }
    }
    return found;
}

void store(const std::string& key, const sgpu_tuning& t) {
    std::ofstream f(cache_path(), std::ios::app);
    if (f) f << key << " " << t.dirs_per_group << " " << t.hours_per_pass << " " << t.waves_per_group << " "
             << t.xcd_band << " " << t.group_blocks << " " << t.receiver_queue << "\n";
}

double time_trace(sgpu_ctx* c, int G, int64_t v0, int64_t n, int32_t* raw) {
    TraceArgs ta{c->grid, c->W, c->H, c->n_top, c->pix_parent, c->wall_z, c->maxz, c->dirs, v0, n, raw,
                 nullptr, nullptr, G, c->trace_perm};
    launch_trace(ta, c->stream);   // warm-up
    EventPair e;
    SGPU_HIP(hipEventRecord(e.a, c->stream));
    launch_trace(ta, c->stream);
    SGPU_HIP(hipEventRecord(e.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    return e.ms();
}

double time_passes(sgpu_ctx* c, unsigned fmt, int K, int waves, int band, int group, int queue, int passes) {
    ensure_accumulators(c, K);   // the hourly buffer grows to this trial's n_top x K
    if (queue >= 2) ensure_group_scratch(c);
    HourArgs a = stage2_args(c, fmt);
    a.waves = waves;
    a.band = band;
    a.group = group;
    a.queue = queue ? c->reflect_queue : nullptr;
    a.queue_band = (int64_t)band * 64 * waves;
    a.queue_blocks_per_cu = queue ? group : 0;
    set_group_args(c, a, queue);
    const int nsel = (int)c->hours.size();
    auto run = [&](int np) {
        for (int p = 0, h0 = 0; p < np && h0 < nsel; ++p, h0 += K) {
            a.hours = c->hours_d + h0;
            a.nh = std::min(K, nsel - h0);
            launch_solar_top(a, K, c->stream);
            launch_reflect_accum(a, K, fmt, c->stream);
        }
    };
    run(1);   // warm-up
    EventPair e;
    SGPU_HIP(hipEventRecord(e.a, c->stream));
    run(passes);
    SGPU_HIP(hipEventRecord(e.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    return e.ms();
}

void autotune(sgpu_ctx* c, unsigned fmt, sgpu_tuning* out) {
    need(c->scene && c->sky_set, SGPU_ERR_STATE, "scene/sky not set");
    need(fmt == SGPU_FMT_CSR || fmt == SGPU_FMT_BARLE, SGPU_ERR_ARG, "bad format");
    const bool lists = (fmt == SGPU_FMT_CSR) ? (c->nnz >= 0 && c->csr_off) : (c->br_size >= 0 && c->br_off);
    const std::string key = tune_key(c, fmt);
    sgpu_tuning t = c->tune;
    if (lookup(key, t)) {
        t.from_cache = 1;
        t.rays_per_s = t.voxel_hours_per_s = 0.0;
        c->tune = t;
        if (out) *out = t;
        return;
    }
    t.from_cache = 0;
    /* DDA tile shape */
    const int64_t n = std::min<int64_t>(c->n_total, (int64_t)1 << 17);
    const int64_t v0 = (c->n_total - n) / 2 / kTraceOrderBlock * kTraceOrderBlock;   // a lane-order block
    int32_t* raw = dalloc<int32_t>((size_t)n * kDirs);
    double best = 1e30;
    try {
        for (int G : {16, 32, 48}) {
            const double ms = time_trace(c, G, v0, n, raw);
            if (ms < best) { best = ms; t.dirs_per_group = G; }
        }
    } catch (...) {
        dfree(raw);
        throw;
    }
    dfree(raw);
    t.rays_per_s = (double)n * kDirs / (best * 1e-3);
    /* stage-2 pass shape */
    if (lists && c->svf_ok && !c->hours.empty()) {
        ensure_hours_d(c, c->hours);
        double best_rate = 0.0;
        const int passes = 3;
        auto trial = [&](int K, int waves, int band, int group, int queue) {
            const double ms = time_passes(c, fmt, K, waves, band, group, queue, passes);
            const int hours = std::min<int>(passes * K, (int)c->hours.size());
            const double rate = (double)c->n_total * hours / (ms * 1e-3);
            if (rate > best_rate) {
                best_rate = rate;
                t.hours_per_pass = K;
                t.waves_per_group = waves;
                t.xcd_band = band;
                t.group_blocks = group;
                t.receiver_queue = queue;
            }
        };
        for (int K : {64, 128, 256})
            for (int waves : {2, 4}) {
                for (int band : {0, 64, 256}) trial(K, waves, band, 0, 0);
                for (int band : {32, 256, 512}) trial(K, waves, band, 0, 1);   // queued receivers
                if (fmt == SGPU_FMT_BARLE && waves == 4)   // queued receiver pairs
                    for (int band : {0, 32, 256, 512}) trial(K, waves, band, 0, 2);
            }
        const int K = t.hours_per_pass;
        for (int waves : {2, 4})   // the best waves count with receiver groups can differ
            for (int group : {8, 32, 64, 128})
                for (int band : {0, 256}) trial(K, waves, band, group, 0);
        t.voxel_hours_per_s = best_rate;
        c->irr_ok = false;   // the scratch passes overwrote the annual accumulators
    }
    c->tune = t;
    store(key, t);
    if (out) *out = t;
}

}  // namespace

/* Apply a cached tuning of this context's scene class, if any, without measuring. */
bool sgpu::apply_cached_tuning(sgpu_ctx* c, unsigned fmt) {
    sgpu_tuning t = c->tune;
    if (!lookup(tune_key(c, fmt), t)) return false;
    t.from_cache = 1;
    t.rays_per_s = t.voxel_hours_per_s = 0.0;
    c->tune = t;
    return true;
}

extern "C" {

int sgpu_autotune(sgpu_ctx* c, unsigned format, sgpu_tuning* out) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        autotune(c, format, out);
    });
}

int sgpu_set_tuning(sgpu_ctx* c, const sgpu_tuning* t) {
    return guarded([&] {
        need(c && t, SGPU_ERR_ARG, "null argument");
        need(t->dirs_per_group == 16 || t->dirs_per_group == 32 || t->dirs_per_group == 48, SGPU_ERR_ARG,
             "dirs_per_group must be 16, 32 or 48");
        need(supported_hours_per_pass(t->hours_per_pass), SGPU_ERR_ARG, "hours_per_pass must be 1, 2, 4, 8, 16, 32, 64, 128 or 256");
        need(t->waves_per_group == 2 || t->waves_per_group == 4, SGPU_ERR_ARG, "waves_per_group must be 2 or 4");
        need(t->xcd_band == 0 || (t->xcd_band >= 16 && t->xcd_band <= 4096), SGPU_ERR_ARG,
             "xcd_band must be 0 or 16..4096");
        need(t->receiver_queue >= 0 && t->receiver_queue <= 2, SGPU_ERR_ARG, "receiver_queue must be 0, 1 or 2");
        if (t->receiver_queue)
            need(t->group_blocks >= 0 && t->group_blocks <= 64, SGPU_ERR_ARG,
                 "group_blocks (workgroups per CU of the receiver queue) must be 0..64");
        else
            need(t->group_blocks == 0 || (t->group_blocks >= 2 && t->group_blocks <= 1024), SGPU_ERR_ARG,
                 "group_blocks must be 0 or 2..1024");
        c->tune = *t;
    });
}

int sgpu_get_tuning(sgpu_ctx* c, sgpu_tuning* out) {
    return guarded([&] {
        need(c && out, SGPU_ERR_ARG, "null argument");
        *out = c->tune;
    });
}

}  // extern "C"
