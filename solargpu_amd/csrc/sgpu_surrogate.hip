/* C ABI of the surrogate path (include/sgpu.h, "surrogate" section):
 *   prepare   sun-bin SVF + exact annual direct/diffuse + MLP features, on the GPU;
 *   labels    annual reflected irradiance of the plain method (training targets);
 *   run       features -> MLP (BF16 MFMA) -> gate -> exact plain fallback (full
 *             2592-direction trace and ba-RLE lists, then either the per-hour
 *             reflection loop or one walk over the neighbours' annual totals)
 *             for the voxels whose predicted error exceeds the threshold. */
#include <algorithm>
#include <vector>

#include "sgpu_ctx.h"
#include "sgpu_internal.h"
#include "sgpu_surrogate.h"

using namespace sgpu;

namespace {

SurArgs sur_args(sgpu_ctx* c) {
    SurArgs a{};
    a.grid = c->grid;
    a.W = c->W;
    a.H = c->H;
    a.n_total = c->n_total;
    a.n_top = c->n_top;
    a.pix_parent = c->pix_parent;
    a.wall_z = c->wall_z;
    a.slopes = c->slopes;
    a.aspects = c->aspects;
    a.maxz = c->maxz;
    a.dirs = c->dirs;
    a.sun_dirs = c->sur.sun_dirs;
    a.n_sun_dirs = c->sky.active;
    a.svf_bytes = c->sky.svf_bytes;
    a.svf = c->svf;
    a.sun_hits = c->sur.sun_hits;
    a.hours = c->hours_d;
    a.n_hours = (int)c->hours.size();
    a.dd = c->sur.dd;
    a.diff = c->sur.diff;
    a.X = c->sur.X;
    a.lat = c->lat;
    a.cell_m = c->cell_m;
    a.albedo = 0.2f;
    return a;
}

/* Sun-bin SVF, annual direct + diffuse and the feature matrix; returns the kernel ms of each. */
void prepare(sgpu_ctx* c, float lat, float cell_m, sgpu_surrogate_stats* st) {
    need(c->scene && c->sky_set, SGPU_ERR_STATE, "scene/sky not set");
    need(cell_m > 0.0f, SGPU_ERR_ARG, "cell size must be positive");
    const int64_t N = c->n_total;
    c->lat = lat;
    c->cell_m = cell_m;
    if (!c->svf) c->svf = dalloc<uint8_t>((size_t)std::max(c->sky.svf_bytes, 1) * N);
    SurrogateState& s = c->sur;
    if (!s.X) {
        s.X = dalloc<uint16_t>((size_t)N * kSurInPad);
        s.dd = dalloc<float>(N);
        s.diff = dalloc<float>(N);
        s.sun_hits = dalloc<uint16_t>(N);
        s.sun_dirs = dalloc<int16_t>(std::max(c->sky.active, 1));
    }
    std::vector<int16_t> sd(std::max(c->sky.active, 1), 0);
    for (int i = 0; i < kAzBins; ++i)
        for (int j = 0; j < kElBins; ++j) {
            const int dense = c->sky.full_to_dense[i * kElBins + j];
            if (dense >= 0) sd[dense] = (int16_t)(i * 2 * kElBins + j);   // upper direction of that bin
        }
    SGPU_HIP(hipMemcpyAsync(s.sun_dirs, sd.data(), sd.size() * sizeof(int16_t), hipMemcpyHostToDevice, c->stream));
    ensure_hours_d(c, c->hours);
    SurArgs a = sur_args(c);
    EventPair e1, e2, e3;
    SGPU_HIP(hipEventRecord(e1.a, c->stream));
    launch_sun_svf(a, c->stream);
    SGPU_HIP(hipEventRecord(e1.b, c->stream));
    SGPU_HIP(hipEventRecord(e2.a, c->stream));
    launch_annual_dd(a, c->stream);
    SGPU_HIP(hipEventRecord(e2.b, c->stream));
    SGPU_HIP(hipEventRecord(e3.a, c->stream));
    if (!s.hz) {
        s.hz_cap = std::min<int64_t>(N, (int64_t)1 << 24);   // 4 GB of march results at most
        s.hz = dalloc<int2>((size_t)s.hz_cap * kSectors);
    }
    launch_features(a, s.hz, s.hz_cap, c->stream);
    SGPU_HIP(hipEventRecord(e3.b, c->stream));
    SGPU_HIP(hipStreamSynchronize(c->stream));
    c->svf_ok = true;
    s.ready = true;
    if (st) {
        st->t_svf_ms = e1.ms();
        st->t_dd_ms = e2.ms();
        st->t_features_ms = e3.ms();
    }
}

}  // namespace

extern "C" {

int sgpu_surrogate_prepare(sgpu_ctx* c, float lat, float cell_m, sgpu_surrogate_stats* st) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        c->set_device();
        prepare(c, lat, cell_m, st);
    });
}

int sgpu_surrogate_get(sgpu_ctx* c, uint16_t* X, float* dd, float* diff) {
    return guarded([&] {
        need(c != nullptr, SGPU_ERR_ARG, "null context");
        need(c->sur.ready, SGPU_ERR_STATE, "run sgpu_surrogate_prepare first");
        c->set_device();
        const int64_t N = c->n_total;
        if (X) SGPU_HIP(hipMemcpyAsync(X, c->sur.X, (size_t)N * kSurInPad * 2, hipMemcpyDeviceToHost, c->stream));
        if (dd) SGPU_HIP(hipMemcpyAsync(dd, c->sur.dd, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
        if (diff) SGPU_HIP(hipMemcpyAsync(diff, c->sur.diff, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
        SGPU_HIP(hipStreamSynchronize(c->stream));
    });
}

int sgpu_surrogate_labels(sgpu_ctx* c, unsigned fmt, float* reflected) {
    return guarded([&] {
        need(c && reflected, SGPU_ERR_ARG, "null argument");
        need(c->sur.ready, SGPU_ERR_STATE, "run sgpu_surrogate_prepare first");
        c->set_device();
        float* out = dalloc<float>(c->n_total);
        try {
            annual_reflected_linear(c, fmt, c->sur.dd, out);
            SGPU_HIP(hipMemcpyAsync(reflected, out, (size_t)c->n_total * 4, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
        } catch (...) {
            dfree(out);
            throw;
        }
        dfree(out);
    });
}

int sgpu_surrogate_run(sgpu_ctx* c, sgpu_mlp* mlp, sgpu_mlp* error_mlp, const sgpu_gate* gate, int hours_per_pass,
                       float* out_host, int32_t* fallback_host, sgpu_surrogate_stats* st) {
    return guarded([&] {
        need(c && mlp && gate && out_host, SGPU_ERR_ARG, "null argument");
        c->set_device();
        const int64_t N = c->n_total;
        const bool annual = hours_per_pass == SGPU_FALLBACK_ANNUAL;
        need(annual || hours_per_pass >= 0, SGPU_ERR_ARG, "hours_per_pass must be >= 0 or SGPU_FALLBACK_ANNUAL");
        const int K = hours_per_pass > 0 ? hours_per_pass : c->tune.hours_per_pass;
        sgpu_surrogate_stats local{};
        sgpu_surrogate_stats* s = st ? st : &local;
        *s = sgpu_surrogate_stats{};
        EventPair all, ef;
        SGPU_HIP(hipEventRecord(all.a, c->stream));
        prepare(c, c->lat, c->cell_m, s);
        float* Y = dalloc<float>((size_t)N * 2);
        float* Yerr = nullptr;
        float* out = nullptr;
        int32_t* flags = nullptr;
        int32_t* pflags = nullptr;
        int64_t* off = nullptr;
        int64_t* scr = nullptr;
        int32_t* ids = nullptr;
        float* exact = nullptr;
        SubsetLists L;
        try {
            if (error_mlp) Yerr = dalloc<float>((size_t)N * 2);
            out = dalloc<float>(N);
            flags = dalloc<int32_t>(N);
            pflags = dalloc<int32_t>(N);
            off = dalloc<int64_t>(N + 1);
            scr = dalloc<int64_t>(scan_scratch_elems(N));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            double mlp_ms = 0.0;
            if (sgpu_mlp_forward_device(mlp, c->sur.X, N, Y, &mlp_ms) != SGPU_OK)
                fail(SGPU_ERR_HIP, "MLP forward failed: %s", sgpu_mlp_last_error());
            s->t_mlp_ms = mlp_ms;
            if (error_mlp) {
                if (sgpu_mlp_forward_device(error_mlp, c->sur.X, N, Yerr, &mlp_ms) != SGPU_OK)
                    fail(SGPU_ERR_HIP, "error-network forward failed: %s", sgpu_mlp_last_error());
                s->t_mlp_ms += mlp_ms;
            }
            SurArgs a = sur_args(c);
            a.Y = Y;
            a.Yerr = Yerr;
            a.out = out;
            a.fallback = flags;
            a.tau = gate->tau;
            a.refl_rel_err = gate->refl_rel_err;
            a.diff_miss_weight = gate->diff_miss_weight;
            a.share_abs_err = gate->share_abs_err;
            launch_gate(a, c->stream);
            /* the gated voxels in the reflection pass's processing order, so consecutive receivers of
               the fallback are neighbours and share their gathered rows in the L2 */
            launch_permute_flags(flags, c->order, N, pflags, c->stream);
            launch_scan_counts(pflags, N, off, scr, c->stream);
            int64_t nf = 0;
            SGPU_HIP(hipMemcpyAsync(&nf, off + N, sizeof nf, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            s->fallback = nf;
            SGPU_HIP(hipEventRecord(ef.a, c->stream));
            if (nf > 0) {
                ids = dalloc<int32_t>(nf);
                exact = dalloc<float>(nf);
                launch_compact(pflags, c->order, off, N, ids, c->stream);
                L = subset_viewshed(c, ids, nf, true);
                if (annual) {
                    subset_reflected_linear(c, ids, nf, L, SGPU_FMT_BARLE, c->sur.dd, exact);
                    launch_scatter_sum(ids, exact, c->sur.dd, nf, out, c->stream);
                } else {
                    subset_hour_loop(c, ids, nf, L, SGPU_FMT_BARLE, K, exact);
                    launch_scatter(ids, exact, nf, out, c->stream);
                }
            }
            SGPU_HIP(hipEventRecord(ef.b, c->stream));
            SGPU_HIP(hipEventRecord(all.b, c->stream));
            SGPU_HIP(hipMemcpyAsync(out_host, out, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
            if (fallback_host)
                SGPU_HIP(hipMemcpyAsync(fallback_host, flags, (size_t)N * 4, hipMemcpyDeviceToHost, c->stream));
            SGPU_HIP(hipStreamSynchronize(c->stream));
            s->t_fallback_ms = ef.ms();
            s->t_total_ms = all.ms();
        } catch (...) {
            L.release();
            dfree(Y); dfree(Yerr); dfree(out); dfree(flags); dfree(pflags); dfree(off); dfree(scr); dfree(ids); dfree(exact);
            throw;
        }
        L.release();
        dfree(Y); dfree(Yerr); dfree(out); dfree(flags); dfree(pflags); dfree(off); dfree(scr); dfree(ids); dfree(exact);
    });
}

}  // extern "C"
