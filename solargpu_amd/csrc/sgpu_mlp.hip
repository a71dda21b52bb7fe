/* Host side of the surrogate MLP (C ABI sgpu_mlp_*): parameters, training
 * loop (minibatch gather -> forward -> backward -> Adam, all on one stream),
 * batched inference and a one-batch gradient for the gradient check. */
#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>
#include <numeric>
#include <string>
#include <vector>

#include "sgpu_internal.h"
#include "sgpu_mlp.h"

using namespace sgpu;

namespace {

template <typename T>
T* dnew(size_t n) {
    void* p = nullptr;
    SGPU_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
    return static_cast<T*>(p);
}
template <typename T>
void dfree(T*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

uint16_t host_bf16(float x) {
    uint32_t u;
    std::memcpy(&u, &x, 4);
    return (uint16_t)((u + 0x7FFFu + ((u >> 16) & 1u)) >> 16);
}

struct Layer {
    int in, out;          // in is padded (64 for layer 1)
    float *w, *b, *mw, *vw, *mb, *vb;
    uint16_t *wb, *wt;    // bf16 [out][in], [in][out]
    float *gw, *gb;       // gradient buffers: gw [splits][out][in], gb [out]
};

__global__ void k_gather(const uint16_t* X, const float* Y, const int64_t* perm, int64_t b0, int B, uint16_t* xb,
                         uint16_t* xbt, float* tb) {
    const int b = blockIdx.x, f = threadIdx.x;   // 64 threads: one feature column each
    const int64_t src = perm[b0 + b];
    const uint16_t v = X[src * kInPad + f];
    xb[(int64_t)b * kInPad + f] = v;
    xbt[(int64_t)f * B + b] = v;
    if (f < kOutputs) tb[2 * b + f] = Y[2 * src + f];
}

__global__ void k_zero_f(float* p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0.0f;
}

}  // namespace

struct sgpu_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    Layer L[kLayers];
    int step = 0;
    int splits = 8;
    /* activation workspace for a batch of cap rows */
    int cap = 0;
    uint16_t *xb = nullptr, *xbt = nullptr;
    uint16_t *Z[4] = {}, *Hh[4] = {}, *Ht[4] = {}, *dZ[4] = {}, *dZt[4] = {};
    float* tb = nullptr;
    double* loss = nullptr;
    float out_wt[2] = {1.0f, 1.0f};
    int infer_rows = 64;   // inference: fused kernel with tiles of 32 or 64 rows; 0 = layer by layer

    void free_work() {
        dfree(xb); dfree(xbt); dfree(tb); dfree(loss);
        for (int l = 0; l < 4; ++l) { dfree(Z[l]); dfree(Hh[l]); dfree(Ht[l]); dfree(dZ[l]); dfree(dZt[l]); }
        cap = 0;
    }
    void ensure_work(int B) {
        if (B <= cap) return;
        free_work();
        xb = dnew<uint16_t>((size_t)B * kInPad);
        xbt = dnew<uint16_t>((size_t)B * kInPad);
        tb = dnew<float>((size_t)B * 2);
        loss = dnew<double>(1);
        for (int l = 0; l < 4; ++l) {
            Z[l] = dnew<uint16_t>((size_t)B * kHidden);
            Hh[l] = dnew<uint16_t>((size_t)B * kHidden);
            Ht[l] = dnew<uint16_t>((size_t)B * kHidden);
            dZ[l] = dnew<uint16_t>((size_t)B * kHidden);
            dZt[l] = dnew<uint16_t>((size_t)B * kHidden);
        }
        cap = B;
    }
    ~sgpu_mlp() {
        (void)hipSetDevice(device);
        free_work();
        for (Layer& l : L) {
            dfree(l.w); dfree(l.b); dfree(l.mw); dfree(l.vw); dfree(l.mb); dfree(l.vb);
            dfree(l.wb); dfree(l.wt); dfree(l.gw); dfree(l.gb);
        }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

thread_local std::string g_mlp_err;

template <typename F>
int mlp_guard(F&& fn) {
    try {
        fn();
        return SGPU_OK;
    } catch (const Failure& f) {
        g_mlp_err = f.what();
        return f.code;
    } catch (const std::exception& e) {
        g_mlp_err = e.what();
        return SGPU_ERR_HIP;
    }
}

void upload_layer(sgpu_mlp* m, int li, const float* w, const float* b) {
    Layer& l = m->L[li];
    const size_t n = (size_t)l.in * l.out;
    std::vector<uint16_t> wb(n), wt(n);
    for (int o = 0; o < l.out; ++o)
        for (int i = 0; i < l.in; ++i) {
            wb[(size_t)o * l.in + i] = host_bf16(w[(size_t)o * l.in + i]);
            wt[(size_t)i * l.out + o] = wb[(size_t)o * l.in + i];
        }
    SGPU_HIP(hipMemcpyAsync(l.w, w, n * 4, hipMemcpyHostToDevice, m->stream));
    SGPU_HIP(hipMemcpyAsync(l.b, b, (size_t)l.out * 4, hipMemcpyHostToDevice, m->stream));
    SGPU_HIP(hipMemcpyAsync(l.wb, wb.data(), n * 2, hipMemcpyHostToDevice, m->stream));
    SGPU_HIP(hipMemcpyAsync(l.wt, wt.data(), n * 2, hipMemcpyHostToDevice, m->stream));
    SGPU_HIP(hipStreamSynchronize(m->stream));
}

InferArgs infer_args(const sgpu_mlp* m, const uint16_t* x, int64_t B, float* y_dev) {
    InferArgs f{};
    f.X = x;
    for (int l = 0; l < 4; ++l) {
        f.W[l] = m->L[l].wb;
        f.b[l] = m->L[l].b;
    }
    f.W5 = m->L[4].wb;
    f.b5 = m->L[4].b;
    f.B = B;
    f.Y = y_dev;
    return f;
}

/* Forward over B rows of x (device, [B][64]).  Training also keeps Z and H^T. */
void forward(sgpu_mlp* m, const uint16_t* x, int B, bool train, float* y_dev) {
    if (!train && m->infer_rows) {
        launch_infer(infer_args(m, x, B, y_dev), m->infer_rows, m->stream);
        return;
    }
    const uint16_t* a = x;
    int lda = kInPad;
    for (int l = 0; l < 4; ++l) {
        GemmArgs g{};
        g.A = a;
        g.lda = lda;
        g.Bt = m->L[l].wb;
        g.ldb = m->L[l].in;
        g.M = B;
        g.N = kHidden;
        g.K = m->L[l].in;
        g.k_begin = 0;
        g.k_len = g.K;
        g.bias = m->L[l].b;
        g.Z = train ? m->Z[l] : nullptr;
        g.H = m->Hh[l];
        g.Ht = train ? m->Ht[l] : nullptr;
        g.ldt = B;
        launch_gemm(g, kEpiForward, 1, m->stream);
        a = m->Hh[l];
        lda = kHidden;
    }
    if (!train) {
        OutArgs o{};
        o.H4 = m->Hh[3];
        o.W5 = m->L[4].wb;
        o.b5 = m->L[4].b;
        o.B = B;
        o.Y = y_dev;
        launch_out(o, false, m->stream);
    }
}

/* Backward for a batch already in xb/xbt/tb; gradients land in L[].gw/gb. */
void backward(sgpu_mlp* m, int B) {
    for (Layer& l : m->L) {
        const int64_t nw = (int64_t)m->splits * l.in * l.out;
        hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((nw + 255) / 256)), dim3(256), 0, m->stream, l.gw, nw);
        hipLaunchKernelGGL(k_zero_f, dim3((unsigned)((l.out + 255) / 256)), dim3(256), 0, m->stream, l.gb,
                           (int64_t)l.out);
    }
    OutArgs o{};
    o.H4 = m->Hh[3];
    o.Z4 = m->Z[3];
    o.W5 = m->L[4].wb;
    o.b5 = m->L[4].b;
    o.B = B;
    o.T = m->tb;
    o.wt[0] = m->out_wt[0];
    o.wt[1] = m->out_wt[1];
    o.scale = 1.0f / (float)B;
    o.dZ4 = m->dZ[3];
    o.dZ4t = m->dZt[3];
    o.ldt = B;
    o.gW5 = m->L[4].gw;
    o.gb5 = m->L[4].gb;
    o.loss = m->loss;
    launch_out(o, true, m->stream);
    for (int l = 3; l >= 0; --l) {
        /* weight gradient: dW[out][in] = dZ^T . Hprev, K = batch, split over the batch */
        GemmArgs g{};
        g.A = m->dZt[l];
        g.lda = B;
        g.Bt = (l == 0) ? m->xbt : m->Ht[l - 1];
        g.ldb = B;
        g.M = kHidden;
        g.N = m->L[l].in;
        g.K = B;
        g.k_begin = 0;
        g.k_len = ((B / m->splits + 31) / 32) * 32;
        const int sp = (B + g.k_len - 1) / g.k_len;
        g.Cf = m->L[l].gw;
        launch_gemm(g, kEpiPartial, sp, m->stream);
        launch_bias_grad(m->dZt[l], B, B, kHidden, m->L[l].gb, m->stream);
        if (l > 0) {
            GemmArgs d{};
            d.A = m->dZ[l];
            d.lda = kHidden;
            d.Bt = m->L[l].wt;
            d.ldb = kHidden;
            d.M = B;
            d.N = kHidden;
            d.K = kHidden;
            d.k_begin = 0;
            d.k_len = kHidden;
            d.Zin = m->Z[l - 1];
            d.dZ = m->dZ[l - 1];
            d.dZt = m->dZt[l - 1];
            d.ldt = B;
            launch_gemm(d, kEpiBackward, 1, m->stream);
        }
    }
}

void adam_step(sgpu_mlp* m, float lr) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    ++m->step;
    const float bc1 = 1.0f / (1.0f - std::pow(b1, (float)m->step));
    const float bc2 = 1.0f / (1.0f - std::pow(b2, (float)m->step));
    for (int li = 0; li < kLayers; ++li) {
        Layer& l = m->L[li];
        AdamArgs a{};
        a.w = l.w; a.m = l.mw; a.v = l.vw; a.grad = l.gw;
        a.splits = (li == kLayers - 1) ? 1 : m->splits;
        a.grad_scale = 1.0f;
        a.n = (int64_t)l.in * l.out;
        a.rows = l.out; a.cols = l.in;
        a.wb = l.wb; a.wt = l.wt;
        a.lr = lr; a.beta1 = b1; a.beta2 = b2; a.eps = eps; a.bc1 = bc1; a.bc2 = bc2;
        launch_adam(a, m->stream);
        AdamArgs ab = a;
        ab.w = l.b; ab.m = l.mb; ab.v = l.vb; ab.grad = l.gb; ab.splits = 1;
        ab.n = l.out; ab.rows = 1; ab.cols = l.out; ab.wb = nullptr; ab.wt = nullptr;
        launch_adam(ab, m->stream);
    }
}

/* Only the used part of the split-K partial buffers must be summed; unused splits stay zero. */

}  // namespace

extern "C" {

const char* sgpu_mlp_last_error(void) { return g_mlp_err.c_str(); }

int sgpu_mlp_create(int device, uint64_t seed, sgpu_mlp** out) {
    return mlp_guard([&] {
        if (!out) fail(SGPU_ERR_ARG, "null output");
        std::unique_ptr<sgpu_mlp> m(new sgpu_mlp());
        m->device = device;
        SGPU_HIP(hipSetDevice(device));
        SGPU_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
        const int ins[kLayers] = {kInPad, kHidden, kHidden, kHidden, kHidden};
        const int outs[kLayers] = {kHidden, kHidden, kHidden, kHidden, kOutputs};
        uint64_t s = seed * 0x9E3779B97F4A7C15ull + 1;
        auto uni = [&]() {
            s ^= s << 13; s ^= s >> 7; s ^= s << 17;
            return (double)(s >> 11) * (1.0 / 9007199254740992.0);
        };
        for (int li = 0; li < kLayers; ++li) {
            Layer& l = m->L[li];
            l.in = ins[li];
            l.out = outs[li];
            const size_t n = (size_t)l.in * l.out;
            l.w = dnew<float>(n); l.mw = dnew<float>(n); l.vw = dnew<float>(n);
            l.b = dnew<float>(l.out); l.mb = dnew<float>(l.out); l.vb = dnew<float>(l.out);
            l.wb = dnew<uint16_t>(n); l.wt = dnew<uint16_t>(n);
            l.gw = dnew<float>((size_t)m->splits * n); l.gb = dnew<float>(l.out);
            SGPU_HIP(hipMemsetAsync(l.mw, 0, n * 4, m->stream));
            SGPU_HIP(hipMemsetAsync(l.vw, 0, n * 4, m->stream));
            SGPU_HIP(hipMemsetAsync(l.mb, 0, l.out * 4, m->stream));
            SGPU_HIP(hipMemsetAsync(l.vb, 0, l.out * 4, m->stream));
            /* He-uniform on the used fan-in; padded input columns start at zero */
            const int fan = (li == 0) ? kFeatures : l.in;
            const double lim = std::sqrt(6.0 / fan) * (li == kLayers - 1 ? 0.1 : 1.0);
            std::vector<float> w(n), b(l.out, 0.0f);
            for (int o = 0; o < l.out; ++o)
                for (int i = 0; i < l.in; ++i)
                    w[(size_t)o * l.in + i] = (li == 0 && i >= kFeatures) ? 0.0f : (float)((2.0 * uni() - 1.0) * lim);
            upload_layer(m.get(), li, w.data(), b.data());
        }
        *out = m.release();
    });
}

int sgpu_mlp_destroy(sgpu_mlp* m) {
    return mlp_guard([&] { delete m; });
}

int sgpu_mlp_param_count(void) {
    return kInPad * kHidden + kHidden + 3 * (kHidden * kHidden + kHidden) + kHidden * kOutputs + kOutputs;
}

/* Parameters as one fp32 vector: for each layer W[out][in] then b[out]. */
int sgpu_mlp_get_params(sgpu_mlp* m, float* out) {
    return mlp_guard([&] {
        if (!m || !out) fail(SGPU_ERR_ARG, "null argument");
        SGPU_HIP(hipSetDevice(m->device));
        size_t pos = 0;
        for (Layer& l : m->L) {
            const size_t n = (size_t)l.in * l.out;
            SGPU_HIP(hipMemcpyAsync(out + pos, l.w, n * 4, hipMemcpyDeviceToHost, m->stream));
            pos += n;
            SGPU_HIP(hipMemcpyAsync(out + pos, l.b, (size_t)l.out * 4, hipMemcpyDeviceToHost, m->stream));
            pos += l.out;
        }
        SGPU_HIP(hipStreamSynchronize(m->stream));
    });
}

int sgpu_mlp_set_params(sgpu_mlp* m, const float* p) {
    return mlp_guard([&] {
        if (!m || !p) fail(SGPU_ERR_ARG, "null argument");
        SGPU_HIP(hipSetDevice(m->device));
        size_t pos = 0;
        for (int li = 0; li < kLayers; ++li) {
            Layer& l = m->L[li];
            const size_t n = (size_t)l.in * l.out;
            upload_layer(m, li, p + pos, p + pos + n);
            pos += n + l.out;
        }
        m->step = 0;
    });
}

int sgpu_mlp_set_loss_weights(sgpu_mlp* m, float w0, float w1) {
    return mlp_guard([&] {
        if (!m) fail(SGPU_ERR_ARG, "null argument");
        m->out_wt[0] = w0;
        m->out_wt[1] = w1;
    });
}

int sgpu_mlp_set_inference(sgpu_mlp* m, int rows) {
    return mlp_guard([&] {
        if (!m) fail(SGPU_ERR_ARG, "null argument");
        if (rows != 0 && rows != 32 && rows != 64) fail(SGPU_ERR_ARG, "rows must be 0, 32 or 64");
        m->infer_rows = rows;
    });
}

/* Inference: X [n][64] bf16 (host), Y [n][2] (host). */
int sgpu_mlp_forward(sgpu_mlp* m, const uint16_t* X, int64_t n, float* Y) {
    return mlp_guard([&] {
        if (!m || !X || !Y || n < 0) fail(SGPU_ERR_ARG, "bad argument");
        SGPU_HIP(hipSetDevice(m->device));
        const int chunk = 65536;
        if (!m->infer_rows) m->ensure_work(chunk);
        uint16_t* xd = dnew<uint16_t>((size_t)chunk * kInPad);
        float* yd = dnew<float>((size_t)chunk * 2);
        for (int64_t r = 0; r < n; r += chunk) {
            const int B = (int)std::min<int64_t>(chunk, n - r);
            SGPU_HIP(hipMemcpyAsync(xd, X + r * kInPad, (size_t)B * kInPad * 2, hipMemcpyHostToDevice, m->stream));
            forward(m, xd, B, false, yd);
            SGPU_HIP(hipMemcpyAsync(Y + 2 * r, yd, (size_t)B * 8, hipMemcpyDeviceToHost, m->stream));
        }
        SGPU_HIP(hipStreamSynchronize(m->stream));
        dfree(xd);
        dfree(yd);
    });
}

/* Device-resident inference (X, Y device pointers); returns the kernel time in ms. */
int sgpu_mlp_forward_device(sgpu_mlp* m, const uint16_t* X, int64_t n, float* Y, double* ms) {
    return mlp_guard([&] {
        if (!m || !X || !Y) fail(SGPU_ERR_ARG, "bad argument");
        SGPU_HIP(hipSetDevice(m->device));
        const int chunk = 1 << 18;
        if (!m->infer_rows) m->ensure_work((int)std::min<int64_t>(chunk, std::max<int64_t>(n, 64)));
        hipEvent_t a, b;
        SGPU_HIP(hipEventCreate(&a));
        SGPU_HIP(hipEventCreate(&b));
        SGPU_HIP(hipEventRecord(a, m->stream));
        if (m->infer_rows) {
            launch_infer(infer_args(m, X, n, Y), m->infer_rows, m->stream);   // every row in one launch
        } else {
            for (int64_t r = 0; r < n; r += m->cap) {
                const int B = (int)std::min<int64_t>(m->cap, n - r);
                forward(m, X + r * kInPad, B, false, Y + 2 * r);
            }
        }
        SGPU_HIP(hipEventRecord(b, m->stream));
        SGPU_HIP(hipEventSynchronize(b));
        float t = 0.0f;
        SGPU_HIP(hipEventElapsedTime(&t, a, b));
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        if (ms) *ms = t;
    });
}

/* Training on host data X [n][64] bf16, T [n][2]; minibatches of `batch` rows (multiple of 64).
   losses[epoch] = mean weighted squared error of the epoch. */
int sgpu_mlp_train(sgpu_mlp* m, const uint16_t* X, const float* T, int64_t n, int epochs, int batch, float lr,
                   uint64_t seed, double* losses) {
    return mlp_guard([&] {
        if (!m || !X || !T || n < batch || batch < 64 || batch % 64) fail(SGPU_ERR_ARG, "bad training arguments");
        SGPU_HIP(hipSetDevice(m->device));
        m->ensure_work(batch);
        uint16_t* xd = dnew<uint16_t>((size_t)n * kInPad);
        float* td = dnew<float>((size_t)n * 2);
        int64_t* perm_d = dnew<int64_t>((size_t)n);
        SGPU_HIP(hipMemcpyAsync(xd, X, (size_t)n * kInPad * 2, hipMemcpyHostToDevice, m->stream));
        SGPU_HIP(hipMemcpyAsync(td, T, (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
        std::vector<int64_t> perm(n);
        std::iota(perm.begin(), perm.end(), 0);
        uint64_t s = seed * 0x2545F4914F6CDD1Dull + 7;
        const int64_t nb = n / batch;
        for (int ep = 0; ep < epochs; ++ep) {
            for (int64_t i = n - 1; i > 0; --i) {   // Fisher-Yates
                s ^= s << 13; s ^= s >> 7; s ^= s << 17;
                std::swap(perm[i], perm[s % (uint64_t)(i + 1)]);
            }
            SGPU_HIP(hipMemcpyAsync(perm_d, perm.data(), (size_t)n * 8, hipMemcpyHostToDevice, m->stream));
            SGPU_HIP(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
            for (int64_t bi = 0; bi < nb; ++bi) {
                hipLaunchKernelGGL(k_gather, dim3(batch), dim3(kInPad), 0, m->stream, xd, td, perm_d, bi * batch, batch,
                                   m->xb, m->xbt, m->tb);
                forward(m, m->xb, batch, true, nullptr);
                backward(m, batch);
                adam_step(m, lr);
            }
            double l = 0.0;
            SGPU_HIP(hipMemcpyAsync(&l, m->loss, sizeof l, hipMemcpyDeviceToHost, m->stream));
            SGPU_HIP(hipStreamSynchronize(m->stream));
            if (losses) losses[ep] = l / (double)(nb * batch);
        }
        dfree(xd);
        dfree(td);
        dfree(perm_d);
    });
}

/* Gradient of the mean weighted squared error over one batch (rows 0..B-1 of X, T),
   in the parameter order of sgpu_mlp_get_params. */
int sgpu_mlp_gradient(sgpu_mlp* m, const uint16_t* X, const float* T, int B, float* grad, double* loss) {
    return mlp_guard([&] {
        if (!m || !X || !T || !grad || B < 64 || B % 64) fail(SGPU_ERR_ARG, "bad gradient arguments");
        SGPU_HIP(hipSetDevice(m->device));
        m->ensure_work(B);
        int64_t* perm_d = dnew<int64_t>(B);
        uint16_t* xd = dnew<uint16_t>((size_t)B * kInPad);
        float* td = dnew<float>((size_t)B * 2);
        std::vector<int64_t> perm(B);
        std::iota(perm.begin(), perm.end(), 0);
        SGPU_HIP(hipMemcpyAsync(perm_d, perm.data(), (size_t)B * 8, hipMemcpyHostToDevice, m->stream));
        SGPU_HIP(hipMemcpyAsync(xd, X, (size_t)B * kInPad * 2, hipMemcpyHostToDevice, m->stream));
        SGPU_HIP(hipMemcpyAsync(td, T, (size_t)B * 8, hipMemcpyHostToDevice, m->stream));
        SGPU_HIP(hipMemsetAsync(m->loss, 0, sizeof(double), m->stream));
        hipLaunchKernelGGL(k_gather, dim3(B), dim3(kInPad), 0, m->stream, xd, td, perm_d, (int64_t)0, B, m->xb, m->xbt,
                           m->tb);
        forward(m, m->xb, B, true, nullptr);
        backward(m, B);
        size_t pos = 0;
        for (int li = 0; li < kLayers; ++li) {
            Layer& l = m->L[li];
            const size_t nw = (size_t)l.in * l.out;
            const int sp = (li == kLayers - 1) ? 1 : m->splits;
            std::vector<float> part((size_t)sp * nw);
            SGPU_HIP(hipMemcpyAsync(part.data(), l.gw, part.size() * 4, hipMemcpyDeviceToHost, m->stream));
            SGPU_HIP(hipMemcpyAsync(grad + pos + nw, l.gb, (size_t)l.out * 4, hipMemcpyDeviceToHost, m->stream));
            SGPU_HIP(hipStreamSynchronize(m->stream));
            for (size_t i = 0; i < nw; ++i) {
                float g = 0.0f;
                for (int s2 = 0; s2 < sp; ++s2) g += part[(size_t)s2 * nw + i];
                grad[pos + i] = g;
            }
            pos += nw + l.out;
        }
        if (loss) {
            SGPU_HIP(hipMemcpyAsync(loss, m->loss, sizeof(double), hipMemcpyDeviceToHost, m->stream));
            SGPU_HIP(hipStreamSynchronize(m->stream));
            *loss /= B;
        }
        dfree(perm_d);
        dfree(xd);
        dfree(td);
    });
}

}  // extern "C"
