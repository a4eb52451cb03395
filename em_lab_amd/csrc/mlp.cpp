// libemfdtd.so host side of the MLP surrogate (emfd_mlp_*): parameter layout, activation
// buffers, the forward / backward launch sequence and Adam. Device work is in mlp.hip.
//
// Device layout: layer l has input width kp[l] (the logical width rounded up to 32, zero padded)
// and output width dims[l + 1]. Master parameters (FP32) are flat: W_l [n_out][kp[l]] then
// b_l [n_out] per layer; gradients and Adam moments share that layout; the BF16 copies hold the
// weights only. Padded weight columns stay zero: their inputs are zero, so are their gradients.
#include "context.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <memory>

using namespace emfd;

struct emfd_mlp {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<int> dims, kp;
    int layers = 0;
    int max_batch = 0;
    std::vector<size_t> w_off, b_off, wb_off;
    size_t n_params = 0, n_wb = 0, n_logical = 0;
    DevBuf p, g, m1, m2, bf_index, wb;
    // activations
    DevBuf xb, tb, y, idx;
    std::vector<DevBuf> z, h;  // per hidden layer: pre-activation FP32, activation BF16
    DevBuf gf[2], gb[2];       // gradient wrt pre-activation, ping-pong (FP32, BF16)
    DevBuf partials;
    int adam_t = 0;
    float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, wd = 0.0f;
    // device training set
    DevBuf data_x, data_t;
    int64_t data_n = 0;
    uint64_t shuffle_state = 0;

    int n_in() const { return dims.front(); }
    int n_out() const { return dims.back(); }
    int widest() const { return *std::max_element(dims.begin() + 1, dims.end()); }
    void release_all() {
        for (DevBuf *d : {&p, &g, &m1, &m2, &bf_index, &wb, &xb, &tb, &y, &idx, &partials, &data_x, &data_t, &gf[0],
                          &gf[1], &gb[0], &gb[1]})
            d->release();
        for (auto &b : z) b.release();
        for (auto &b : h) b.release();
    }
    int forward(int rows, bool keep_preactivation = true);
    bool fused_ok = false;  // the fused all-layer inference kernel fits these widths
    int forward_infer(const float *dx, int rows);
    int backward(int rows, double *loss_partials_dst, int blocks);
};

namespace {
constexpr int kLossBlocks = 64;

int round32(int v) { return (v + 31) / 32 * 32; }

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
}  // namespace

// Forward over `rows` rows already in xb: hidden layers write h (BF16) and, for training, z
// (FP32 pre-activation); the last layer writes y (FP32).
int emfd_mlp::forward(int rows, bool keep_preactivation) {
    const void *in = xb.ptr;
    for (int l = 0; l < layers; ++l) {
        const bool last = l == layers - 1;
        MlpGemm gm{};
        gm.A = in;
        gm.B = static_cast<const uint16_t *>(wb.ptr) + wb_off[static_cast<size_t>(l)];
        gm.a_t = false;
        gm.b_t = true;
        gm.M = rows;
        gm.N = dims[static_cast<size_t>(l) + 1];
        gm.K = kp[static_cast<size_t>(l)];
        gm.lda = kp[static_cast<size_t>(l)];
        gm.ldb = kp[static_cast<size_t>(l)];
        gm.epi = last ? MLP_EPI_BIAS : (keep_preactivation ? MLP_EPI_BIAS_GELU : MLP_EPI_BIAS_GELU_INFER);
        gm.bias = p.as<float>() + b_off[static_cast<size_t>(l)];
        gm.out_f32 = last ? y.as<float>() : z[static_cast<size_t>(l)].as<float>();
        gm.out_bf16 = last ? nullptr : h[static_cast<size_t>(l)].ptr;
        gm.ldc = gm.N;
        EMFD_HIP(launch_mlp_gemm(gm, stream));
        if (!last) in = h[static_cast<size_t>(l)].ptr;
    }
    return EMFD_OK;
}

// Inference over `rows` FP32 device rows dx into y: the fused all-layer kernel when the widths
// fit it, else the layered kernels.
int emfd_mlp::forward_infer(const float *dx, int rows) {
    if (fused_ok) {
        MlpFusedArgs fa{};
        fa.x = dx;
        fa.y = y.as<float>();
        fa.M = rows;
        fa.n_in = n_in();
        fa.layers = layers;
        for (int l = 0; l < layers; ++l) {
            fa.kp[l] = kp[static_cast<size_t>(l)];
            fa.n_out[l] = dims[static_cast<size_t>(l) + 1];
            fa.w[l] = static_cast<const __bf16 *>(wb.ptr) + wb_off[static_cast<size_t>(l)];
            fa.bias[l] = p.as<float>() + b_off[static_cast<size_t>(l)];
        }
        EMFD_HIP(launch_mlp_forward_fused(fa, stream));
        return EMFD_OK;
    }
    EMFD_HIP(launch_mlp_gather(dx, n_in(), kp[0], nullptr, n_out(), nullptr, rows, xb.ptr, tb.as<float>(), stream));
    return forward(rows, false);
}

// Loss gradient from y and tb, then the backward pass into g; the loss partial sums go to
// loss_partials_dst (kLossBlocks doubles, device memory).
int emfd_mlp::backward(int rows, double *loss_partials_dst, int blocks) {
    const size_t nout = static_cast<size_t>(rows) * n_out();
    const float scale = 2.0f / static_cast<float>(nout);
    int cur = 0;
    EMFD_HIP(launch_mlp_mse(y.as<float>(), tb.as<float>(), nout, scale, gf[cur].as<float>(), gb[cur].ptr,
                            loss_partials_dst, blocks, stream));
    for (int l = layers - 1; l >= 0; --l) {
        const size_t L = static_cast<size_t>(l);
        const int nl = dims[L + 1];
        const void *in = l == 0 ? xb.ptr : h[L - 1].ptr;
        // dW_l[n][k] = sum_m G[m][n] in[m][k]
        MlpGemm dw{};
        dw.A = gb[cur].ptr;
        dw.a_t = true;
        dw.lda = nl;
        dw.B = in;
        dw.b_t = false;
        dw.ldb = kp[L];
        dw.M = nl;
        dw.N = kp[L];
        dw.K = rows;
        dw.epi = MLP_EPI_STORE;
        dw.out_f32 = g.as<float>() + w_off[L];
        dw.ldc = kp[L];
        EMFD_HIP(launch_mlp_gemm(dw, stream));
        EMFD_HIP(launch_mlp_colsum(gf[cur].as<float>(), rows, nl, g.as<float>() + b_off[L], stream));
        if (l == 0) break;
        // G_{l-1} = (G W_l) * gelu'(z_{l-1})
        MlpGemm dx{};
        dx.A = gb[cur].ptr;
        dx.a_t = false;
        dx.lda = nl;
        dx.B = static_cast<const uint16_t *>(wb.ptr) + wb_off[L];
        dx.b_t = false;
        dx.ldb = kp[L];
        dx.M = rows;
        dx.N = dims[L];
        dx.K = nl;
        dx.epi = MLP_EPI_GELU_GRAD;
        dx.aux = z[L - 1].as<float>();
        dx.out_f32 = gf[cur ^ 1].as<float>();
        dx.out_bf16 = gb[cur ^ 1].ptr;
        dx.ldc = dims[L];
        EMFD_HIP(launch_mlp_gemm(dx, stream));
        cur ^= 1;
    }
    return EMFD_OK;
}

namespace {
int check_mlp(emfd_mlp *m) { return m ? EMFD_OK : fail(EMFD_ERR_INVALID, "null mlp"); }

int adam_step(emfd_mlp *m, float lr) {
    ++m->adam_t;
    const float c1 = 1.0f - std::pow(m->b1, static_cast<float>(m->adam_t));
    const float c2 = 1.0f - std::pow(m->b2, static_cast<float>(m->adam_t));
    EMFD_HIP(launch_mlp_adam(m->p.as<float>(), m->g.as<float>(), m->m1.as<float>(), m->m2.as<float>(), m->n_params,
                             lr, m->b1, m->b2, m->eps, m->wd, c1, c2, m->bf_index.as<int>(), m->wb.ptr, m->stream));
    return EMFD_OK;
}

}  // namespace

extern "C" {

int emfd_mlp_create(int device, const int *dims, int n_dims, int max_batch, emfd_mlp **out) {
    if (!out || !dims) return fail(EMFD_ERR_INVALID, "null argument");
    *out = nullptr;
    if (n_dims < 2 || n_dims > 16) return fail(EMFD_ERR_INVALID, "an MLP needs 2..16 widths");
    for (int i = 0; i < n_dims; ++i)
        if (dims[i] < 1 || dims[i] > 8192) return fail(EMFD_ERR_INVALID, "layer width out of range");
    if (max_batch < 1 || max_batch > (1 << 20)) return fail(EMFD_ERR_INVALID, "max_batch out of range");
    int count = 0;
    EMFD_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(EMFD_ERR_INVALID, "no such HIP device");
    EMFD_HIP(hipSetDevice(device));
    std::unique_ptr<emfd_mlp> m(new emfd_mlp());
    m->device = device;
    EMFD_HIP(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking));
    m->dims.assign(dims, dims + n_dims);
    m->layers = n_dims - 1;
    m->max_batch = max_batch;
    for (int l = 0; l < m->layers; ++l) {
        // hidden widths are the next layer's inputs: keep them multiples of 32 as they are
        m->kp.push_back(l == 0 ? round32(dims[0]) : dims[l]);
        if (l > 0 && dims[l] % 32 != 0) return fail(EMFD_ERR_INVALID, "hidden widths must be multiples of 32");
    }
    size_t off = 0, woff = 0;
    for (int l = 0; l < m->layers; ++l) {
        const size_t nw = static_cast<size_t>(dims[l + 1]) * m->kp[static_cast<size_t>(l)];
        m->w_off.push_back(off);
        m->wb_off.push_back(woff);
        off += nw;
        woff += nw;
        m->b_off.push_back(off);
        off += static_cast<size_t>(dims[l + 1]);
        m->n_logical += static_cast<size_t>(dims[l + 1]) * dims[l] + dims[l + 1];
    }
    m->n_params = off;
    m->n_wb = woff;
    {
        std::vector<int> nout(m->dims.begin() + 1, m->dims.end());
        m->fused_ok = mlp_fused_supported(m->layers, m->kp.data(), nout.data()) != 0;
    }
    std::vector<int> bfi(m->n_params, -1);
    for (int l = 0; l < m->layers; ++l)
        for (size_t i = 0; i < static_cast<size_t>(dims[l + 1]) * m->kp[static_cast<size_t>(l)]; ++i)
            bfi[m->w_off[static_cast<size_t>(l)] + i] = static_cast<int>(m->wb_off[static_cast<size_t>(l)] + i);
    for (DevBuf *d : {&m->p, &m->g, &m->m1, &m->m2})
        if (int rc = d->ensure(m->n_params * sizeof(float))) return rc;
    if (int rc = m->bf_index.ensure(m->n_params * sizeof(int))) return rc;
    if (int rc = m->wb.ensure(m->n_wb * 2)) return rc;
    EMFD_HIP(hipMemcpy(m->bf_index.ptr, bfi.data(), m->n_params * sizeof(int), hipMemcpyHostToDevice));
    for (DevBuf *d : {&m->p, &m->g, &m->m1, &m->m2}) EMFD_HIP(hipMemset(d->ptr, 0, m->n_params * sizeof(float)));
    EMFD_HIP(hipMemset(m->wb.ptr, 0, m->n_wb * 2));
    const size_t B = static_cast<size_t>(max_batch);
    if (int rc = m->xb.ensure(B * m->kp[0] * 2)) return rc;
    if (int rc = m->tb.ensure(B * m->n_out() * sizeof(float))) return rc;
    if (int rc = m->y.ensure(B * m->n_out() * sizeof(float))) return rc;
    if (int rc = m->idx.ensure(B * sizeof(int))) return rc;
    m->z.resize(static_cast<size_t>(m->layers - 1));
    m->h.resize(static_cast<size_t>(m->layers - 1));
    for (int l = 0; l + 1 < m->layers; ++l) {
        if (int rc = m->z[static_cast<size_t>(l)].ensure(B * dims[l + 1] * sizeof(float))) return rc;
        if (int rc = m->h[static_cast<size_t>(l)].ensure(B * dims[l + 1] * 2)) return rc;
    }
    for (int c = 0; c < 2; ++c) {
        if (int rc = m->gf[c].ensure(B * m->widest() * sizeof(float))) return rc;
        if (int rc = m->gb[c].ensure(B * m->widest() * 2)) return rc;
    }
    if (int rc = m->partials.ensure(kLossBlocks * sizeof(double))) return rc;
    *out = m.release();
    return EMFD_OK;
}

int emfd_mlp_destroy(emfd_mlp *m) {
    if (!m) return EMFD_OK;
    (void)hipSetDevice(m->device);
    if (m->stream) (void)hipStreamSynchronize(m->stream);
    m->release_all();
    if (m->stream) (void)hipStreamDestroy(m->stream);
    delete m;
    return EMFD_OK;
}

int emfd_mlp_param_count(emfd_mlp *m, int64_t *n) {
    if (int rc = check_mlp(m)) return rc;
    if (!n) return fail(EMFD_ERR_INVALID, "null output");
    *n = static_cast<int64_t>(m->n_logical);
    return EMFD_OK;
}

int emfd_mlp_set_params(emfd_mlp *m, const float *params, int64_t n) {
    if (int rc = check_mlp(m)) return rc;
    if (!params || n != static_cast<int64_t>(m->n_logical)) return fail(EMFD_ERR_INVALID, "parameter count mismatch");
    std::vector<float> dev(m->n_params, 0.0f);
    size_t src = 0;
    for (int l = 0; l < m->layers; ++l) {
        const size_t L = static_cast<size_t>(l);
        const int no = m->dims[L + 1], ni = m->dims[L], k = m->kp[L];
        for (int o = 0; o < no; ++o)
            for (int i = 0; i < ni; ++i) dev[m->w_off[L] + static_cast<size_t>(o) * k + i] = params[src++];
        for (int o = 0; o < no; ++o) dev[m->b_off[L] + o] = params[src++];
    }
    EMFD_HIP(hipSetDevice(m->device));
    EMFD_HIP(hipStreamSynchronize(m->stream));
    EMFD_HIP(hipMemcpy(m->p.ptr, dev.data(), m->n_params * sizeof(float), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemset(m->m1.ptr, 0, m->n_params * sizeof(float)));
    EMFD_HIP(hipMemset(m->m2.ptr, 0, m->n_params * sizeof(float)));
    EMFD_HIP(launch_mlp_to_bf16(m->p.as<float>(), m->bf_index.as<int>(), m->n_params, m->wb.ptr, m->stream));
    EMFD_HIP(hipStreamSynchronize(m->stream));
    m->adam_t = 0;
    return EMFD_OK;
}

static void to_logical(const emfd_mlp *m, const std::vector<float> &dev, float *out) {
    size_t dst = 0;
    for (int l = 0; l < m->layers; ++l) {
        const size_t L = static_cast<size_t>(l);
        const int no = m->dims[L + 1], ni = m->dims[L], k = m->kp[L];
        for (int o = 0; o < no; ++o)
            for (int i = 0; i < ni; ++i) out[dst++] = dev[m->w_off[L] + static_cast<size_t>(o) * k + i];
        for (int o = 0; o < no; ++o) out[dst++] = dev[m->b_off[L] + o];
    }
}

int emfd_mlp_get_params(emfd_mlp *m, float *params, int64_t n) {
    if (int rc = check_mlp(m)) return rc;
    if (!params || n != static_cast<int64_t>(m->n_logical)) return fail(EMFD_ERR_INVALID, "parameter count mismatch");
    std::vector<float> dev(m->n_params);
    EMFD_HIP(hipSetDevice(m->device));
    EMFD_HIP(hipStreamSynchronize(m->stream));
    EMFD_HIP(hipMemcpy(dev.data(), m->p.ptr, m->n_params * sizeof(float), hipMemcpyDeviceToHost));
    to_logical(m, dev, params);
    return EMFD_OK;
}

// Copies host rows to the device staging buffers (data_x/data_t are not touched) and gathers
// them into xb / tb.
static int stage_rows(emfd_mlp *m, const float *x, const float *t, int rows, DevBuf &sx, DevBuf &st) {
    const size_t xin = static_cast<size_t>(rows) * m->n_in(), tin = static_cast<size_t>(rows) * m->n_out();
    if (int rc = sx.ensure(static_cast<size_t>(m->max_batch) * m->n_in() * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpyAsync(sx.ptr, x, xin * sizeof(float), hipMemcpyHostToDevice, m->stream));
    if (t) {
        if (int rc = st.ensure(static_cast<size_t>(m->max_batch) * m->n_out() * sizeof(float))) return rc;
        EMFD_HIP(hipMemcpyAsync(st.ptr, t, tin * sizeof(float), hipMemcpyHostToDevice, m->stream));
    }
    EMFD_HIP(launch_mlp_gather(sx.as<float>(), m->n_in(), m->kp[0], t ? st.as<float>() : nullptr, m->n_out(), nullptr,
                               rows, m->xb.ptr, m->tb.as<float>(), m->stream));
    return EMFD_OK;
}

namespace {
struct Staging {
    DevBuf x, t;
    ~Staging() {
        x.release();
        t.release();
    }
};
}  // namespace

int emfd_mlp_forward(emfd_mlp *m, const float *x, int64_t batch, float *y) {
    if (int rc = check_mlp(m)) return rc;
    if (!x || !y || batch < 0) return fail(EMFD_ERR_INVALID, "bad forward arguments");
    EMFD_HIP(hipSetDevice(m->device));
    Staging s;
    if (int rc = s.x.ensure(static_cast<size_t>(m->max_batch) * m->n_in() * sizeof(float))) return rc;
    for (int64_t r0 = 0; r0 < batch; r0 += m->max_batch) {
        const int rows = static_cast<int>(std::min<int64_t>(m->max_batch, batch - r0));
        EMFD_HIP(hipMemcpyAsync(s.x.ptr, x + r0 * m->n_in(), static_cast<size_t>(rows) * m->n_in() * sizeof(float),
                                hipMemcpyHostToDevice, m->stream));
        if (int rc = m->forward_infer(s.x.as<float>(), rows)) return rc;
        EMFD_HIP(hipMemcpyAsync(y + r0 * m->n_out(), m->y.ptr, static_cast<size_t>(rows) * m->n_out() * sizeof(float),
                                hipMemcpyDeviceToHost, m->stream));
        EMFD_HIP(hipStreamSynchronize(m->stream));
    }
    return EMFD_OK;
}

static int loss_of(emfd_mlp *m, int rows, double *loss) {
    double part[kLossBlocks];
    EMFD_HIP(hipMemcpyAsync(part, m->partials.ptr, sizeof(part), hipMemcpyDeviceToHost, m->stream));
    EMFD_HIP(hipStreamSynchronize(m->stream));
    double s = 0.0;
    for (double v : part) s += v;
    if (loss) *loss = s / (static_cast<double>(rows) * m->n_out());
    return EMFD_OK;
}

int emfd_mlp_gradients(emfd_mlp *m, const float *x, const float *t, int64_t batch, float *grads, double *loss) {
    if (int rc = check_mlp(m)) return rc;
    if (!x || !t || batch < 1 || batch > m->max_batch) return fail(EMFD_ERR_INVALID, "bad gradient arguments");
    EMFD_HIP(hipSetDevice(m->device));
    Staging s;
    const int rows = static_cast<int>(batch);
    if (int rc = stage_rows(m, x, t, rows, s.x, s.t)) return rc;
    if (int rc = m->forward(rows)) return rc;
    if (int rc = m->backward(rows, m->partials.as<double>(), kLossBlocks)) return rc;
    if (int rc = loss_of(m, rows, loss)) return rc;
    if (grads) {
        std::vector<float> dev(m->n_params);
        EMFD_HIP(hipMemcpy(dev.data(), m->g.ptr, m->n_params * sizeof(float), hipMemcpyDeviceToHost));
        to_logical(m, dev, grads);
    }
    return EMFD_OK;
}

int emfd_mlp_set_adam(emfd_mlp *m, float beta1, float beta2, float eps, float weight_decay) {
    if (int rc = check_mlp(m)) return rc;
    if (!(beta1 >= 0.0f && beta1 < 1.0f && beta2 >= 0.0f && beta2 < 1.0f && eps > 0.0f && weight_decay >= 0.0f))
        return fail(EMFD_ERR_INVALID, "Adam settings out of range");
    m->b1 = beta1;
    m->b2 = beta2;
    m->eps = eps;
    m->wd = weight_decay;
    return EMFD_OK;
}

int emfd_mlp_train_step(emfd_mlp *m, const float *x, const float *t, int64_t batch, float lr, double *loss) {
    if (int rc = check_mlp(m)) return rc;
    if (!x || !t || batch < 1 || batch > m->max_batch) return fail(EMFD_ERR_INVALID, "bad training arguments");
    EMFD_HIP(hipSetDevice(m->device));
    Staging s;
    const int rows = static_cast<int>(batch);
    if (int rc = stage_rows(m, x, t, rows, s.x, s.t)) return rc;
    if (int rc = m->forward(rows)) return rc;
    if (int rc = m->backward(rows, m->partials.as<double>(), kLossBlocks)) return rc;
    if (int rc = adam_step(m, lr)) return rc;
    return loss_of(m, rows, loss);
}

int emfd_mlp_set_dataset(emfd_mlp *m, const float *x, const float *t, int64_t n) {
    if (int rc = check_mlp(m)) return rc;
    if (!x || !t || n < 1 || n > (int64_t(1) << 28)) return fail(EMFD_ERR_INVALID, "bad dataset");
    EMFD_HIP(hipSetDevice(m->device));
    EMFD_HIP(hipStreamSynchronize(m->stream));
    if (int rc = m->data_x.ensure(static_cast<size_t>(n) * m->n_in() * sizeof(float))) return rc;
    if (int rc = m->data_t.ensure(static_cast<size_t>(n) * m->n_out() * sizeof(float))) return rc;
    EMFD_HIP(hipMemcpy(m->data_x.ptr, x, static_cast<size_t>(n) * m->n_in() * sizeof(float), hipMemcpyHostToDevice));
    EMFD_HIP(hipMemcpy(m->data_t.ptr, t, static_cast<size_t>(n) * m->n_out() * sizeof(float), hipMemcpyHostToDevice));
    m->data_n = n;
    return EMFD_OK;
}

int emfd_mlp_train(emfd_mlp *m, int batch, int epochs, float lr_start, float lr_end, uint64_t seed,
                   double *epoch_loss, int cap) {
    if (int rc = check_mlp(m)) return rc;
    if (m->data_n < 1) return fail(EMFD_ERR_STATE, "no training set on the device");
    if (batch < 1 || batch > m->max_batch || epochs < 0) return fail(EMFD_ERR_INVALID, "bad training arguments");
    EMFD_HIP(hipSetDevice(m->device));
    const int64_t n = m->data_n;
    const int steps = static_cast<int>((n + batch - 1) / batch);
    DevBuf perm, parts;
    if (int rc = perm.ensure(static_cast<size_t>(n) * sizeof(int))) return rc;
    if (int rc = parts.ensure(static_cast<size_t>(steps) * kLossBlocks * sizeof(double))) {
        perm.release();
        return rc;
    }
    std::vector<int> order(static_cast<size_t>(n));
    std::vector<double> part(static_cast<size_t>(steps) * kLossBlocks);
    uint64_t st = seed;
    const int total = std::max(epochs * steps, 1);
    int done = 0, rc = EMFD_OK;
    for (int e = 0; e < epochs && rc == EMFD_OK; ++e) {
        for (int64_t i = 0; i < n; ++i) order[static_cast<size_t>(i)] = static_cast<int>(i);
        for (int64_t i = n - 1; i > 0; --i) {
            const int64_t j = static_cast<int64_t>(splitmix(st) % static_cast<uint64_t>(i + 1));
            std::swap(order[static_cast<size_t>(i)], order[static_cast<size_t>(j)]);
        }
        if (hipMemcpyAsync(perm.ptr, order.data(), order.size() * sizeof(int), hipMemcpyHostToDevice, m->stream) !=
            hipSuccess) {
            rc = fail(EMFD_ERR_HIP, "permutation upload failed");
            break;
        }
        for (int s = 0; s < steps && rc == EMFD_OK; ++s) {
            const int r0 = s * batch;
            const int rows = static_cast<int>(std::min<int64_t>(batch, n - r0));
            const float frac = static_cast<float>(done) / static_cast<float>(total);
            const float lr = lr_end + 0.5f * (lr_start - lr_end) * (1.0f + std::cos(3.14159265358979f * frac));
            hipError_t he = launch_mlp_gather(m->data_x.as<float>(), m->n_in(), m->kp[0], m->data_t.as<float>(),
                                              m->n_out(), perm.as<int>() + r0, rows, m->xb.ptr, m->tb.as<float>(),
                                              m->stream);
            if (he != hipSuccess) {
                rc = hip_fail(he, "launch_mlp_gather");
                break;
            }
            if ((rc = m->forward(rows))) break;
            if ((rc = m->backward(rows, parts.as<double>() + static_cast<size_t>(s) * kLossBlocks, kLossBlocks))) break;
            if ((rc = adam_step(m, lr))) break;
            ++done;
        }
        if (rc) break;
        if (hipMemcpyAsync(part.data(), parts.ptr, part.size() * sizeof(double), hipMemcpyDeviceToHost, m->stream) !=
                hipSuccess ||
            hipStreamSynchronize(m->stream) != hipSuccess) {
            rc = fail(EMFD_ERR_HIP, "loss readback failed");
            break;
        }
        if (epoch_loss && e < cap) {
            double sum = 0.0;
            for (int s = 0; s < steps; ++s) {
                const int rows = static_cast<int>(std::min<int64_t>(batch, n - static_cast<int64_t>(s) * batch));
                double v = 0.0;
                for (int b = 0; b < kLossBlocks; ++b) v += part[static_cast<size_t>(s) * kLossBlocks + b];
                sum += v / (static_cast<double>(rows) * m->n_out());
            }
            epoch_loss[e] = sum / steps;
        }
    }
    (void)hipStreamSynchronize(m->stream);
    perm.release();
    parts.release();
    return rc;
}

int emfd_mlp_time_forward(emfd_mlp *m, int64_t batch, int iterations, double *avg_ms) {
    if (int rc = check_mlp(m)) return rc;
    if (batch < 1 || batch > m->max_batch || iterations < 1 || !avg_ms) return fail(EMFD_ERR_INVALID, "bad arguments");
    EMFD_HIP(hipSetDevice(m->device));
    const int rows = static_cast<int>(batch);
    hipEvent_t a, b;
    EMFD_HIP(hipEventCreate(&a));
    EMFD_HIP(hipEventCreate(&b));
    Staging st;  // device input rows (zeros): timing does not depend on the values
    if (int rc = st.x.ensure(static_cast<size_t>(rows) * m->n_in() * sizeof(float))) return rc;
    EMFD_HIP(hipMemsetAsync(st.x.ptr, 0, static_cast<size_t>(rows) * m->n_in() * sizeof(float), m->stream));
    if (int rc = m->forward_infer(st.x.as<float>(), rows)) return rc;  // warm-up
    EMFD_HIP(hipEventRecord(a, m->stream));
    for (int i = 0; i < iterations; ++i)
        if (int rc = m->forward_infer(st.x.as<float>(), rows)) return rc;
    EMFD_HIP(hipEventRecord(b, m->stream));
    EMFD_HIP(hipEventSynchronize(b));
    float ms = 0.0f;
    EMFD_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    *avg_ms = static_cast<double>(ms) / iterations;
    return EMFD_OK;
}

}  // extern "C"
