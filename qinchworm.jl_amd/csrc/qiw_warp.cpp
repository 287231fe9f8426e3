// qiw_warp.cpp -- the learned warp of the unit cube: parameters, the batched forward,
// backward and inverse passes over the flow kernels (qiw_flow.hip), Adam training on
// |integrand| samples of the plain method, and the warped sample points of bold steps.
//
// Direction conventions: the training direction maps a plain sample point z to x and
// accumulates log q(z) = sum of the layers' log dy/dx, where q is the density of the
// warped points; sampling runs the layers backwards with the inverse splines, x -> z, and
// accumulates log |dz/dx| = -log q(z).  The loss of a batch is -mean_i coef_i log q(z_i)
// with coef_i = |f(z_i)| / mean |f| of its (step, group) set.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "qiw_ctx.h"
#include "qiw_flow.h"

namespace qiwi {

struct WarpState {
    qiw_warp_config cfg{};
    int L = 0, K = 0, H = 0, P = 0, n_out = 0;
    int64_t n_params = 0;
    std::vector<int64_t> off;  // per layer: W1, b1, W2, b2, W3, b3 offsets; then the embedding
    int64_t emb_off = 0;
    float *params = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
    int64_t adam_t = 0;
    bool enabled = false;
    // groups prepared by warp_prepare for the current run
    std::vector<WarpGroup> groups;
    int64_t n = 0;
};

void warp_destroy(WarpState *w) {
    if (!w) return;
    for (float *p : {w->params, w->grad, w->m1, w->m2})
        if (p) (void)hipFree(p);
    delete w;
}

bool warp_enabled(const WarpState *w) { return w && w->enabled; }

namespace {

enum { W1 = 0, B1, W2, B2, W3, B3 };

// xorshift64* stream for the initialisation and the minibatch order
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed ? seed : 0x9E3779B97F4A7C15ull) {}
    uint64_t next() {
        s ^= s >> 12;
        s ^= s << 25;
        s ^= s >> 27;
        return s * 2685821657736338717ull;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

// Device work buffers of a batch of n rows (named context buffers, grown on demand).
struct Work {
    double *x = nullptr;      // [L+1][n][8]
    float *feat = nullptr;    // [L][n][FEAT]
    float *h1 = nullptr, *h2 = nullptr;  // [L][n][H]
    float *raw = nullptr;     // [L][n][n_out]
    double *logq = nullptr;   // [n]
    double *gx = nullptr, *gx2 = nullptr;  // [n][8]
    float *graw = nullptr;    // [n][n_out]
    float *dh = nullptr, *dpre = nullptr;  // [n][H]
    float *dfeat = nullptr;   // [n][FEAT]
    float *part = nullptr;    // weight-gradient partial sums
    int32_t *order = nullptr, *group = nullptr;
    float *step = nullptr, *coef = nullptr;
    double *loss = nullptr;
};

int work(qiw_ctx *c, const WarpState &w, int64_t n, bool training, Work &k) {
    int rc = 0;
    const int64_t L = w.L;
    const int64_t chunks = (n + QIW_FLOW_CHUNK - 1) / QIW_FLOW_CHUNK;
    const int64_t xs = training ? L + 1 : 2;  // inference ping-pongs two coordinate buffers
    const int64_t ls = training ? L : 1;      // and keeps one layer of activations
    k.x = (double *)dev(c, "warp.x", (size_t)(xs * n * QIW_FLOW_D) * 8, &rc);
    k.feat = (float *)dev(c, "warp.feat", (size_t)(ls * n * QIW_FLOW_FEAT) * 4, &rc);
    k.h1 = (float *)dev(c, "warp.h1", (size_t)(ls * n * w.H) * 4, &rc);
    k.h2 = (float *)dev(c, "warp.h2", (size_t)(ls * n * w.H) * 4, &rc);
    k.raw = (float *)dev(c, "warp.raw", (size_t)(ls * n * w.n_out) * 4, &rc);
    k.logq = (double *)dev(c, "warp.logq", (size_t)n * 8, &rc);
    k.order = (int32_t *)dev(c, "warp.order", (size_t)n * 4, &rc);
    k.group = (int32_t *)dev(c, "warp.group", (size_t)n * 4, &rc);
    k.step = (float *)dev(c, "warp.step", (size_t)n * 4, &rc);
    if (training) {
        k.gx = (double *)dev(c, "warp.gx", (size_t)(n * QIW_FLOW_D) * 8, &rc);
        k.gx2 = (double *)dev(c, "warp.gx2", (size_t)(n * QIW_FLOW_D) * 8, &rc);
        k.graw = (float *)dev(c, "warp.graw", (size_t)(n * w.n_out) * 4, &rc);
        k.dh = (float *)dev(c, "warp.dh", (size_t)(n * w.H) * 4, &rc);
        k.dpre = (float *)dev(c, "warp.dpre", (size_t)(n * std::max(w.H, w.n_out)) * 4, &rc);
        k.dfeat = (float *)dev(c, "warp.dfeat", (size_t)(n * QIW_FLOW_FEAT) * 4, &rc);
        const int64_t ne = (int64_t)(std::max(w.H, QIW_FLOW_FEAT) + 1) * std::max(w.H, w.n_out);
        k.part = (float *)dev(c, "warp.part",
                              (size_t)(chunks * std::max<int64_t>(ne, QIW_FLOW_GROUPS * QIW_FLOW_EMBED)) * 4, &rc);
        k.coef = (float *)dev(c, "warp.coef", (size_t)n * 4, &rc);
        k.loss = (double *)dev(c, "warp.loss", 8, &rc);
    }
    return rc;
}

// Training direction over n rows starting from k.x[0]: every layer's intermediates are
// kept (training) or only the last (inference); log q accumulates in k.logq.
int forward(qiw_ctx *c, const WarpState &w, int n, bool training, Work &k) {
    hipStream_t st = c->stream;
    const float *p = w.params;
    HIPCHK(c, qiw_flow_fill_f64(st, k.logq, n, 0.0));
    for (int l = 0; l < w.L; ++l) {
        const int64_t li = training ? l : 0;
        const double *xin = k.x + (training ? l : (l & 1)) * (int64_t)n * QIW_FLOW_D;
        double *xout = k.x + (training ? l + 1 : ((l + 1) & 1)) * (int64_t)n * QIW_FLOW_D;
        float *feat = k.feat + li * n * QIW_FLOW_FEAT, *h1 = k.h1 + li * n * w.H, *h2 = k.h2 + li * n * w.H;
        float *raw = k.raw + li * n * w.n_out;
        const int64_t *o = &w.off[6 * l];
        HIPCHK(c, qiw_flow_features(st, l, xin, k.order, k.group, k.step, p + w.emb_off, feat, n));
        HIPCHK(c, qiw_flow_dense(st, feat, p + o[W1], p + o[B1], h1, n, QIW_FLOW_FEAT, w.H, 1));
        HIPCHK(c, qiw_flow_dense(st, h1, p + o[W2], p + o[B2], h2, n, w.H, w.H, 1));
        HIPCHK(c, qiw_flow_dense(st, h2, p + o[W3], p + o[B3], raw, n, w.H, w.n_out, 0));
        HIPCHK(c, qiw_flow_rqs(st, l, xin, xout, raw, k.order, k.logq, n, w.K, 0));
    }
    return QIW_OK;
}

// Sampling direction: layers L-1..0 with the inverse splines, from k.x[0]; the result is
// in buffer (L & 1) of k.x, log |dz/dx| in k.logq.
int inverse(qiw_ctx *c, const WarpState &w, int n, Work &k) {
    hipStream_t st = c->stream;
    const float *p = w.params;
    HIPCHK(c, qiw_flow_fill_f64(st, k.logq, n, 0.0));
    for (int q = 0; q < w.L; ++q) {
        const int l = w.L - 1 - q;
        const double *xin = k.x + (int64_t)(q & 1) * n * QIW_FLOW_D;
        double *xout = k.x + (int64_t)((q + 1) & 1) * n * QIW_FLOW_D;
        const int64_t *o = &w.off[6 * l];
        HIPCHK(c, qiw_flow_features(st, l, xin, k.order, k.group, k.step, p + w.emb_off, k.feat, n));
        HIPCHK(c, qiw_flow_dense(st, k.feat, p + o[W1], p + o[B1], k.h1, n, QIW_FLOW_FEAT, w.H, 1));
        HIPCHK(c, qiw_flow_dense(st, k.h1, p + o[W2], p + o[B2], k.h2, n, w.H, w.H, 1));
        HIPCHK(c, qiw_flow_dense(st, k.h2, p + o[W3], p + o[B3], k.raw, n, w.H, w.n_out, 0));
        HIPCHK(c, qiw_flow_rqs(st, l, xin, xout, k.raw, k.order, k.logq, n, w.K, 1));
    }
    return QIW_OK;
}

// Loss and gradient (into w.grad, which is cleared first) of the batch in k after forward().
int backward(qiw_ctx *c, WarpState &w, int n, Work &k) {
    hipStream_t st = c->stream;
    const float *p = w.params;
    float *g = w.grad;
    const double scale = 1.0 / n;
    HIPCHK(c, qiw_flow_fill(st, g, w.n_params, 0.0f));
    HIPCHK(c, qiw_flow_loss(st, k.logq, k.coef, scale, k.loss, n));
    HIPCHK(c, qiw_flow_fill_f64(st, k.gx, (int64_t)n * QIW_FLOW_D, 0.0));
    for (int l = w.L - 1; l >= 0; --l) {
        const double *xin = k.x + (int64_t)l * n * QIW_FLOW_D;
        const float *feat = k.feat + (int64_t)l * n * QIW_FLOW_FEAT, *h1 = k.h1 + (int64_t)l * n * w.H;
        const float *h2 = k.h2 + (int64_t)l * n * w.H, *raw = k.raw + (int64_t)l * n * w.n_out;
        const int64_t *o = &w.off[6 * l];
        HIPCHK(c, qiw_flow_rqs_bwd(st, l, xin, raw, k.order, k.coef, scale, k.gx, k.gx2, k.graw, n, w.K));
        // output layer (no activation)
        HIPCHK(c, qiw_flow_dense_bwd_weight(st, h2, k.graw, k.part, g + o[W3], g + o[B3], n, w.H, w.n_out));
        HIPCHK(c, qiw_flow_dense_bwd_data(st, k.graw, p + o[W3], k.dh, n, w.H, w.n_out));
        // second hidden layer
        HIPCHK(c, qiw_flow_tanh_bwd(st, k.dh, h2, k.dpre, (int64_t)n * w.H));
        HIPCHK(c, qiw_flow_dense_bwd_weight(st, h1, k.dpre, k.part, g + o[W2], g + o[B2], n, w.H, w.H));
        HIPCHK(c, qiw_flow_dense_bwd_data(st, k.dpre, p + o[W2], k.dh, n, w.H, w.H));
        // first hidden layer
        HIPCHK(c, qiw_flow_tanh_bwd(st, k.dh, h1, k.dpre, (int64_t)n * w.H));
        HIPCHK(c, qiw_flow_dense_bwd_weight(st, feat, k.dpre, k.part, g + o[W1], g + o[B1], n, QIW_FLOW_FEAT,
                                            w.H));
        HIPCHK(c, qiw_flow_dense_bwd_data(st, k.dpre, p + o[W1], k.dfeat, n, QIW_FLOW_FEAT, w.H));
        HIPCHK(c, qiw_flow_feat_bwd(st, l, k.dfeat, k.order, k.group, k.gx2, k.part, g + w.emb_off, n));
        std::swap(k.gx, k.gx2);
    }
    return QIW_OK;
}

int need_warp(qiw_ctx *c) {
    if (!c) return QIW_E_ARG;
    if (!c->warp) return fail(c, QIW_E_STATE, "no warp initialised (qiw_warp_init)");
    return QIW_OK;
}

// Copies the host batch description into the work buffers.
int stage_batch(qiw_ctx *c, Work &k, int64_t n, const double *pts, const int32_t *order, const int32_t *group,
                const double *step, const double *coef) {
    for (int64_t r = 0; r < n; ++r) {
        if (order[r] < 1 || order[r] > QIW_FLOW_KMAX || group[r] < 0 || group[r] >= QIW_FLOW_GROUPS)
            return fail(c, QIW_E_ARG, "warp order must be 1..4 and group 0..15");
    }
    std::vector<float> sf(step, step + n);
    HIPCHK(c, hipMemcpy(k.x, pts, (size_t)n * QIW_FLOW_D * 8, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k.order, order, (size_t)n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k.group, group, (size_t)n * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k.step, sf.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    if (coef) {
        std::vector<float> cf(coef, coef + n);
        HIPCHK(c, hipMemcpy(k.coef, cf.data(), (size_t)n * 4, hipMemcpyHostToDevice));
    }
    return QIW_OK;
}

// Sobol coordinates of point i of the staged sequence (the closed form of next!,
// src/scrambled_sobol.jl:158-173, as computed by the device kernels)
double sobol_coord(const qiw_ctx *c, int d, uint64_t i) {
    uint64_t g = i ^ (i >> 1);
    uint32_t x = c->sob_x[d];
    for (int j = 0; j < 32; ++j)
        if ((g >> j) & 1ull) x ^= c->sob_m[(size_t)d * 32 + j];
    return std::ldexp((double)x, -32);
}

}  // namespace

// ---------------------------------------------------------------------------------------
// Driver hooks
// ---------------------------------------------------------------------------------------
int warp_prepare(qiw_ctx *c, const std::vector<WarpGroup> &gs, int64_t n) {
    WarpState &w = *c->warp;
    w.groups = gs;
    w.n = n;
    const int64_t rows = (int64_t)gs.size() * n;
    if (rows > (int64_t)1 << 31) return fail(c, QIW_E_LIMIT, "too many warped samples");
    int rc = 0;
    Work k;
    if ((rc = work(c, w, rows, false, k))) return rc;
    double *cube = (double *)dev(c, "warp.cube", (size_t)(rows * QIW_FLOW_D) * 8, &rc);
    uint32_t *sm = (uint32_t *)dev(c, "sobol.m", c->sob_m.size() * 4, &rc);
    uint32_t *sx = (uint32_t *)dev(c, "sobol.x", c->sob_x.size() * 4, &rc);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(sm, c->sob_m.data(), c->sob_m.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(sx, c->sob_x.data(), c->sob_x.size() * 4, hipMemcpyHostToDevice));
    std::vector<int32_t> ord((size_t)rows), grp((size_t)rows);
    for (size_t gi = 0; gi < gs.size(); ++gi) {
        const int dim = gs[gi].d_before + gs[gi].d_after;
        if (dim > c->sob_dim) return fail(c, QIW_E_DIM, "Sobol dimension beyond the staged direction numbers");
        HIPCHK(c, qiw_launch_sobol_cube(c->stream, sm, sx, dim, QIW_FLOW_D, n, cube + (int64_t)gi * n * QIW_FLOW_D));
        std::fill(ord.begin() + gi * n, ord.begin() + (gi + 1) * n, gs[gi].order);
        std::fill(grp.begin() + gi * n, grp.begin() + (gi + 1) * n, gs[gi].gid);
    }
    HIPCHK(c, hipMemcpy(k.order, ord.data(), (size_t)rows * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(k.group, grp.data(), (size_t)rows * 4, hipMemcpyHostToDevice));
    return QIW_OK;
}

int warp_points(qiw_ctx *c, double step_frac, double *ut) {
    WarpState &w = *c->warp;
    const int64_t n = w.n, rows = (int64_t)w.groups.size() * n;
    int rc = 0;
    Work k;
    if ((rc = work(c, w, rows, false, k))) return rc;
    double *cube = (double *)dev(c, "warp.cube", (size_t)(rows * QIW_FLOW_D) * 8, &rc);
    if (rc) return rc;
    HIPCHK(c, hipMemcpyAsync(k.x, cube, (size_t)(rows * QIW_FLOW_D) * 8, hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(c, qiw_flow_fill(c->stream, k.step, rows, (float)step_frac));
    if ((rc = inverse(c, w, (int)rows, k))) return rc;
    const double *z = k.x + (int64_t)(w.L & 1) * rows * QIW_FLOW_D;
    for (size_t gi = 0; gi < w.groups.size(); ++gi) {
        const WarpGroup &g = w.groups[gi];
        HIPCHK(c, qiw_launch_warp_coords(c->stream, z + (int64_t)gi * n * QIW_FLOW_D, QIW_FLOW_D,
                                         k.logq + (int64_t)gi * n, g.d_before, g.d_after, n, ut + g.off));
    }
    return QIW_OK;
}

}  // namespace qiwi

using namespace qiwi;

// ---------------------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------------------
static int qiw_warp_init_impl(qiw_ctx *c, const qiw_warp_config *cfg) {
    if (!c || !cfg) return QIW_E_ARG;
    if (cfg->n_layers < 1 || cfg->n_layers > 16 || cfg->n_bins < 1 || cfg->n_bins > QIW_FLOW_MAX_BINS ||
        cfg->hidden < 1 || cfg->hidden > QIW_FLOW_MAX_HIDDEN)
        return fail(c, QIW_E_LIMIT, "warp layers 1..16, bins 1..16, hidden width 1..128");
    (void)hipSetDevice(c->device);
    warp_destroy(c->warp);
    c->warp = new WarpState();
    WarpState &w = *c->warp;
    w.cfg = *cfg;
    w.L = cfg->n_layers;
    w.K = cfg->n_bins;
    w.H = cfg->hidden;
    w.P = 3 * w.K + 1;
    w.n_out = QIW_FLOW_SLOTS * w.P;
    int64_t o = 0;
    const int64_t sizes[6] = {(int64_t)QIW_FLOW_FEAT * w.H, w.H, (int64_t)w.H * w.H, w.H, (int64_t)w.H * w.n_out,
                              w.n_out};
    for (int l = 0; l < w.L; ++l)
        for (int q = 0; q < 6; ++q) {
            w.off.push_back(o);
            o += sizes[q];
        }
    w.emb_off = o;
    w.n_params = o + QIW_FLOW_GROUPS * QIW_FLOW_EMBED;
    std::vector<float> h((size_t)w.n_params, 0.0f);
    Rng rng(cfg->seed);
    auto glorot = [&](int64_t at, int fan_in, int fan_out) {
        const double a = std::sqrt(6.0 / (fan_in + fan_out));
        for (int64_t i = 0; i < (int64_t)fan_in * fan_out; ++i) h[at + i] = (float)((2.0 * rng.uniform() - 1.0) * a);
    };
    for (int l = 0; l < w.L; ++l) {
        glorot(w.off[6 * l + W1], QIW_FLOW_FEAT, w.H);
        glorot(w.off[6 * l + W2], w.H, w.H);
    }
    for (int64_t i = 0; i < QIW_FLOW_GROUPS * QIW_FLOW_EMBED; ++i)
        h[w.emb_off + i] = (float)((2.0 * rng.uniform() - 1.0) * 0.5);
    for (float **p : {&w.params, &w.grad, &w.m1, &w.m2})
        if (hipMalloc(p, (size_t)w.n_params * 4) != hipSuccess) return fail(c, QIW_E_DEVICE, "hipMalloc failed for the warp");
    HIPCHK(c, hipMemcpy(w.params, h.data(), (size_t)w.n_params * 4, hipMemcpyHostToDevice));
    HIPCHK(c, hipMemset(w.m1, 0, (size_t)w.n_params * 4));
    HIPCHK(c, hipMemset(w.m2, 0, (size_t)w.n_params * 4));
    w.adam_t = 0;
    return QIW_OK;
}

static int qiw_warp_params_impl(qiw_ctx *c, int64_t *n_out, float *get, const float *set) {
    int rc = need_warp(c);
    if (rc) return rc;
    WarpState &w = *c->warp;
    if (n_out) *n_out = w.n_params;
    if (get) HIPCHK(c, hipMemcpy(get, w.params, (size_t)w.n_params * 4, hipMemcpyDeviceToHost));
    if (set) HIPCHK(c, hipMemcpy(w.params, set, (size_t)w.n_params * 4, hipMemcpyHostToDevice));
    return QIW_OK;
}

static int qiw_warp_apply_impl(qiw_ctx *c, int32_t inv, int64_t n, const double *pts, const int32_t *order,
                              const int32_t *group, const double *step, double *out, double *logdet) {
    int rc = need_warp(c);
    if (rc) return rc;
    if (n < 1 || n > (1 << 24) || !pts || !order || !group || !step) return fail(c, QIW_E_ARG, "bad warp batch");
    WarpState &w = *c->warp;
    Work k;
    if ((rc = work(c, w, n, false, k))) return rc;
    if ((rc = stage_batch(c, k, n, pts, order, group, step, nullptr))) return rc;
    if ((rc = inv ? inverse(c, w, (int)n, k) : forward(c, w, (int)n, false, k))) return rc;
    const double *res = k.x + (int64_t)(w.L & 1) * n * QIW_FLOW_D;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (out) HIPCHK(c, hipMemcpy(out, res, (size_t)n * QIW_FLOW_D * 8, hipMemcpyDeviceToHost));
    if (logdet) HIPCHK(c, hipMemcpy(logdet, k.logq, (size_t)n * 8, hipMemcpyDeviceToHost));
    return QIW_OK;
}

static int qiw_warp_loss_grad_impl(qiw_ctx *c, int64_t n, const double *z, const int32_t *order,
                                  const int32_t *group, const double *step, const double *coef, double *loss,
                                  float *grad) {
    int rc = need_warp(c);
    if (rc) return rc;
    if (n < 1 || n > (1 << 24) || !z || !order || !group || !step || !coef) return fail(c, QIW_E_ARG, "bad warp batch");
    WarpState &w = *c->warp;
    Work k;
    if ((rc = work(c, w, n, true, k))) return rc;
    if ((rc = stage_batch(c, k, n, z, order, group, step, coef))) return rc;
    if ((rc = forward(c, w, (int)n, true, k))) return rc;
    if ((rc = backward(c, w, (int)n, k))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss) HIPCHK(c, hipMemcpy(loss, k.loss, 8, hipMemcpyDeviceToHost));
    if (grad) HIPCHK(c, hipMemcpy(grad, w.grad, (size_t)w.n_params * 4, hipMemcpyDeviceToHost));
    return QIW_OK;
}

static int qiw_warp_train_impl(qiw_ctx *c, const qiw_run *run, double *loss_history) {
    int rc = need_warp(c);
    if (rc) return rc;
    if (!run) return fail(c, QIW_E_ARG, "null run");
    if (!c->has_model) return fail(c, QIW_E_STATE, "no model staged (qiw_set_model)");
    WarpState &w = *c->warp;
    const qiw_warp_config &cfg = w.cfg;
    const int64_t ntr = cfg.n_train;
    if (ntr < 1 || (ntr & (ntr - 1)) || cfg.batch < 1 || cfg.iters < 0 || cfg.n_train_steps < 1)
        return fail(c, QIW_E_ARG, "warp training: n_train must be 2^n, batch and n_train_steps positive");
    const HostModel &m = c->m;
    // training steps spread over the bold steps 1 .. n_tau-2
    std::vector<int> steps;
    const int last = m.ntau - 2;
    if (last < 1) return fail(c, QIW_E_ARG, "grid too short for bold steps");
    for (int i = 0; i < cfg.n_train_steps; ++i) {
        const int jw = (cfg.n_train_steps == 1) ? (1 + last) / 2
                                                : 1 + (int)std::lround((double)i * (last - 1) / (cfg.n_train_steps - 1));
        if (steps.empty() || steps.back() != jw) steps.push_back(jw);
    }
    std::vector<WarpGroup> gs;
    std::vector<double> mag;
    const bool was = w.enabled;
    w.enabled = false;  // training data from the plain method
    rc = warp_collect(c, run, steps, ntr, gs, mag);
    w.enabled = was;
    if (rc) return rc;
    // the training set: plain points, conditions and normalised |f| weights
    const int64_t M = (int64_t)steps.size() * gs.size() * ntr;
    std::vector<double> z((size_t)M * QIW_FLOW_D, 0.0);
    std::vector<int32_t> ord((size_t)M), grp((size_t)M);
    std::vector<double> stp((size_t)M), cf((size_t)M);
    int64_t r = 0;
    for (size_t si = 0; si < steps.size(); ++si)
        for (size_t gi = 0; gi < gs.size(); ++gi) {
            const double *mg = &mag[(si * gs.size() + gi) * (size_t)ntr];
            double mean = 0.0;
            for (int64_t i = 0; i < ntr; ++i) mean += mg[i];
            mean /= (double)ntr;
            const int dim = gs[gi].d_before + gs[gi].d_after;
            for (int64_t i = 0; i < ntr; ++i, ++r) {
                for (int d = 0; d < dim; ++d) z[(size_t)r * QIW_FLOW_D + d] = sobol_coord(c, d, (uint64_t)i);
                ord[r] = gs[gi].order;
                grp[r] = gs[gi].gid;
                stp[r] = steps[si] * m.h / m.beta;
                cf[r] = mean > 0.0 ? mg[i] / mean : 1.0;
            }
        }
    // shuffled once; iteration t takes the next `batch` rows cyclically
    Rng rng(cfg.seed ^ 0xA5A5A5A5ull);
    std::vector<int64_t> perm((size_t)M);
    for (int64_t i = 0; i < M; ++i) perm[i] = i;
    for (int64_t i = M - 1; i > 0; --i) std::swap(perm[i], perm[(int64_t)(rng.next() % (uint64_t)(i + 1))]);
    const int B = (int)std::min<int64_t>(cfg.batch, M);
    const int64_t nb = M / B;
    std::vector<double> zb((size_t)B * QIW_FLOW_D), sb((size_t)B), cb((size_t)B);
    std::vector<int32_t> ob((size_t)B), gb((size_t)B);
    double *dev_z = (double *)dev(c, "warp.train.z", (size_t)(nb * B * QIW_FLOW_D) * 8, &rc);
    int32_t *dev_o = (int32_t *)dev(c, "warp.train.o", (size_t)(nb * B) * 4, &rc);
    int32_t *dev_g = (int32_t *)dev(c, "warp.train.g", (size_t)(nb * B) * 4, &rc);
    float *dev_s = (float *)dev(c, "warp.train.s", (size_t)(nb * B) * 4, &rc);
    float *dev_c = (float *)dev(c, "warp.train.c", (size_t)(nb * B) * 4, &rc);
    double *dev_l = (double *)dev(c, "warp.train.loss", (size_t)std::max(cfg.iters, 1) * 8, &rc);
    if (rc) return rc;
    {
        std::vector<double> zz((size_t)(nb * B * QIW_FLOW_D));
        std::vector<int32_t> oo((size_t)(nb * B)), gg((size_t)(nb * B));
        std::vector<float> ss((size_t)(nb * B)), cc((size_t)(nb * B));
        for (int64_t i = 0; i < nb * B; ++i) {
            const int64_t s = perm[i];
            std::memcpy(&zz[(size_t)i * QIW_FLOW_D], &z[(size_t)s * QIW_FLOW_D], QIW_FLOW_D * 8);
            oo[i] = ord[s];
            gg[i] = grp[s];
            ss[i] = (float)stp[s];
            cc[i] = (float)cf[s];
        }
        HIPCHK(c, hipMemcpy(dev_z, zz.data(), zz.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dev_o, oo.data(), oo.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dev_g, gg.data(), gg.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dev_s, ss.data(), ss.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(dev_c, cc.data(), cc.size() * 4, hipMemcpyHostToDevice));
    }
    Work k;
    if ((rc = work(c, w, B, true, k))) return rc;
    const float b1 = 0.9f, b2 = 0.999f;
    for (int t = 0; t < cfg.iters; ++t) {
        const int64_t r0 = (int64_t)(t % nb) * B;
        hipStream_t st = c->stream;
        HIPCHK(c, hipMemcpyAsync(k.x, dev_z + r0 * QIW_FLOW_D, (size_t)B * QIW_FLOW_D * 8, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(k.order, dev_o + r0, (size_t)B * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(k.group, dev_g + r0, (size_t)B * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(k.step, dev_s + r0, (size_t)B * 4, hipMemcpyDeviceToDevice, st));
        HIPCHK(c, hipMemcpyAsync(k.coef, dev_c + r0, (size_t)B * 4, hipMemcpyDeviceToDevice, st));
        if ((rc = forward(c, w, B, true, k))) return rc;
        if ((rc = backward(c, w, B, k))) return rc;
        HIPCHK(c, hipMemcpyAsync(dev_l + t, k.loss, 8, hipMemcpyDeviceToDevice, st));
        ++w.adam_t;
        const float bc1 = 1.0f - std::pow(b1, (float)w.adam_t), bc2 = 1.0f - std::pow(b2, (float)w.adam_t);
        HIPCHK(c, qiw_flow_adam(st, w.params, w.grad, w.m1, w.m2, w.n_params, (float)cfg.lr, b1, b2, 1e-8f, bc1, bc2));
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (loss_history && cfg.iters > 0)
        HIPCHK(c, hipMemcpy(loss_history, dev_l, (size_t)cfg.iters * 8, hipMemcpyDeviceToHost));
    return QIW_OK;
}

static int qiw_warp_enable_impl(qiw_ctx *c, int32_t on) {
    int rc = need_warp(c);
    if (rc) return rc;
    c->warp->enabled = on != 0;
    return QIW_OK;
}

// ---- C ABI: every entry point runs its body through guarded() ----
extern "C" int qiw_warp_init(qiw_ctx *c, const qiw_warp_config *cfg) {
    return guarded(c, [&] { return qiw_warp_init_impl(c, cfg); });
}

extern "C" int qiw_warp_params(qiw_ctx *c, int64_t *n_out, float *get, const float *set) {
    return guarded(c, [&] { return qiw_warp_params_impl(c, n_out, get, set); });
}

extern "C" int qiw_warp_apply(qiw_ctx *c, int32_t inv, int64_t n, const double *pts, const int32_t *order,
                              const int32_t *group, const double *step, double *out, double *logdet) {
    return guarded(c, [&] { return qiw_warp_apply_impl(c, inv, n, pts, order, group, step, out, logdet); });
}

extern "C" int qiw_warp_loss_grad(qiw_ctx *c, int64_t n, const double *z, const int32_t *order,
                                  const int32_t *group, const double *step, const double *coef, double *loss,
                                  float *grad) {
    return guarded(c, [&] { return qiw_warp_loss_grad_impl(c, n, z, order, group, step, coef, loss, grad); });
}

extern "C" int qiw_warp_train(qiw_ctx *c, const qiw_run *run, double *loss_history) {
    return guarded(c, [&] { return qiw_warp_train_impl(c, run, loss_history); });
}

extern "C" int qiw_warp_enable(qiw_ctx *c, int32_t on) {
    return guarded(c, [&] { return qiw_warp_enable_impl(c, on); });
}
