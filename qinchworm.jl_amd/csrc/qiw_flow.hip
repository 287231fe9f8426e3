// qiw_flow.hip -- gfx950 kernels of the learned warp: batched conditioner layers (float32)
// with their backward passes, the rational-quadratic spline transforms and their
// gradients (float64), the loss and the Adam update.
//
// The spline of one transformed coordinate (Durkan et al. 2019, eqs. 4-6 and appendix A,
// on [0, 1] with end knots (0, 0) and (1, 1)):
//   widths  w_j = MIN_W + (1 - K MIN_W) softmax(raw[0..K))_j,   knots X_0 = 0, X_{j+1} = X_j + w_j
//   heights h_j = MIN_W + (1 - K MIN_W) softmax(raw[K..2K))_j,  knots Y likewise
//   slopes  delta_j = MIN_D + (1 - MIN_D) softplus(raw[2K + j]) / log 2,  j = 0..K
//   in bin k with xi = (x - X_k) / w_k, s = h_k / w_k, t = xi (1 - xi):
//   y = Y_k + h_k (s xi^2 + delta_k t) / (s + (delta_k + delta_{k+1} - 2 s) t)
// Gradients of y and log dy/dx with respect to (x, X_k, Y_k, w_k, h_k, delta_k,
// delta_{k+1}) come from forward-mode dual numbers, then pass through the cumulative sums,
// the softmax and the softplus.
#include <hip/hip_runtime.h>
#include <math.h>

#include "qiw_flow.h"

namespace {

constexpr int NB = QIW_FLOW_MAX_BINS;

// ---- spline parameters of one coordinate --------------------------------------------------
struct Spline {
    double X[NB + 1], Y[NB + 1], dl[NB + 1];   // knots and knot slopes
    double smw[NB], smh[NB];                     // softmax of the raw widths / heights
    double sgd[NB + 1];                          // d delta / d raw
};

__device__ double softplus(double x) { return log1p(exp(-fabs(x))) + fmax(x, 0.0); }

__device__ void spline_setup(const float *__restrict__ raw, int K, Spline &s) {
    const double scale = 1.0 - K * QIW_FLOW_MIN_W;
    double mw = -INFINITY, mh = -INFINITY;
    for (int j = 0; j < K; ++j) {
        mw = fmax(mw, (double)raw[j]);
        mh = fmax(mh, (double)raw[K + j]);
    }
    double zw = 0.0, zh = 0.0;
    for (int j = 0; j < K; ++j) {
        s.smw[j] = exp((double)raw[j] - mw);
        s.smh[j] = exp((double)raw[K + j] - mh);
        zw += s.smw[j];
        zh += s.smh[j];
    }
    s.X[0] = 0.0;
    s.Y[0] = 0.0;
    for (int j = 0; j < K; ++j) {
        s.smw[j] /= zw;
        s.smh[j] /= zh;
        s.X[j + 1] = s.X[j] + (QIW_FLOW_MIN_W + scale * s.smw[j]);
        s.Y[j + 1] = s.Y[j] + (QIW_FLOW_MIN_W + scale * s.smh[j]);
    }
    s.X[K] = 1.0;
    s.Y[K] = 1.0;
    const double il2 = 1.0 / log(2.0);
    for (int j = 0; j <= K; ++j) {
        const double r = (double)raw[2 * K + j];
        s.dl[j] = QIW_FLOW_MIN_D + (1.0 - QIW_FLOW_MIN_D) * softplus(r) * il2;
        s.sgd[j] = (1.0 - QIW_FLOW_MIN_D) * il2 / (1.0 + exp(-r));
    }
}

__device__ int find_bin(const double *knots, int K, double v) {
    int k = 0;
    for (int j = 1; j < K; ++j) k += (knots[j] <= v) ? 1 : 0;
    return k;
}

__device__ void rqs_eval(const Spline &s, int k, double x, double &y, double &ld) {
    const double X = s.X[k], w = s.X[k + 1] - s.X[k];
    const double Y = s.Y[k], h = s.Y[k + 1] - s.Y[k];
    const double d0 = s.dl[k], d1 = s.dl[k + 1];
    const double sl = h / w, xi = (x - X) / w, t = xi * (1.0 - xi);
    const double den = sl + (d0 + d1 - 2.0 * sl) * t;
    y = Y + h * (sl * xi * xi + d0 * t) / den;
    const double dydx = sl * sl * (d1 * xi * xi + 2.0 * sl * t + d0 * (1.0 - xi) * (1.0 - xi)) / (den * den);
    ld = log(dydx);
}

__device__ double rqs_invert(const Spline &s, int k, double y) {
    const double X = s.X[k], w = s.X[k + 1] - s.X[k];
    const double Y = s.Y[k], h = s.Y[k + 1] - s.Y[k];
    const double d0 = s.dl[k], d1 = s.dl[k + 1];
    const double sl = h / w, dy = y - Y, D = d0 + d1 - 2.0 * sl;
    const double a = h * (sl - d0) + dy * D;
    const double b = h * d0 - dy * D;
    const double c = -sl * dy;
    const double disc = fmax(b * b - 4.0 * a * c, 0.0);
    const double xi = 2.0 * c / (-b - sqrt(disc));
    return X + fmin(fmax(xi, 0.0), 1.0) * w;
}

// ---- forward-mode dual numbers over 7 inputs -------------------------------------------------
struct D7 {
    double v, d[7];
};
__device__ D7 dconst(double v) {
    D7 r;
    r.v = v;
    for (int i = 0; i < 7; ++i) r.d[i] = 0.0;
    return r;
}
__device__ D7 dvar(double v, int i) {
    D7 r = dconst(v);
    r.d[i] = 1.0;
    return r;
}
__device__ D7 operator+(const D7 &a, const D7 &b) {
    D7 r;
    r.v = a.v + b.v;
    for (int i = 0; i < 7; ++i) r.d[i] = a.d[i] + b.d[i];
    return r;
}
__device__ D7 operator-(const D7 &a, const D7 &b) {
    D7 r;
    r.v = a.v - b.v;
    for (int i = 0; i < 7; ++i) r.d[i] = a.d[i] - b.d[i];
    return r;
}
__device__ D7 operator*(const D7 &a, const D7 &b) {
    D7 r;
    r.v = a.v * b.v;
    for (int i = 0; i < 7; ++i) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
    return r;
}
__device__ D7 operator*(double c, const D7 &a) {
    D7 r;
    r.v = c * a.v;
    for (int i = 0; i < 7; ++i) r.d[i] = c * a.d[i];
    return r;
}
__device__ D7 operator/(const D7 &a, const D7 &b) {
    D7 r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int i = 0; i < 7; ++i) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
    return r;
}
__device__ D7 dlog(const D7 &a) {
    D7 r;
    r.v = log(a.v);
    for (int i = 0; i < 7; ++i) r.d[i] = a.d[i] / a.v;
    return r;
}

// y and log dy/dx as functions of (x, X_k, Y_k, w_k, h_k, delta_k, delta_{k+1})
__device__ void rqs_dual(const Spline &s, int k, double x, D7 &y, D7 &ld) {
    const D7 dx = dvar(x, 0), X = dvar(s.X[k], 1), Y = dvar(s.Y[k], 2);
    const D7 w = dvar(s.X[k + 1] - s.X[k], 3), h = dvar(s.Y[k + 1] - s.Y[k], 4);
    const D7 d0 = dvar(s.dl[k], 5), d1 = dvar(s.dl[k + 1], 6);
    const D7 one = dconst(1.0);
    const D7 sl = h / w, xi = (dx - X) / w, omx = one - xi, t = xi * omx;
    const D7 den = sl + (d0 + d1 - 2.0 * sl) * t;
    y = Y + h * (sl * xi * xi + d0 * t) / den;
    const D7 dydx = sl * sl * (d1 * xi * xi + 2.0 * sl * t + d0 * omx * omx) / (den * den);
    ld = dlog(dydx);
}

__device__ int slot_coord(int layer, int b) { return 2 * b + 1 - (layer & 1); }  // transformed
__device__ int cond_coord(int layer, int a) { return 2 * a + (layer & 1); }       // conditioning

// ---------------------------------------------------------------------------------------
__global__ void k_flow_features(int layer, const double *__restrict__ x, const int32_t *__restrict__ order,
                                const int32_t *__restrict__ group, const float *__restrict__ step,
                                const float *__restrict__ emb, float *__restrict__ feat, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int k = order[r], d = 2 * k;
    float *f = feat + (int64_t)r * QIW_FLOW_FEAT;
    for (int a = 0; a < QIW_FLOW_SLOTS; ++a) {
        const int i = cond_coord(layer, a);
        f[a] = (i < d) ? (float)x[(int64_t)r * QIW_FLOW_D + i] : 0.0f;
    }
    for (int o = 0; o < QIW_FLOW_KMAX; ++o) f[QIW_FLOW_SLOTS + o] = (o == k - 1) ? 1.0f : 0.0f;
    const float *e = emb + (int64_t)group[r] * QIW_FLOW_EMBED;
    for (int q = 0; q < QIW_FLOW_EMBED; ++q) f[QIW_FLOW_SLOTS + QIW_FLOW_KMAX + q] = e[q];
    f[QIW_FLOW_FEAT - 1] = step[r];
}

// out[r][o] = act(sum_i in[r][i] W[i][o] + b[o]) for a tile of QIW_FLOW_TILE rows; the weight
// matrix and the input rows are staged in LDS.
__global__ __launch_bounds__(256) void k_flow_dense(const float *__restrict__ in, const float *__restrict__ W,
                                                    const float *__restrict__ b, float *__restrict__ out, int n,
                                                    int n_in, int n_out, int tanh_act) {
    extern __shared__ float sh[];
    float *Ws = sh;
    float *Xs = sh + n_in * n_out;
    const int r0 = blockIdx.x * QIW_FLOW_TILE;
    for (int i = threadIdx.x; i < n_in * n_out; i += blockDim.x) Ws[i] = W[i];
    for (int i = threadIdx.x; i < QIW_FLOW_TILE * n_in; i += blockDim.x) {
        const int r = r0 + i / n_in;
        Xs[i] = (r < n) ? in[(int64_t)r0 * n_in + i] : 0.0f;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < QIW_FLOW_TILE * n_out; idx += blockDim.x) {
        const int rr = idx / n_out, o = idx % n_out;
        if (r0 + rr >= n) continue;
        float acc = b[o];
        const float *xr = Xs + rr * n_in;
        for (int i = 0; i < n_in; ++i) acc = fmaf(xr[i], Ws[i * n_out + o], acc);
        out[(int64_t)(r0 + rr) * n_out + o] = tanh_act ? tanhf(acc) : acc;
    }
}

// One sample per thread: spline transform of the slots a layer transforms, forward (the
// training direction z -> x) or inverse (sampling, x -> z); logdet accumulates
// log dy/dx (forward) or log dx/dy (inverse).
__global__ void k_flow_rqs(int layer, const double *__restrict__ xin, double *__restrict__ xout,
                           const float *__restrict__ raw, const int32_t *__restrict__ order,
                           double *__restrict__ logdet, int n, int K, int inverse) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int P = 3 * K + 1, d = 2 * order[r];
    for (int i = 0; i < QIW_FLOW_D; ++i) xout[(int64_t)r * QIW_FLOW_D + i] = xin[(int64_t)r * QIW_FLOW_D + i];
    double acc = 0.0;
    Spline s;
    for (int b = 0; b < QIW_FLOW_SLOTS; ++b) {
        const int i = slot_coord(layer, b);
        if (i >= d) continue;
        spline_setup(raw + ((int64_t)r * QIW_FLOW_SLOTS + b) * P, K, s);
        const double v = xin[(int64_t)r * QIW_FLOW_D + i];
        double y, ld;
        if (!inverse) {
            rqs_eval(s, find_bin(s.X, K, v), v, y, ld);
        } else {
            const int k = find_bin(s.Y, K, v);
            y = rqs_invert(s, k, v);
            double yy;
            rqs_eval(s, k, y, yy, ld);
            ld = -ld;
        }
        xout[(int64_t)r * QIW_FLOW_D + i] = y;
        acc += ld;
    }
    logdet[r] += acc;
}

// Backward of one forward (training-direction) layer.  gx_out = dL/dx after the layer;
// the loss term of the layer's log-Jacobian is -scale * coef[r] * logdet.  Writes
// gx_in = dL/dx before the layer (spline part; the conditioner's part is added by
// k_flow_feat_bwd) and graw = dL/draw of the conditioner outputs.
__global__ void k_flow_rqs_bwd(int layer, const double *__restrict__ xin, const float *__restrict__ raw,
                               const int32_t *__restrict__ order, const float *__restrict__ coef, double scale,
                               const double *__restrict__ gx_out, double *__restrict__ gx_in,
                               float *__restrict__ graw, int n, int K) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int P = 3 * K + 1, d = 2 * order[r];
    const double gl = -scale * (double)coef[r];
    for (int i = 0; i < QIW_FLOW_D; ++i) gx_in[(int64_t)r * QIW_FLOW_D + i] = gx_out[(int64_t)r * QIW_FLOW_D + i];
    Spline s;
    for (int b = 0; b < QIW_FLOW_SLOTS; ++b) {
        float *gr = graw + ((int64_t)r * QIW_FLOW_SLOTS + b) * P;
        const int i = slot_coord(layer, b);
        if (i >= d) {
            for (int j = 0; j < P; ++j) gr[j] = 0.0f;
            continue;
        }
        spline_setup(raw + ((int64_t)r * QIW_FLOW_SLOTS + b) * P, K, s);
        const double v = xin[(int64_t)r * QIW_FLOW_D + i];
        const int k = find_bin(s.X, K, v);
        D7 y, ld;
        rqs_dual(s, k, v, y, ld);
        const double gy = gx_out[(int64_t)r * QIW_FLOW_D + i];
        double g[7];
        for (int q = 0; q < 7; ++q) g[q] = gy * y.d[q] + gl * ld.d[q];
        gx_in[(int64_t)r * QIW_FLOW_D + i] = g[0];
        // widths / heights: X_k = sum_{j<k} w_j, so dL/dw_j = g_X (j < k) + g_w (j = k)
        const double scl = 1.0 - K * QIW_FLOW_MIN_W;
        double sw = 0.0, shh = 0.0;
        for (int j = 0; j < K; ++j) {
            const double dw = (j < k ? g[1] : 0.0) + (j == k ? g[3] : 0.0);
            const double dh = (j < k ? g[2] : 0.0) + (j == k ? g[4] : 0.0);
            sw += s.smw[j] * dw;
            shh += s.smh[j] * dh;
        }
        for (int j = 0; j < K; ++j) {
            const double dw = (j < k ? g[1] : 0.0) + (j == k ? g[3] : 0.0);
            const double dh = (j < k ? g[2] : 0.0) + (j == k ? g[4] : 0.0);
            gr[j] = (float)(scl * s.smw[j] * (dw - sw));
            gr[K + j] = (float)(scl * s.smh[j] * (dh - shh));
        }
        for (int j = 0; j <= K; ++j) {
            const double dd = (j == k ? g[5] : 0.0) + (j == k + 1 ? g[6] : 0.0);
            gr[2 * K + j] = (float)(dd * s.sgd[j]);
        }
    }
}

__global__ void k_flow_tanh_bwd(const float *__restrict__ dout, const float *__restrict__ out,
                                float *__restrict__ dpre, int64_t count) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float y = out[i];
    dpre[i] = dout[i] * (1.0f - y * y);
}

// din[r][i] = sum_o dpre[r][o] W[i][o]
__global__ __launch_bounds__(256) void k_flow_dense_bwd_data(const float *__restrict__ dpre,
                                                             const float *__restrict__ W, float *__restrict__ din,
                                                             int n, int n_in, int n_out) {
    extern __shared__ float sh[];
    float *Ws = sh;
    float *Gs = sh + n_in * n_out;
    const int r0 = blockIdx.x * QIW_FLOW_TILE;
    for (int i = threadIdx.x; i < n_in * n_out; i += blockDim.x) Ws[i] = W[i];
    for (int i = threadIdx.x; i < QIW_FLOW_TILE * n_out; i += blockDim.x) {
        const int r = r0 + i / n_out;
        Gs[i] = (r < n) ? dpre[(int64_t)r0 * n_out + i] : 0.0f;
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < QIW_FLOW_TILE * n_in; idx += blockDim.x) {
        const int rr = idx / n_in, i = idx % n_in;
        if (r0 + rr >= n) continue;
        float acc = 0.0f;
        const float *gr = Gs + rr * n_out;
        const float *wr = Ws + i * n_out;
        for (int o = 0; o < n_out; ++o) acc = fmaf(gr[o], wr[o], acc);
        din[(int64_t)(r0 + rr) * n_in + i] = acc;
    }
}

// Partial sums over one chunk of rows: part[chunk][i*n_out + o] = sum_r in[r][i] dpre[r][o],
// part[chunk][n_in*n_out + o] = sum_r dpre[r][o].
__global__ __launch_bounds__(256) void k_flow_dense_bwd_weight(const float *__restrict__ in,
                                                               const float *__restrict__ dpre,
                                                               float *__restrict__ part, int n, int n_in,
                                                               int n_out) {
    const int ne = (n_in + 1) * n_out;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int r0 = blockIdx.y * QIW_FLOW_CHUNK, r1 = min(n, r0 + QIW_FLOW_CHUNK);
    const int i = e / n_out, o = e % n_out;
    float acc = 0.0f;
    if (i < n_in) {
        for (int r = r0; r < r1; ++r) acc = fmaf(in[(int64_t)r * n_in + i], dpre[(int64_t)r * n_out + o], acc);
    } else {
        for (int r = r0; r < r1; ++r) acc += dpre[(int64_t)r * n_out + o];
    }
    part[(int64_t)blockIdx.y * ne + e] = acc;
}

// dW += sum over chunks (fixed order), db likewise
__global__ void k_flow_reduce_chunks(const float *__restrict__ part, int n_chunks, int ne, int n_w,
                                     float *__restrict__ dW, float *__restrict__ db) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    float acc = 0.0f;
    for (int c = 0; c < n_chunks; ++c) acc += part[(int64_t)c * ne + e];
    if (e < n_w)
        dW[e] += acc;
    else
        db[e - n_w] += acc;
}

// conditioning coordinates: gx[A] += dfeat[slot]
__global__ void k_flow_feat_bwd_x(int layer, const float *__restrict__ dfeat, const int32_t *__restrict__ order,
                                  double *__restrict__ gx, int n) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const int d = 2 * order[r];
    for (int a = 0; a < QIW_FLOW_SLOTS; ++a) {
        const int i = cond_coord(layer, a);
        if (i < d) gx[(int64_t)r * QIW_FLOW_D + i] += (double)dfeat[(int64_t)r * QIW_FLOW_FEAT + a];
    }
}

// embedding rows: partial sums per chunk of samples, [chunk][GROUPS*EMBED]
__global__ void k_flow_emb_partial(const float *__restrict__ dfeat, const int32_t *__restrict__ group,
                                   float *__restrict__ part, int n) {
    const int e = threadIdx.x;  // GROUPS * EMBED = 256 threads
    const int g = e / QIW_FLOW_EMBED, q = e % QIW_FLOW_EMBED;
    const int r0 = blockIdx.x * QIW_FLOW_CHUNK, r1 = min(n, r0 + QIW_FLOW_CHUNK);
    float acc = 0.0f;
    for (int r = r0; r < r1; ++r)
        if (group[r] == g) acc += dfeat[(int64_t)r * QIW_FLOW_FEAT + QIW_FLOW_SLOTS + QIW_FLOW_KMAX + q];
    part[(int64_t)blockIdx.x * (QIW_FLOW_GROUPS * QIW_FLOW_EMBED) + e] = acc;
}

__global__ void k_flow_emb_reduce(const float *__restrict__ part, int n_chunks, float *__restrict__ demb) {
    const int e = threadIdx.x;
    float acc = 0.0f;
    for (int c = 0; c < n_chunks; ++c) acc += part[(int64_t)c * (QIW_FLOW_GROUPS * QIW_FLOW_EMBED) + e];
    demb[e] += acc;
}

// loss = -scale * sum_r coef[r] logq[r], one workgroup, fixed summation order
__global__ __launch_bounds__(256) void k_flow_loss(const double *__restrict__ logq, const float *__restrict__ coef,
                                                   double scale, double *__restrict__ loss, int n) {
    __shared__ double sh[256];
    double acc = 0.0;
    for (int r = threadIdx.x; r < n; r += 256) acc += (double)coef[r] * logq[r];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) sh[threadIdx.x] += sh[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *loss = -scale * sh[0];
}

__global__ void k_flow_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int64_t count, float lr, float b1, float b2, float eps, float bc1,
                            float bc2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void k_flow_fill(float *__restrict__ p, int64_t count, float value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

__global__ void k_flow_fill_f64(double *__restrict__ p, int64_t count, double value) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = value;
}

unsigned nblk(int64_t n, int b) { return (unsigned)((n + b - 1) / b); }

}  // namespace

// ---------------------------------------------------------------------------------------
// Launch wrappers
// ---------------------------------------------------------------------------------------
hipError_t qiw_flow_features(hipStream_t st, int layer, const double *x, const int32_t *order,
                             const int32_t *group, const float *step, const float *emb, float *feat, int n) {
    hipLaunchKernelGGL(k_flow_features, dim3(nblk(n, 256)), dim3(256), 0, st, layer, x, order, group, step, emb,
                       feat, n);
    return hipGetLastError();
}

hipError_t qiw_flow_dense(hipStream_t st, const float *in, const float *W, const float *b, float *out, int n,
                          int n_in, int n_out, int tanh_act) {
    const size_t lds = ((size_t)n_in * n_out + (size_t)QIW_FLOW_TILE * n_in) * sizeof(float);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flow_dense, dim3(nblk(n, QIW_FLOW_TILE)), dim3(256), lds, st, in, W, b, out, n, n_in,
                       n_out, tanh_act);
    return hipGetLastError();
}

hipError_t qiw_flow_rqs(hipStream_t st, int layer, const double *xin, double *xout, const float *raw,
                        const int32_t *order, double *logdet, int n, int bins, int inverse) {
    if (bins < 1 || bins > QIW_FLOW_MAX_BINS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flow_rqs, dim3(nblk(n, 64)), dim3(64), 0, st, layer, xin, xout, raw, order, logdet, n,
                       bins, inverse);
    return hipGetLastError();
}

hipError_t qiw_flow_rqs_bwd(hipStream_t st, int layer, const double *xin, const float *raw, const int32_t *order,
                            const float *coef, double scale, const double *gx_out, double *gx_in, float *graw,
                            int n, int bins) {
    if (bins < 1 || bins > QIW_FLOW_MAX_BINS) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flow_rqs_bwd, dim3(nblk(n, 64)), dim3(64), 0, st, layer, xin, raw, order, coef, scale,
                       gx_out, gx_in, graw, n, bins);
    return hipGetLastError();
}

hipError_t qiw_flow_tanh_bwd(hipStream_t st, const float *dout, const float *out, float *dpre, int64_t count) {
    hipLaunchKernelGGL(k_flow_tanh_bwd, dim3(nblk(count, 256)), dim3(256), 0, st, dout, out, dpre, count);
    return hipGetLastError();
}

hipError_t qiw_flow_dense_bwd_data(hipStream_t st, const float *dpre, const float *W, float *din, int n, int n_in,
                                   int n_out) {
    const size_t lds = ((size_t)n_in * n_out + (size_t)QIW_FLOW_TILE * n_out) * sizeof(float);
    if (lds > 160 * 1024) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_flow_dense_bwd_data, dim3(nblk(n, QIW_FLOW_TILE)), dim3(256), lds, st, dpre, W, din, n,
                       n_in, n_out);
    return hipGetLastError();
}

// partial must hold ceil(n / QIW_FLOW_CHUNK) * (n_in + 1) * n_out floats
hipError_t qiw_flow_dense_bwd_weight(hipStream_t st, const float *in, const float *dpre, float *partial,
                                     float *dW, float *db, int n, int n_in, int n_out) {
    const int ne = (n_in + 1) * n_out;
    const int chunks = (int)nblk(n, QIW_FLOW_CHUNK);
    hipLaunchKernelGGL(k_flow_dense_bwd_weight, dim3(nblk(ne, 256), chunks), dim3(256), 0, st, in, dpre, partial,
                       n, n_in, n_out);
    hipLaunchKernelGGL(k_flow_reduce_chunks, dim3(nblk(ne, 256)), dim3(256), 0, st, partial, chunks, ne,
                       n_in * n_out, dW, db);
    return hipGetLastError();
}

// partial must hold ceil(n / QIW_FLOW_CHUNK) * GROUPS * EMBED floats
hipError_t qiw_flow_feat_bwd(hipStream_t st, int layer, const float *dfeat, const int32_t *order,
                             const int32_t *group, double *gx, float *partial, float *demb, int n) {
    hipLaunchKernelGGL(k_flow_feat_bwd_x, dim3(nblk(n, 256)), dim3(256), 0, st, layer, dfeat, order, gx, n);
    const int chunks = (int)nblk(n, QIW_FLOW_CHUNK);
    hipLaunchKernelGGL(k_flow_emb_partial, dim3(chunks), dim3(QIW_FLOW_GROUPS * QIW_FLOW_EMBED), 0, st, dfeat,
                       group, partial, n);
    hipLaunchKernelGGL(k_flow_emb_reduce, dim3(1), dim3(QIW_FLOW_GROUPS * QIW_FLOW_EMBED), 0, st, partial, chunks,
                       demb);
    return hipGetLastError();
}

hipError_t qiw_flow_loss(hipStream_t st, const double *logq, const float *coef, double scale, double *loss, int n) {
    hipLaunchKernelGGL(k_flow_loss, dim3(1), dim3(256), 0, st, logq, coef, scale, loss, n);
    return hipGetLastError();
}

hipError_t qiw_flow_adam(hipStream_t st, float *p, const float *g, float *m, float *v, int64_t count, float lr,
                         float b1, float b2, float eps, float bc1, float bc2) {
    hipLaunchKernelGGL(k_flow_adam, dim3(nblk(count, 256)), dim3(256), 0, st, p, g, m, v, count, lr, b1, b2, eps,
                       bc1, bc2);
    return hipGetLastError();
}

hipError_t qiw_flow_fill(hipStream_t st, float *p, int64_t count, float value) {
    hipLaunchKernelGGL(k_flow_fill, dim3(nblk(count, 256)), dim3(256), 0, st, p, count, value);
    return hipGetLastError();
}

hipError_t qiw_flow_fill_f64(hipStream_t st, double *p, int64_t count, double value) {
    hipLaunchKernelGGL(k_flow_fill_f64, dim3(nblk(count, 256)), dim3(256), 0, st, p, count, value);
    return hipGetLastError();
}
