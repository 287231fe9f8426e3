"""flow_oracle.py -- float64 NumPy restatement of libqiw's learned warp of the unit cube.

TEST INFRASTRUCTURE ONLY.  This module is the checker for the GPU flow kernels
(qinchworm.jl_amd/csrc/qiw_flow.hip): tests compare the device forward pass, inverse
pass, log-Jacobians, loss and parameter gradients with the float64 functions below.  The
product never imports it.

The warp has no counterpart in the reference: QInchworm.jl integrates with plain
(unwarped) quasi Monte Carlo (src/qmc_integrate.jl:497-507).  It is a normalising flow
made of coupling layers with monotone rational-quadratic splines (Durkan, Bekasov, Murray,
Papamakarios, "Neural Spline Flows", NeurIPS 2019, eqs. 4-6 and appendix A), restricted to
[0, 1] with both end knots fixed at (0, 0) and (1, 1), so that every layer maps the unit
cube onto itself.

Layout (shared with the device, float32 there):
  coordinates    d <= D_MAX = 8 (2k for order k); a layer l conditions on the coordinates
                 i with i % 2 == l % 2 (slot a <-> i = 2a + l % 2) and transforms the
                 others (slot b <-> i = 2b + 1 - l % 2); slots beyond d are unused
  features       [x_A (4 slots, 0 where unused), one-hot(order - 1) (4), embedding of
                 the integration group (16), step fraction tau_w / beta (1)]  -> 25
  conditioner    tanh(f W1 + b1) -> tanh(. W2 + b2) -> . W3 + b3, widths 25 -> H -> H -> 100
  spline         per transformed slot 3K + 1 raw values: K widths, K heights, K + 1
                 knot derivatives; w = MIN_W + (1 - K MIN_W) softmax, same for heights,
                 delta = MIN_D + (1 - MIN_D) softplus(raw) / log 2 (raw 0 -> delta 1)
  parameters     per layer W1 [25][H], b1 [H], W2 [H][H], b2 [H], W3 [H][100], b3 [100]
                 (row-major), then the group embedding table [N_GROUPS][16]

Direction: `forward` maps a sample point z of the plain method to x (the density q of the
warped points at z is exp(logq)); `inverse` maps a qMC point x to the warped point z with
log |dz/dx| = logJ = -logq(z).  The warped estimator is mean(f(z_i) exp(logJ_i)).
"""
import numpy as np

D_MAX = 8
SLOTS = D_MAX // 2
K_MAX = 4
N_EMBED = 16
N_GROUPS = 16
N_FEAT = SLOTS + K_MAX + N_EMBED + 1
MIN_W = 1e-3
MIN_D = 1e-3


class Spec:
    def __init__(self, n_layers=4, n_bins=8, hidden=64):
        self.L, self.K, self.H = n_layers, n_bins, hidden
        self.P = 3 * n_bins + 1           # raw spline values per transformed slot
        self.n_out = SLOTS * self.P

    def layer_sizes(self):
        H, O = self.H, self.n_out
        return [("W1", N_FEAT * H), ("b1", H), ("W2", H * H), ("b2", H), ("W3", H * O), ("b3", O)]

    def n_params(self):
        return self.L * sum(n for _, n in self.layer_sizes()) + N_GROUPS * N_EMBED

    def unpack(self, theta):
        """Flat parameter vector -> list of per-layer dicts and the embedding table."""
        theta = np.asarray(theta, dtype=np.float64)
        layers, o = [], 0
        shapes = {"W1": (N_FEAT, self.H), "W2": (self.H, self.H), "W3": (self.H, self.n_out)}
        for _ in range(self.L):
            d = {}
            for name, n in self.layer_sizes():
                v = theta[o:o + n]
                d[name] = v.reshape(shapes[name]) if name in shapes else v
                o += n
            layers.append(d)
        emb = theta[o:o + N_GROUPS * N_EMBED].reshape(N_GROUPS, N_EMBED)
        return layers, emb


def softplus(x):
    return np.log1p(np.exp(-np.abs(x))) + np.maximum(x, 0.0)


def spline_params(raw, K):
    """raw [..., 3K+1] -> knots X, Y [..., K+1] and derivatives delta [..., K+1]."""
    def norm(r):
        e = np.exp(r - r.max(axis=-1, keepdims=True))
        sm = e / e.sum(axis=-1, keepdims=True)
        return MIN_W + (1.0 - K * MIN_W) * sm
    w = norm(raw[..., :K])
    h = norm(raw[..., K:2 * K])
    delta = MIN_D + (1.0 - MIN_D) * softplus(raw[..., 2 * K:]) / np.log(2.0)
    X = np.concatenate([np.zeros(w.shape[:-1] + (1,)), np.cumsum(w, axis=-1)], axis=-1)
    Y = np.concatenate([np.zeros(h.shape[:-1] + (1,)), np.cumsum(h, axis=-1)], axis=-1)
    X[..., -1] = 1.0
    Y[..., -1] = 1.0
    return X, Y, delta


def rqs_forward(x, raw, K):
    """Rational-quadratic spline y(x) and log dy/dx, elementwise over leading axes."""
    X, Y, dl = spline_params(raw, K)
    k = np.clip(np.sum(X[..., 1:-1] <= x[..., None], axis=-1), 0, K - 1)
    take = lambda A, i: np.take_along_axis(A, i[..., None], axis=-1)[..., 0]  # noqa: E731
    xk, wk = take(X, k), take(X, k + 1) - take(X, k)
    yk, hk = take(Y, k), take(Y, k + 1) - take(Y, k)
    d0, d1 = take(dl, k), take(dl, k + 1)
    s = hk / wk
    xi = (x - xk) / wk
    t = xi * (1.0 - xi)
    den = s + (d0 + d1 - 2.0 * s) * t
    y = yk + hk * (s * xi * xi + d0 * t) / den
    dydx = s * s * (d1 * xi * xi + 2.0 * s * t + d0 * (1.0 - xi) ** 2) / (den * den)
    return y, np.log(dydx)


def rqs_inverse(y, raw, K):
    """x(y) of the same spline and log dx/dy."""
    X, Y, dl = spline_params(raw, K)
    k = np.clip(np.sum(Y[..., 1:-1] <= y[..., None], axis=-1), 0, K - 1)
    take = lambda A, i: np.take_along_axis(A, i[..., None], axis=-1)[..., 0]  # noqa: E731
    xk, wk = take(X, k), take(X, k + 1) - take(X, k)
    yk, hk = take(Y, k), take(Y, k + 1) - take(Y, k)
    d0, d1 = take(dl, k), take(dl, k + 1)
    s = hk / wk
    dy = y - yk
    D = d0 + d1 - 2.0 * s
    a = hk * (s - d0) + dy * D
    b = hk * d0 - dy * D
    c = -s * dy
    xi = 2.0 * c / (-b - np.sqrt(b * b - 4.0 * a * c))
    x = xk + xi * wk
    _, ld = rqs_forward(x, raw, K)
    return x, -ld


def features(x, l, order, group, step, emb):
    n = x.shape[0]
    f = np.zeros((n, N_FEAT))
    d = 2 * order
    for a in range(SLOTS):
        i = 2 * a + (l % 2)
        f[:, a] = np.where(i < d, x[:, min(i, D_MAX - 1)], 0.0)
    f[np.arange(n), SLOTS + order - 1] = 1.0
    f[:, SLOTS + K_MAX:SLOTS + K_MAX + N_EMBED] = emb[group]
    f[:, -1] = step
    return f


def conditioner(f, lp, spec):
    h1 = np.tanh(f @ lp["W1"] + lp["b1"])
    h2 = np.tanh(h1 @ lp["W2"] + lp["b2"])
    return (h2 @ lp["W3"] + lp["b3"]).reshape(-1, SLOTS, spec.P)


def _apply(spec, theta, x, order, group, step, inverse):
    layers, emb = spec.unpack(theta)
    x = np.array(x, dtype=np.float64, copy=True)
    order = np.asarray(order)
    group = np.asarray(group)
    step = np.asarray(step, dtype=np.float64)
    logdet = np.zeros(x.shape[0])
    seq = range(spec.L - 1, -1, -1) if inverse else range(spec.L)
    for l in seq:
        raw = conditioner(features(x, l, order, group, step, emb), layers[l], spec)
        for b in range(SLOTS):
            i = 2 * b + 1 - (l % 2)
            valid = i < 2 * order
            if not np.any(valid):
                continue
            col = x[:, min(i, D_MAX - 1)]
            y, ld = (rqs_inverse if inverse else rqs_forward)(col, raw[:, b, :], spec.K)
            x[:, min(i, D_MAX - 1)] = np.where(valid, y, col)
            logdet += np.where(valid, ld, 0.0)
    return x, logdet


def forward(spec, theta, z, order, group, step):
    """z [n][8] -> (x, log q(z)) with log q = sum of the layers' log dy/dx."""
    return _apply(spec, theta, z, order, group, step, inverse=False)


def inverse(spec, theta, x, order, group, step):
    """x [n][8] -> (z, log |dz/dx|)."""
    return _apply(spec, theta, x, order, group, step, inverse=True)


def loss(spec, theta, z, order, group, step, coef):
    """Weighted negative log-likelihood of the plain sample points: -mean(coef * log q(z))."""
    _, lq = forward(spec, theta, z, order, group, step)
    return -np.mean(np.asarray(coef) * lq)


def fd_grad(spec, theta, idx, z, order, group, step, coef, eps=1e-5):
    """Central differences of `loss` with respect to the parameters in idx."""
    g = np.zeros(len(idx))
    th = np.asarray(theta, dtype=np.float64).copy()
    for n, i in enumerate(idx):
        old = th[i]
        th[i] = old + eps
        lp = loss(spec, th, z, order, group, step, coef)
        th[i] = old - eps
        lm = loss(spec, th, z, order, group, step, coef)
        th[i] = old
        g[n] = (lp - lm) / (2.0 * eps)
    return g
