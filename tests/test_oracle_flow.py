"""The float64 flow oracle against what the mathematics fixes (CPU): the identity map at a
zero output layer, inverse(forward(z)) = z with opposite log-Jacobians, log dy/dx against
a numerical derivative, volume preservation E_x[|dz/dx|] = 1, and monotonicity."""
import numpy as np

from flow_ref import flow_oracle as F

SPEC = F.Spec()


def batch(n, rng):
    z = rng.random((n, 8))
    order = rng.integers(1, 5, n)
    group = np.array([F.N_GROUPS - 1 if o == 4 else (o - 1) ** 2 for o in order])
    step = rng.random(n)
    return z, order, group, step


def zero_output(theta):
    th = theta.copy()
    o = 0
    for _ in range(SPEC.L):
        for name, n in SPEC.layer_sizes():
            if name in ("W3", "b3"):
                th[o:o + n] = 0.0
            o += n
    return th


def used(order):
    return np.arange(8)[None, :] < 2 * np.asarray(order)[:, None]


def test_zero_output_layer_is_identity():
    rng = np.random.default_rng(0)
    th = zero_output(rng.normal(0, 0.3, SPEC.n_params()))
    z, order, group, step = batch(500, rng)
    x, lq = F.forward(SPEC, th, z, order, group, step)
    assert np.abs(x - z).max() < 1e-14
    assert np.abs(lq).max() < 1e-13


def test_inverse_roundtrip_and_logdet():
    rng = np.random.default_rng(1)
    th = rng.normal(0, 0.1, SPEC.n_params())
    z, order, group, step = batch(2000, rng)
    x, lq = F.forward(SPEC, th, z, order, group, step)
    z2, lj = F.inverse(SPEC, th, x, order, group, step)
    m = used(order)
    assert np.abs((z2 - z) * m).max() < 1e-11
    assert np.abs(lq + lj).max() < 1e-10
    assert (x[m] >= 0).all() and (x[m] <= 1).all()


def test_spline_derivative_and_monotone():
    rng = np.random.default_rng(2)
    raw = rng.normal(0, 1.5, (200, 3 * SPEC.K + 1))
    xs = np.sort(rng.random(200))
    y, ld = F.rqs_forward(xs, raw, SPEC.K)
    e = 1e-6
    yp, _ = F.rqs_forward(xs + e, raw, SPEC.K)
    ym, _ = F.rqs_forward(xs - e, raw, SPEC.K)
    assert np.abs(np.log((yp - ym) / (2 * e)) - ld).max() < 1e-5
    # one spline on a fine grid is increasing from 0 to 1
    g = np.linspace(0, 1, 4001)
    yy, _ = F.rqs_forward(g, np.broadcast_to(raw[0], (g.size, raw.shape[1])), SPEC.K)
    assert abs(yy[0]) < 1e-15 and abs(yy[-1] - 1) < 1e-12 and (np.diff(yy) > 0).all()


def test_warp_preserves_volume():
    # the warped points cover the cube once: E_x[|dz/dx|] = vol = 1
    rng = np.random.default_rng(3)
    th = rng.normal(0, 0.05, SPEC.n_params())
    n = 200000
    x = rng.random((n, 8))
    _, lj = F.inverse(SPEC, th, x, np.full(n, 2), np.full(n, 3), np.full(n, 0.4))
    w = np.exp(lj)
    assert abs(w.mean() - 1.0) < 5 * w.std() / np.sqrt(n)


def test_fd_gradient_of_loss_is_consistent():
    # central differences agree with a one-sided difference at smaller step (smoothness)
    rng = np.random.default_rng(4)
    th = rng.normal(0, 0.1, SPEC.n_params())
    z, order, group, step = batch(64, rng)
    coef = rng.random(64) * 2
    idx = rng.choice(SPEC.n_params(), 12, replace=False)
    g1 = F.fd_grad(SPEC, th, idx, z, order, group, step, coef, eps=1e-5)
    g2 = F.fd_grad(SPEC, th, idx, z, order, group, step, coef, eps=1e-4)
    assert np.allclose(g1, g2, rtol=1e-4, atol=1e-9)
