"""GPU kernels of the learned warp against the float64 NumPy flow oracle, and the warped
inchworm estimator against the plain one, through the C ABI."""
import numpy as np
import pytest

import qiw
from flow_ref import flow_oracle as F
from qiw import model

pytestmark = pytest.mark.gpu
SPEC = F.Spec()


def batch(n, rng):
    order = rng.integers(1, 5, n)
    n_after = np.array([rng.integers(1, 2 * o) for o in order])
    group = np.array([qiw.group_index(int(o), int(a)) for o, a in zip(order, n_after)])
    z = rng.random((n, 8))
    z[np.arange(8)[None, :] >= 2 * order[:, None]] = 0.0
    return z, order, group, rng.random(n)


def random_params(sd, seed):
    qiw.warp_init(seed=seed)
    rng = np.random.default_rng(seed)
    th = rng.normal(0, sd, SPEC.n_params()).astype(np.float32)
    qiw.warp_params(set=th)
    return th.astype(np.float64)


def test_param_count_and_identity_init():
    qiw.warp_init(seed=3)
    th = qiw.warp_params()
    assert th.size == SPEC.n_params()
    rng = np.random.default_rng(0)
    z, order, group, step = batch(1000, rng)
    x, lq = qiw.warp_apply(z, order, group, step)
    assert np.abs(x - z).max() < 1e-14 and np.abs(lq).max() < 1e-13


@pytest.mark.parametrize("inverse", [False, True])
def test_apply_matches_oracle(inverse):
    th = random_params(0.1, 5)
    rng = np.random.default_rng(1)
    z, order, group, step = batch(3000, rng)
    x_g, l_g = qiw.warp_apply(z, order, group, step, inverse=inverse)
    f = F.inverse if inverse else F.forward
    x_c, l_c = f(SPEC, th, z, order, group, step.astype(np.float32).astype(np.float64))
    m = np.arange(8)[None, :] < 2 * order[:, None]
    # float32 conditioner on the device, float64 splines on both sides
    assert np.abs((x_g - x_c) * m).max() < 1e-4
    assert np.abs(l_g - l_c).max() < 1e-3


def test_device_roundtrip():
    random_params(0.1, 7)
    rng = np.random.default_rng(2)
    z, order, group, step = batch(4000, rng)
    x, lq = qiw.warp_apply(z, order, group, step)
    z2, lj = qiw.warp_apply(x, order, group, step, inverse=True)
    assert np.abs(z2 - z).max() < 1e-10
    assert np.abs(lq + lj).max() < 1e-9


def test_gradient_matches_oracle_finite_differences():
    th = random_params(0.1, 11)
    rng = np.random.default_rng(3)
    z, order, group, step = batch(256, rng)
    coef = rng.random(256) * 2.0
    step32 = step.astype(np.float32).astype(np.float64)
    loss_g, grad_g = qiw.warp_loss_grad(z, order, group, step, coef)
    assert abs(loss_g - F.loss(SPEC, th, z, order, group, step32, coef)) < 1e-4 * (1 + abs(loss_g))
    # parameters of every block of every layer, and used embedding rows
    idx, o = [], 0
    for _ in range(SPEC.L):
        for name, n in SPEC.layer_sizes():
            idx += list(o + rng.choice(n, 4, replace=False))
            o += n
    idx += [o + int(group[0]) * F.N_EMBED + q for q in range(4)]
    g_fd = F.fd_grad(SPEC, th, idx, z, order, group, step32, coef, eps=1e-4)
    g_dev = grad_g[idx].astype(np.float64)
    scale = np.abs(g_fd).max()
    assert np.abs(g_dev - g_fd).max() < 2e-3 * scale + 1e-6


def kanamori2(n_tau=8):
    return model.kanamori(2, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5)


def test_identity_warp_reproduces_plain_inchworm():
    k1, k2 = kanamori2(), kanamori2()
    qiw.warp_init(seed=1)
    qiw.warp_enable(False)
    qiw.inchworm(k1, k1.grid, range(0, 3), range(0, 3), 2 ** 8)
    qiw.warp_enable(True)
    try:
        qiw.inchworm(k2, k2.grid, range(0, 3), range(0, 3), 2 ** 8)
    finally:
        qiw.warp_enable(False)
    assert np.abs(k2.P - k1.P).max() < 1e-12 * np.abs(k1.P).max()


def test_training_lowers_loss_and_keeps_estimate():
    k = kanamori2()
    qiw.inchworm(k, k.grid, range(0, 3), range(0, 3), 2 ** 12)
    P_plain = k.P.copy()
    qiw.warp_init(iters=300, batch=4096, n_train_steps=3, n_train=2 ** 11, lr=3e-3, seed=2)
    hist = qiw.warp_train(k, range(0, 3))
    assert np.all(np.isfinite(hist))
    assert hist[-20:].mean() < hist[:20].mean()
    kw = kanamori2()
    qiw.warp_enable(True)
    try:
        qiw.inchworm(kw, kw.grid, range(0, 3), range(0, 3), 2 ** 12)
    finally:
        qiw.warp_enable(False)
    # unbiased: same propagator within the qMC error of either estimate
    assert np.abs(kw.P - P_plain).max() < 1e-3 * np.abs(P_plain).max()


def test_trained_warp_gains_over_plain():
    """What the trained network gains: at 2^10 points per group, the warped estimate of two
    bold steps of the two-orbital model is closer to a 2^18-point plain reference than the plain
    estimate at the same cost (measured gain about 10x in error)."""
    k = kanamori2(n_tau=20)
    orders = [0, 1, 2]
    qiw.inchworm(k, k.grid, orders, orders, 2 ** 12)
    P = k.P.copy()
    qiw.warp_init(iters=3000, batch=8192, n_train_steps=8, n_train=2 ** 14, lr=2e-3, seed=1)
    qiw.warp_train(k, orders)
    e_plain, e_warp = [], []
    for jw in (5, 15):
        k.P = P.copy()
        ref, _ = qiw.inchworm_step(k, jw, orders, 2 ** 18)
        v_plain, _ = qiw.inchworm_step(k, jw, orders, 2 ** 10)
        qiw.warp_enable(True)
        try:
            v_warp, _ = qiw.inchworm_step(k, jw, orders, 2 ** 10)
        finally:
            qiw.warp_enable(False)
        e_plain.append(np.linalg.norm(v_plain - ref) / np.linalg.norm(ref))
        e_warp.append(np.linalg.norm(v_warp - ref) / np.linalg.norm(ref))
    print(f"rel. error at 2^10: plain {np.mean(e_plain):.2e}, warped {np.mean(e_warp):.2e}")
    assert np.mean(e_warp) < 0.5 * np.mean(e_plain)
