"""The learned warp of the unit cube over the libqiw C ABI (qiw_warp_* in include/qiw.h).

    warp_init(...)                      a fresh flow (identity map until trained)
    warp_params(set=None)               float32 parameters (copy), optionally replaced
    warp_apply(pts, order, group, step, inverse)
                                        inverse=False: z -> x with log q(z);
                                        inverse=True:  x -> z with log |dz/dx|
    warp_loss_grad(z, order, group, step, coef)
                                        -mean(coef * log q(z)) and its gradient
    warp_train(expansion, orders, N_train, ...)
                                        fit the flow to |integrand| of the plain method at
                                        the expansion's current bold propagator
    warp_enable(on)                     use warped points in inchworm() bold steps

The reference has no such component; it integrates with plain quasi Monte Carlo
(src/qmc_integrate.jl:497-507).  Everything runs in libqiw on the GPU.
"""
import ctypes as C

import numpy as np

from . import abi
from .inchworm import _set_P, _stage, context

D = 8
_last_cfg = {}


def group_index(order, n_after):
    """Embedding row of the bold group (order k <= 4, n_after points after tau_w)."""
    return (order - 1) ** 2 + n_after - 1


def warp_init(n_layers=4, n_bins=8, hidden=64, iters=2000, batch=8192, n_train_steps=8, n_train=2 ** 12,
              lr=2e-3, seed=1, ctx=None):
    ctx = ctx or context()
    cfg = abi.qiw_warp_config(n_layers=n_layers, n_bins=n_bins, hidden=hidden, iters=iters, batch=batch,
                              n_train_steps=n_train_steps, n_train=int(n_train), lr=float(lr), seed=int(seed))
    ctx.check(ctx.lib.qiw_warp_init(ctx.h, C.byref(cfg)))
    _last_cfg[id(ctx)] = cfg
    return cfg


def warp_params(set=None, ctx=None):
    ctx = ctx or context()
    n = C.c_int64()
    ctx.check(ctx.lib.qiw_warp_params(ctx.h, C.byref(n), None, None))
    out = np.zeros(n.value, np.float32)
    src = None if set is None else np.ascontiguousarray(set, dtype=np.float32)
    if src is not None and src.size != n.value:
        raise ValueError(f"expected {n.value} parameters")
    ctx.check(ctx.lib.qiw_warp_params(ctx.h, None, abi.ptr(out, C.c_float),
                                      None if src is None else abi.ptr(src, C.c_float)))
    return out


def _batch(pts, order, group, step):
    pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, D)
    n = pts.shape[0]
    order = np.ascontiguousarray(np.broadcast_to(order, (n,)), dtype=np.int32)
    group = np.ascontiguousarray(np.broadcast_to(group, (n,)), dtype=np.int32)
    step = np.ascontiguousarray(np.broadcast_to(step, (n,)), dtype=np.float64)
    return pts, order, group, step


def warp_apply(pts, order, group, step, inverse=False, ctx=None):
    ctx = ctx or context()
    pts, order, group, step = _batch(pts, order, group, step)
    n = pts.shape[0]
    out = np.zeros_like(pts)
    ld = np.zeros(n)
    ctx.check(ctx.lib.qiw_warp_apply(ctx.h, int(bool(inverse)), n, abi.ptr(pts, C.c_double),
                                     abi.ptr(order, C.c_int32), abi.ptr(group, C.c_int32),
                                     abi.ptr(step, C.c_double), abi.ptr(out, C.c_double), abi.ptr(ld, C.c_double)))
    return out, ld


def warp_loss_grad(z, order, group, step, coef, ctx=None):
    ctx = ctx or context()
    z, order, group, step = _batch(z, order, group, step)
    coef = np.ascontiguousarray(np.broadcast_to(coef, (z.shape[0],)), dtype=np.float64)
    n = C.c_int64()
    ctx.check(ctx.lib.qiw_warp_params(ctx.h, C.byref(n), None, None))
    grad = np.zeros(n.value, np.float32)
    loss = C.c_double()
    ctx.check(ctx.lib.qiw_warp_loss_grad(ctx.h, z.shape[0], abi.ptr(z, C.c_double), abi.ptr(order, C.c_int32),
                                         abi.ptr(group, C.c_int32), abi.ptr(step, C.c_double),
                                         abi.ptr(coef, C.c_double), C.byref(loss), abi.ptr(grad, C.c_float)))
    return loss.value, grad


def warp_train(expansion, orders, ctx=None, seq=None):
    """Fit the initialised flow (warp_init) to |integrand| of every bold group of `orders`,
    sampled by the plain method at the expansion's current propagator.  Returns the loss of
    every Adam iteration."""
    ctx = ctx or context()
    orders = list(orders)
    _stage(ctx, expansion, max(orders + [0]), seq)
    _set_P(ctx, expansion)
    run = abi.RunStruct(orders, [], 0)
    iters = _last_cfg[id(ctx)].iters
    hist = np.zeros(max(int(iters), 1))
    ctx.check(ctx.lib.qiw_warp_train(ctx.h, run.ref(), abi.ptr(hist, C.c_double)))
    return hist[:iters]


def warp_enable(on=True, ctx=None):
    ctx = ctx or context()
    ctx.check(ctx.lib.qiw_warp_enable(ctx.h, int(bool(on))))
