"""Host-side mirror of the reference's drivers (src/inchworm.jl) over the libqiw C ABI.

    inchworm(expansion, grid, orders, orders_bare, N_samples; n_pts_after_max, rand_params)
        -> (P_orders, P_orders_std), updates expansion.P          (inchworm!, :332-498)
    inchworm_step(expansion, n_w, orders, N_samples)            (inchworm_step, :123-203)
    inchworm_step_bare(expansion, orders_bare, N_samples)       (inchworm_step_bare, :228-304)
    correlator_2p(expansion, grid, orders, N_samples)           (correlator_2p, :1075-1084)
    correlator_2p(expansion, grid, orders, N_samples, RequestStdDev(); rand_params)
        -> (correlators, standard deviations)                    (correlator_2p, :918-1052)
    diff_inchworm(expansion, grid, orders, N_samples; rand_params)
        -> (Σ_orders, Σ_orders_std), updates expansion.P          (diff_inchworm!, :666-778)
    partition_function(expansion), density_matrix(expansion)    (src/ppgf.jl:611-617, :729-733)

All evaluation runs in libqiw on the GPU; there is no CPU fallback.
"""
import ctypes as C

import numpy as np

from . import abi
from .randomization import RandomizationParams, RequestStdDev, draw_points, draw_sequences
from .sobol import ScrambledSobolSeq

_ctx = None


def context(device=0):
    """The process-wide libqiw context (one GPU)."""
    global _ctx
    if _ctx is None:
        _ctx = abi.Context(device)
    return _ctx


def _stage(ctx, expansion, max_order, seq=None):
    ms = expansion.to_struct()
    ctx.check(ctx.lib.qiw_set_model(ctx.h, ms.ref()))
    D = max(2 * max_order, 1)
    seq = seq or ScrambledSobolSeq(D)
    if seq.D < D:
        raise ValueError("Sobol sequence has too few dimensions")
    m = np.ascontiguousarray(seq.m[:D])
    x0 = np.ascontiguousarray(seq.x0[:D])
    ctx.check(ctx.lib.qiw_set_sobol(ctx.h, D, abi.ptr(m, C.c_uint32), abi.ptr(x0, C.c_uint32)))
    return ms


def _set_P(ctx, expansion):
    P = np.ascontiguousarray(expansion.P, dtype=np.complex128)
    d0 = np.ascontiguousarray(expansion.P_der0, dtype=np.complex128)
    ctx.check(ctx.lib.qiw_set_ppgf(ctx.h, int(expansion.P_n_known), P.ctypes.data_as(abi._f64p),
                                   d0.ctypes.data_as(abi._f64p)))


def _get_P(ctx, expansion):
    P = np.zeros_like(expansion.P)
    d0 = np.zeros(expansion.bt_size, np.complex128)
    nk = C.c_int32()
    ctx.check(ctx.lib.qiw_get_ppgf(ctx.h, C.byref(nk), P.ctypes.data_as(abi._f64p), d0.ctypes.data_as(abi._f64p)))
    expansion.P, expansion.P_der0, expansion.P_n_known = P, d0, nk.value


def stats(ctx=None):
    ctx = ctx or context()
    st = abi.qiw_stats()
    ctx.check(ctx.lib.qiw_get_stats(ctx.h, C.byref(st)))
    return {f: getattr(st, f) for f, _ in abi.qiw_stats._fields_}


def _rand(rand_params):
    if rand_params is None:
        return None
    if rand_params.N_seqs <= 0:
        raise ValueError("rand_params.N_seqs must be > 0")  # @assert, src/inchworm.jl:344
    return rand_params.struct_args()


_CALLS = {"inchworm": 0, "step": 1, "step_bare": 2, "correlator": 3, "diff": 4}


def scramble_dims(run, call, ctx=None):
    """qiw_scramble_dims: the dimensions 2k of the sequences a call (kind "inchworm",
    "step", "step_bare", "correlator" or "diff") with the run struct consumes, in the
    reference's construction order."""
    ctx = ctx or context()
    n = ctx.lib.qiw_scramble_dims(ctx.h, run.ref(), _CALLS[call], 0, None)
    if n < 0:
        ctx.check(n)
    out = np.zeros(max(n, 1), np.int32)
    ctx.check(min(0, ctx.lib.qiw_scramble_dims(ctx.h, run.ref(), _CALLS[call], n, out.ctypes.data_as(abi._i32p))))
    return [int(v) for v in out[:n]]


def _stage_given(ctx, rand_params, run, call):
    """Stage the caller-drawn sequences of one call (qiw_set_scrambles) when rand_params
    asks for them: ScrambledSobolSeq(D, scramble_rng=rng) for each D qiw_scramble_dims
    lists."""
    if rand_params is None or not rand_params.caller_drawn:
        return
    dims = scramble_dims(run, call, ctx)
    stride = max(dims + [2])
    if rand_params.seq_type is not None:
        n = max(int(run.struct.N_samples), 1)
        set_points(draw_points(rand_params.rng, rand_params.seq_type, dims, stride, n), ctx)
        return
    m, x = draw_sequences(rand_params.rng, dims, stride)
    set_scrambles(m, x, ctx)


def set_scrambles(m, x=None, ctx=None):
    """qiw_set_scrambles: m [n][stride][32], x [n][stride] (None: zeros)."""
    ctx = ctx or context()
    m = np.ascontiguousarray(m, dtype=np.uint32)
    n, stride = (m.shape[0], m.shape[1]) if m.ndim == 3 else (0, 1)
    xp = None
    if x is not None:
        x = np.ascontiguousarray(x, dtype=np.uint32)
        xp = x.ctypes.data_as(abi._u32p)
    ctx.check(ctx.lib.qiw_set_scrambles(ctx.h, n, stride, m.ctypes.data_as(abi._u32p) if n else None, xp))


def set_points(x, ctx=None):
    """qiw_set_points: x [n][n_points][stride] in [0, 1] (the caller's point sequences of
    qiw_rand.scramble = 3)."""
    ctx = ctx or context()
    x = np.ascontiguousarray(x, dtype=np.float64)
    n, npts, stride = x.shape if x.ndim == 3 and x.shape[0] > 0 else (0, 1, 1)
    ctx.check(ctx.lib.qiw_set_points(ctx.h, n, npts, stride, x.ctypes.data_as(abi._f64p) if n else None))


def inchworm(expansion, grid, orders, orders_bare, N_samples, n_pts_after_max=None,
             precision=abi.QIW_PREC_FP64, seq=None, ctx=None, rand_params=None):
    """inchworm! (src/inchworm.jl:332-498): writes the bold propagator into expansion.P and
    returns the order-resolved contributions and their standard deviations as
    {order: array (n_tau, sum m^2)}.  rand_params: RandomizationParams (default: one
    unscrambled sequence; seq, if given, is the staged sequence)."""
    ctx = ctx or context()
    orders, orders_bare = list(orders), list(orders_bare)
    _stage(ctx, expansion, max(orders + orders_bare + [0]), seq)
    run = abi.RunStruct(orders, orders_bare, N_samples, n_pts_after_max or 0, precision, rand=_rand(rand_params))
    if rand_params is not None and rand_params.caller_drawn:
        _stage_given(ctx, rand_params, run, "inchworm")
    oall = sorted(set(orders) | set(orders_bare))
    nt, bt = grid.n_tau, expansion.bt_size
    Po = np.zeros((len(oall), nt, bt), np.complex128)
    Ps = np.zeros((len(oall), nt, bt), np.complex128)
    P = np.zeros((nt, bt), np.complex128)
    ctx.check(ctx.lib.qiw_inchworm(ctx.h, run.ref(), P.ctypes.data_as(abi._f64p), Po.ctypes.data_as(abi._f64p),
                                   Ps.ctypes.data_as(abi._f64p)))
    _get_P(ctx, expansion)
    return {o: Po[i] for i, o in enumerate(oall)}, {o: Ps[i] for i, o in enumerate(oall)}


def inchworm_step(expansion, n_w, orders, N_samples, n_pts_after_max=None, seq=None, ctx=None,
                  precision=abi.QIW_PREC_FP64, rand_params=None):
    """One regular step (src/inchworm.jl:123-203) with tau_i = grid[0], tau_w = grid[n_w],
    tau_f = grid[n_w + 1] on the current expansion.P.  Returns (value, per-order)."""
    ctx = ctx or context()
    orders = list(orders)
    _stage(ctx, expansion, max(orders + [0]), seq)
    _set_P(ctx, expansion)
    run = abi.RunStruct(orders, [], N_samples, n_pts_after_max or 0, precision, rand=_rand(rand_params))
    if rand_params is not None and rand_params.caller_drawn:
        _stage_given(ctx, rand_params, run, "step")
    val = np.zeros(expansion.bt_size, np.complex128)
    per = np.zeros((len(orders), expansion.bt_size), np.complex128)
    ctx.check(ctx.lib.qiw_inchworm_step(ctx.h, run.ref(), int(n_w), val.ctypes.data_as(abi._f64p),
                                        per.ctypes.data_as(abi._f64p)))
    return val, per


def inchworm_step_bare(expansion, orders_bare, N_samples, seq=None, ctx=None, rand_params=None):
    """The initial step (src/inchworm.jl:228-304) on [grid[0], grid[1]]."""
    ctx = ctx or context()
    orders_bare = list(orders_bare)
    _stage(ctx, expansion, max(orders_bare + [0]), seq)
    run = abi.RunStruct([], orders_bare, N_samples, rand=_rand(rand_params))
    if rand_params is not None and rand_params.caller_drawn:
        _stage_given(ctx, rand_params, run, "step_bare")
    val = np.zeros(expansion.bt_size, np.complex128)
    per = np.zeros((len(orders_bare), expansion.bt_size), np.complex128)
    ctx.check(ctx.lib.qiw_inchworm_step_bare(ctx.h, run.ref(), val.ctypes.data_as(abi._f64p),
                                             per.ctypes.data_as(abi._f64p)))
    return val, per


def diff_inchworm(expansion, grid, orders, N_samples, precision=abi.QIW_PREC_FP64, seq=None, ctx=None,
                  rand_params=None):
    """diff_inchworm! (src/inchworm.jl:666-778): the differential formulation.  Writes the
    propagator into expansion.P and returns the order-resolved self-energy and its
    standard deviations as {order: array (n_tau, sum m^2)}."""
    ctx = ctx or context()
    orders = list(orders)
    _stage(ctx, expansion, max(orders + [0]), seq)
    run = abi.RunStruct(orders, [], N_samples, 0, precision, rand=_rand(rand_params))
    if rand_params is not None and rand_params.caller_drawn:
        _stage_given(ctx, rand_params, run, "diff")
    nt, bt = grid.n_tau, expansion.bt_size
    So = np.zeros((len(orders), nt, bt), np.complex128)
    Ss = np.zeros_like(So)
    P = np.zeros((nt, bt), np.complex128)
    ctx.check(ctx.lib.qiw_diff_inchworm(ctx.h, run.ref(), P.ctypes.data_as(abi._f64p), So.ctypes.data_as(abi._f64p),
                                        Ss.ctypes.data_as(abi._f64p)))
    _get_P(ctx, expansion)
    return {o: So[i] for i, o in enumerate(orders)}, {o: Ss[i] for i, o in enumerate(orders)}


def correlator_2p(expansion, grid, orders, N_samples, *flag, seq=None, ctx=None, precision=abi.QIW_PREC_FP64,
                  rand_params=None):
    """correlator_2p (src/inchworm.jl:1075-1084): one array of n_tau complex values
    <A(tau) B(0)> per (A, B) pair of expansion.corr_operators.  With RequestStdDev() as
    fifth argument (src/inchworm.jl:918-1052): (values, standard deviations)."""
    if len(flag) > 1 or (flag and not isinstance(flag[0], RequestStdDev)):
        raise TypeError("the only flag argument of correlator_2p is RequestStdDev()")
    ctx = ctx or context()
    orders = list(orders)
    _stage(ctx, expansion, max(orders + [0]), seq)
    _set_P(ctx, expansion)
    run = abi.RunStruct(orders, [], N_samples, 0, precision, rand=_rand(rand_params))
    if rand_params is not None and rand_params.caller_drawn:
        _stage_given(ctx, rand_params, run, "correlator")
    out = np.zeros((len(expansion.corr_operators), grid.n_tau), np.complex128)
    if not flag:
        ctx.check(ctx.lib.qiw_correlator_2p(ctx.h, run.ref(), out.ctypes.data_as(abi._f64p)))
        return [out[i] for i in range(out.shape[0])]
    std = np.zeros_like(out)
    ctx.check(ctx.lib.qiw_correlator_2p_std(ctx.h, run.ref(), out.ctypes.data_as(abi._f64p),
                                            std.ctypes.data_as(abi._f64p)))
    return [out[i] for i in range(out.shape[0])], [std[i] for i in range(std.shape[0])]


def partition_function(expansion):
    """Z = sum_s tr i P_s(beta) (src/ppgf.jl:611-617)."""
    row = expansion.P[-1]
    return sum(1j * np.trace(expansion.block(row, s)) for s in range(len(expansion.ed.dims)))


def normalize(expansion, beta):
    """normalize!(P, beta) (src/ppgf.jl:629-636): rescale so that Z = 1; returns lambda."""
    Z = partition_function(expansion)
    lam = np.log(Z) / beta
    expansion.P = expansion.P * np.exp(-expansion.grid.tau * lam)[:, None]
    return lam


def density_matrix(expansion):
    """rho = i P(beta) of a normalized propagator (src/ppgf.jl:729-733)."""
    row = expansion.P[-1]
    return [1j * expansion.block(row, s) for s in range(len(expansion.ed.dims))]


def _tuning_dict(t):
    return {f: getattr(t, f) for f, _ in abi.qiw_tuning._fields_ if f != "pad_"}


def set_graphs(on=True, ctx=None):
    """Replay the bold-step loop from a hipGraph (default on; results are identical)."""
    ctx = ctx or context()
    ctx.check(ctx.lib.qiw_set_graphs(ctx.h, 1 if on else 0))


def set_tuning(chunk_words=0, chunk_cost=0, samples_per_task=0, ctx=None):
    """Fix the work mapping of the evaluators (qiw_set_tuning); 0 keeps a default."""
    ctx = ctx or context()
    t = abi.qiw_tuning(chunk_words=int(chunk_words), chunk_cost=int(chunk_cost),
                       samples_per_task=int(samples_per_task))
    ctx.check(ctx.lib.qiw_set_tuning(ctx.h, C.byref(t)))


def get_tuning(ctx=None):
    ctx = ctx or context()
    t = abi.qiw_tuning()
    ctx.check(ctx.lib.qiw_get_tuning(ctx.h, C.byref(t)))
    return _tuning_dict(t)


def autotune(expansion, orders, N_samples, precision=abi.QIW_PREC_FP64, seq=None, ctx=None):
    """Time candidate work mappings on one bold step of the expansion at its current P
    (qiw_autotune) and keep the fastest for the following calls.  Returns the setting."""
    ctx = ctx or context()
    orders = list(orders)
    _stage(ctx, expansion, max(orders + [0]), seq)
    _set_P(ctx, expansion)
    run = abi.RunStruct(orders, [], N_samples, 0, precision)
    t = abi.qiw_tuning()
    ctx.check(ctx.lib.qiw_autotune(ctx.h, run.ref(), C.byref(t)))
    return _tuning_dict(t)
