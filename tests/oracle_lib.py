"""ctypes binding of the CPU float64 oracle (oracle/build/liborc.so).

Test infrastructure: the oracle is the checker the GPU path is compared against and is
never part of the product path."""
import ctypes as C
import os
import subprocess

import numpy as np

from qiw import abi

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORC_PATH = os.path.join(REPO, "oracle", "build", "liborc.so")
_orc = None


def lib():
    global _orc
    if _orc is None:
        if not os.path.exists(ORC_PATH):
            subprocess.run(["make", "-C", os.path.join(REPO, "oracle")], check=True,
                           capture_output=True)
        L = C.CDLL(ORC_PATH)
        f64p, u32p, i32p = C.POINTER(C.c_double), C.POINTER(C.c_uint32), C.POINTER(C.c_int32)
        L.orc_last_error.restype = C.c_char_p
        L.orc_sobol_direction_numbers.argtypes = [C.c_int, u32p, u32p, C.c_int, C.POINTER(C.c_uint8),
                                                  C.POINTER(C.c_uint8), u32p, u32p]
        L.orc_sobol_points.argtypes = [C.c_int, u32p, u32p, C.c_int64, C.c_int, C.c_int64, f64p]
        L.orc_simplex_points.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_double, f64p,
                                         C.c_int64, f64p, f64p]
        L.orc_topologies.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, i32p, i32p]
        L.orc_inchworm.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p, C.c_int,
                                   f64p, f64p, f64p]
        L.orc_inchworm_step.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p,
                                        C.c_int, f64p, f64p, C.c_int, f64p, f64p]
        L.orc_inchworm_step_bare.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p,
                                             f64p, f64p]
        L.orc_correlator_2p.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p,
                                        C.c_int, f64p, f64p, C.c_int, C.c_int, f64p]
        L.orc_correlator_2p_std.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p,
                                            C.c_int, f64p, f64p, C.c_int, C.c_int, f64p, f64p]
        L.orc_diff_inchworm.argtypes = [C.POINTER(abi.qiw_model), C.POINTER(abi.qiw_run), u32p, u32p, C.c_int,
                                        f64p, f64p, f64p]
        L.orc_rq_scramble.argtypes = [C.c_uint64, C.c_uint64, C.c_int, u32p, u32p, u32p]
        L.orc_dlr_eval.argtypes = [C.c_int, f64p, f64p, C.c_int, C.c_double, C.c_double, C.c_double, f64p]
        L.orc_set_scrambles.argtypes = [C.c_int64, C.c_int, u32p, u32p]
        L.orc_set_points.argtypes = [C.c_int64, C.c_int64, C.c_int, f64p]
        L.orc_scrambles_used.restype = C.c_int64
        L.orc_scrambles_log.argtypes = [C.c_int64, C.POINTER(C.c_int64)]
        L.orc_scrambles_log.restype = C.c_int64
        L.orc_rq_stream.argtypes = [C.c_int, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.orc_rq_stream.restype = C.c_uint64
        L.orc_eval_topologies.argtypes = [C.POINTER(abi.qiw_model), C.c_int, C.c_int, C.c_int, C.c_double,
                                          C.c_double, C.c_double, C.c_int, i32p, i32p, f64p, f64p]
        L.orc_gf_eval.argtypes = [C.c_int, C.c_int, C.c_double, f64p, C.c_int, f64p, C.c_int, C.c_int,
                                  f64p, f64p, f64p]
        L.orc_get_counters.argtypes = [C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_double)]
        _orc = L
    return _orc


def _p(a, t):
    return None if a is None else a.ctypes.data_as(C.POINTER(t))


def check(rc):
    if rc != 0:
        raise abi.QiwError(rc, lib().orc_last_error().decode())


def set_threads(n):
    lib().orc_set_threads(int(n))


def counters():
    a, b, c = C.c_int64(), C.c_int64(), C.c_double()
    lib().orc_get_counters(C.byref(a), C.byref(b), C.byref(c))
    return a.value, b.value, c.value


def reset_counters():
    lib().orc_reset_counters()


def sobol_tables(max_dim):
    from qiw.sobol import direction_numbers
    m = np.ascontiguousarray(direction_numbers(max_dim))
    return m, np.zeros(max_dim, dtype=np.uint32)


def inchworm(exp, orders, orders_bare, N, n_pts_after_max=0, n_steps_max=-1, sobol=None, rand=None):
    ms = exp.to_struct()
    run = abi.RunStruct(orders, orders_bare, N, n_pts_after_max, rand=rand)
    max_order = max(list(orders) + list(orders_bare) + [1])
    m, x0 = sobol if sobol is not None else sobol_tables(2 * max_order)
    n_tau, bt = exp.grid.n_tau, exp.bt_size
    oall = sorted(set(orders) | set(orders_bare))
    P = np.zeros((n_tau, bt), np.complex128)
    Po = np.zeros((len(oall), n_tau, bt), np.complex128)
    Ps = np.zeros((len(oall), n_tau, bt), np.complex128)
    check(lib().orc_inchworm(ms.ref(), run.ref(), _p(m, C.c_uint32), _p(x0, C.c_uint32), int(n_steps_max),
                             _p(P, C.c_double), _p(Po, C.c_double), _p(Ps, C.c_double)))
    return P, {o: Po[i] for i, o in enumerate(oall)}, {o: Ps[i] for i, o in enumerate(oall)}


def inchworm_step(exp, orders, N, jw, P=None, n_known=None, der0=None, sobol=None, n_pts_after_max=0, rand=None):
    ms = exp.to_struct()
    run = abi.RunStruct(orders, [], N, n_pts_after_max, rand=rand)
    m, x0 = sobol if sobol is not None else sobol_tables(2 * max(list(orders) + [1]))
    P = np.ascontiguousarray(exp.P if P is None else P, dtype=np.complex128)
    der0 = np.ascontiguousarray(exp.P_der0 if der0 is None else der0, dtype=np.complex128)
    n_known = exp.P_n_known if n_known is None else n_known
    val = np.zeros(exp.bt_size, np.complex128)
    per = np.zeros((len(orders), exp.bt_size), np.complex128)
    check(lib().orc_inchworm_step(ms.ref(), run.ref(), _p(m, C.c_uint32), _p(x0, C.c_uint32), int(n_known),
                                  _p(P, C.c_double), _p(der0, C.c_double), int(jw), _p(val, C.c_double),
                                  _p(per, C.c_double)))
    return val, per


def inchworm_step_bare(exp, orders_bare, N, sobol=None, rand=None):
    ms = exp.to_struct()
    run = abi.RunStruct([], orders_bare, N, rand=rand)
    m, x0 = sobol if sobol is not None else sobol_tables(2 * max(list(orders_bare) + [1]))
    val = np.zeros(exp.bt_size, np.complex128)
    per = np.zeros((len(orders_bare), exp.bt_size), np.complex128)
    check(lib().orc_inchworm_step_bare(ms.ref(), run.ref(), _p(m, C.c_uint32), _p(x0, C.c_uint32),
                                       _p(val, C.c_double), _p(per, C.c_double)))
    return val, per


def correlator_2p(exp, orders, N, P=None, n_known=None, der0=None, tau_first=0, tau_count=-1, sobol=None,
                  rand=None, with_std=False):
    ms = exp.to_struct()
    run = abi.RunStruct(orders, [], N, rand=rand)
    m, x0 = sobol if sobol is not None else sobol_tables(2 * max(list(orders) + [1]))
    P = np.ascontiguousarray(exp.P if P is None else P, dtype=np.complex128)
    der0 = np.ascontiguousarray(exp.P_der0 if der0 is None else der0, dtype=np.complex128)
    n_known = exp.P_n_known if n_known is None else n_known
    out = np.zeros((len(exp.corr_operators), exp.grid.n_tau), np.complex128)
    std = np.zeros_like(out)
    check(lib().orc_correlator_2p_std(ms.ref(), run.ref(), _p(m, C.c_uint32), _p(x0, C.c_uint32), int(n_known),
                                      _p(P, C.c_double), _p(der0, C.c_double), int(tau_first), int(tau_count),
                                      _p(out, C.c_double), _p(std, C.c_double)))
    return (out, std) if with_std else out


def rq_scramble(seed, stream, m):
    """Direction numbers m [D][32] scrambled with the bits of (seed, stream), and the shift."""
    m = np.ascontiguousarray(m, dtype=np.uint32)
    D = m.shape[0]
    mo = np.zeros_like(m)
    xo = np.zeros(D, np.uint32)
    lib().orc_rq_scramble(int(seed), int(stream), D, _p(m, C.c_uint32), _p(mo, C.c_uint32), _p(xo, C.c_uint32))
    return mo, xo


def dlr_eval(g, t1, t2):
    """The oracle's lookup of a DLR function g (model.DLRGF) at (t1, t2)."""
    w = np.ascontiguousarray(g.omega, dtype=np.float64)
    c = np.ascontiguousarray(g.coef, dtype=np.complex128)
    out = np.zeros(2)
    lib().orc_dlr_eval(len(w), _p(w, C.c_double), c.ctypes.data_as(C.POINTER(C.c_double)), int(g.xi), float(g.beta),
                       float(t1), float(t2), _p(out, C.c_double))
    return complex(out[0], out[1])


def set_scrambles(m, x):
    """Stage caller-drawn sequences (m [n][stride][32], x [n][stride]) for scramble = 2;
    the oracle consumes them in the order its loops (the reference's) construct sequences."""
    m = np.ascontiguousarray(m, dtype=np.uint32)
    x = np.ascontiguousarray(x, dtype=np.uint32)
    lib().orc_set_scrambles(int(m.shape[0]), int(m.shape[1]), _p(m, C.c_uint32), _p(x, C.c_uint32))


def set_points(x):
    """Stage caller-supplied point sequences (x [n][n_points][stride]) for scramble = 3,
    consumed as under scramble = 2."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    lib().orc_set_points(int(x.shape[0]), int(x.shape[1]), int(x.shape[2]), _p(x, C.c_double))


def scrambles_used():
    return int(lib().orc_scrambles_used())


def scrambles_log():
    """(domain, key, order, n_pts_after, s) of every sequence consumed since set_scrambles."""
    n = int(lib().orc_scrambles_log(0, None))
    out = np.zeros((max(n, 1), 5), np.int64)
    lib().orc_scrambles_log(n, _p(out, C.c_int64))
    return [tuple(int(v) for v in row) for row in out[:n]]


def rq_stream(domain, key, order, n_after, seq):
    return int(lib().orc_rq_stream(int(domain), int(key), int(order), int(n_after), int(seq)))


def diff_inchworm(exp, orders, N, n_steps_max=-1, sobol=None, rand=None):
    """diff_inchworm! on the oracle: (P table, {order: Σ contributions}, {order: std})."""
    ms = exp.to_struct()
    run = abi.RunStruct(orders, [], N, rand=rand)
    m, x0 = sobol if sobol is not None else sobol_tables(2 * max(list(orders) + [1]))
    n_tau, bt = exp.grid.n_tau, exp.bt_size
    P = np.zeros((n_tau, bt), np.complex128)
    So = np.zeros((len(list(orders)), n_tau, bt), np.complex128)
    Ss = np.zeros_like(So)
    check(lib().orc_diff_inchworm(ms.ref(), run.ref(), _p(m, C.c_uint32), _p(x0, C.c_uint32), int(n_steps_max),
                                  _p(P, C.c_double), _p(So, C.c_double), _p(Ss, C.c_double)))
    return P, {o: So[i] for i, o in enumerate(orders)}, {o: Ss[i] for i, o in enumerate(orders)}
