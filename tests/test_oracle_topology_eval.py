"""Oracle TopologyEvaluator against the reference's golden values of test/topology_eval.h5
(test/topology_eval.jl:37-155): order-3 inchworm topologies with n_pts_after = 1 of the
single-state model (beta = 10, mu = 0.1, DeltaDOS(0.1), V = -0.1, n_tau = 30 on the
imaginary branch), evaluated at 100 given time configurations with the atomic P and a
grid-stored Delta, both read by Keldysh's grid interpolation.  The reference checks
isapprox(values, golden, rtol = 1e-10).

This pins the grid-cell interpolation of grid containers, for the bold P and for a
grid-stored pair propagator alike: linear interpolation in the time difference misses the
golden values by 3e-4 (P and Delta) or 1.4e-1 (Delta only, by the reversed-pair end
point)."""
import ctypes as C
import json
import os

import numpy as np

import oracle_lib as O
from qiw import diagrammatics, model

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "topology_eval_h5.json")))


def golden_inputs():
    x1 = np.array(FX["/x1_list"]["data"])                    # [100]
    xs = np.array(FX["/xs_list"]["data"]).reshape(100, 5)    # h5 is row-major: [sample][k]
    vals = np.array([complex(a, b) for a, b in FX["/values"]["data"]]).reshape(2, 100)
    return x1, xs, vals


def build_model():
    beta, mu, eps, V, n_tau = 10.0, 0.1, 0.1, -0.1, 30
    fock = model.FockSpace(["0"])
    grid = model.ImaginaryTimeGrid(beta, n_tau)
    delta = V ** 2 * model.delta_dos_gf(eps, beta, grid)
    # Delta_rev(tau) = Delta[0, tau] (test/topology_eval.jl:63): the lesser value
    # xi Delta(beta - tau) for tau > 0 and the greater Delta(0) at equal times
    rev = -delta.values[::-1].copy()
    rev[0] = delta.values[0]
    # KeldyshED puts the occupied (lower-energy) sector first for H = -mu n
    rank = {1: 0, 0: 1}
    ed = model.EDCore(-mu * fock.n("0"), fock, lambda b: rank[fock.occupation(b, "0")])
    pairs = [model.InteractionPair(fock.c_dag("0"), fock.c("0"), delta),
             model.InteractionPair(fock.c("0"), fock.c_dag("0"), model.ScalarGF(rev))]
    return model.Expansion(ed, grid, pairs), beta


def evaluate(exp, beta, x1, xs):
    order = 3
    tops = diagrammatics.get_topologies_at_order(order, 1)
    assert len(tops) == 4
    pairs = np.array([[a for p in t.pairs for a in p] for t in tops], np.int32)
    par = np.array([t.parity for t in tops], np.int32)
    x_w, x_f = 6 / 29, 7 / 29  # tau_grid[7], tau_grid[8] as branch fractions
    ms = exp.to_struct()
    out = np.zeros((len(x1), exp.bt_size), np.complex128)
    for s in range(len(x1)):
        t = np.r_[x_w + (x_f - x_w) * x1[s], x_w * xs[s]] * beta  # descending, var_time_pos order
        O.check(O.lib().orc_eval_topologies(ms.ref(), 0, order, 2 * order - 1, 0.0, x_w * beta, x_f * beta,
                                            len(tops), pairs.ctypes.data_as(C.POINTER(C.c_int32)),
                                            par.ctypes.data_as(C.POINTER(C.c_int32)),
                                            t.ctypes.data_as(C.POINTER(C.c_double)),
                                            out[s].ctypes.data_as(C.POINTER(C.c_double))))
    return out.T


def test_topology_eval_golden():
    x1, xs, vals = golden_inputs()
    exp, beta = build_model()
    got = evaluate(exp, beta, x1, xs)
    rel = np.linalg.norm(got - vals) / np.linalg.norm(vals)
    assert rel <= 1e-10          # the reference's own tolerance
    assert rel <= 1e-14          # measured 5e-16


def test_topology_eval_fixture_discriminates_the_end_point():
    # the reversed pair's value at tau = 0 enters every short arc: the lesser convention
    # -Delta(beta) instead of the greater Delta(0) misses the golden values by far
    x1, xs, vals = golden_inputs()
    exp, beta = build_model()
    exp.pairs[1].propagator.values[0] = -exp.pairs[0].propagator.values[-1]
    got = evaluate(exp, beta, x1, xs)
    assert np.linalg.norm(got - vals) / np.linalg.norm(vals) > 1e-2
