"""Oracle inchworm!/correlator_2p against the reference's golden vectors
(test/inchworm.jl:180-212 via test/inchworm.h5) and closed forms.

The reference compares with isapprox (rtol = sqrt(eps), norm-wise per array); the oracle
reproduces both P sectors and g to round-off (about 1e-15).  The bold P lookups follow
the grid-cell interpolation of Keldysh's interpolate!: linear interpolation in the time
difference alone misses /inchworm/P by 2e-8 and /inchworm/g by 7e-7."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from qiw import abi, model

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "inchworm_h5.json")))


def fx(name):
    return np.array([complex(a, b) for a, b in FX[name]["data"]])


P1, P2, G = fx("/inchworm/P/1"), fx("/inchworm/P/2"), fx("/inchworm/g")


def golden_model():
    # KeldyshED orders the single-mode sectors occupied-first here: /inchworm/P/1 is the
    # sector with the lower atomic energy (its iP(beta) is the normalized maximum, -1).
    return model.single_state_golden_model(("occupied", "empty"))


def test_bare_step_exact():
    exp = golden_model()
    val, _ = O.inchworm_step_bare(exp, [0, 1, 2], 2 ** 8)
    # the per-step normalization rescales both sectors alike: compare the ratio
    assert abs((val[1] / val[0]).real - (P2[1] / P1[1]).real) < 1e-14


def test_first_bold_step_exact():
    exp = golden_model()
    P, _, _ = O.inchworm(exp, range(0, 4), range(0, 3), 2 ** 8, n_steps_max=1)
    assert abs((P[2, 1] / P[2, 0]).real - (P2[2] / P1[2]).real) < 1e-14
    assert P[2, 0] == -1j  # normalized maximum


def test_gate_discriminates_delta_interpolation():
    # the same bare step with grid-linear instead of spline hybridization is off by ~1e-7
    exp = golden_model()
    for p in exp.pairs:
        p.propagator.kind = abi.QIW_GF_GRID_LINEAR
    val, _ = O.inchworm_step_bare(exp, [0, 1, 2], 2 ** 8)
    assert abs((val[1] / val[0]).real - (P2[1] / P1[1]).real) > 1e-8


RTOL = np.sqrt(np.finfo(float).eps)  # Julia's isapprox default


def test_full_inchworm_and_correlator():
    exp = golden_model()
    P, Po, Ps = O.inchworm(exp, range(0, 4), range(0, 3), 2 ** 8)
    for got, ref in ((P[:, 0], P1), (P[:, 1], P2)):
        assert np.linalg.norm(got - ref) <= RTOL * np.linalg.norm(ref)
        assert np.linalg.norm(got - ref) <= 1e-14 * np.linalg.norm(ref)  # measured 4e-16
    assert np.all(P.real == 0.0)
    # order-resolved contributions are the raw (unnormalized) step results
    bare, _ = O.inchworm_step_bare(exp, [0, 1, 2], 2 ** 8)
    np.testing.assert_allclose(sum(Po[o][1] for o in Po), bare, rtol=1e-15, atol=0)
    assert np.isnan(Ps[1][5]).all() and (Ps[0][5] == 0).all()
    exp.P, exp.P_n_known = P, exp.grid.n_tau
    exp.add_corr_operators((exp.fock.c("0"), exp.fock.c_dag("0")))
    g = -O.correlator_2p(exp, range(0, 4), 2 ** 8)[0]
    assert np.linalg.norm(g - G) <= RTOL * np.linalg.norm(G)
    assert np.linalg.norm(g - G) <= 1e-14 * np.linalg.norm(G)  # measured 8e-16


def test_order_zero_is_atomic():
    # with order 0 only, P stays the atomic propagator up to the per-step normalization
    exp = model.anderson_single_orbital(n_tau=16)
    P, _, _ = O.inchworm(exp, [0], [0], 0)
    tau = exp.grid.tau
    E = [e[0] for e in exp.ed.energies]
    for s in range(1, 4):
        np.testing.assert_allclose((P[:, s] / P[:, 0]).real, np.exp(-tau * (E[s] - E[0])), rtol=1e-13)


def test_rejects_non_power_of_two():
    exp = golden_model()
    with pytest.raises(abi.QiwError) as e:
        O.inchworm(exp, [0, 1], [0, 1], 100)
    assert e.value.code == abi.QIW_E_NOT_POW2
