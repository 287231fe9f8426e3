"""Oracle diff_inchworm! (src/inchworm.jl:528-778) against the reference's golden vector
(test/inchworm.jl:214-239 via test/inchworm.h5:/diff_inchworm/P) and closed forms."""
import json
import os

import numpy as np

import oracle_lib as O
from qiw import model

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "inchworm_h5.json")))


def fx(name):
    return np.array([complex(a, b) for a, b in FX[name]["data"]])


def golden_model():
    return model.single_state_golden_model(("occupied", "empty"))


def test_diff_inchworm_golden():
    """orders 0:3, N = 2^8, incremental-spline P: each sector within the reference test's
    own tolerance (isapprox, rtol = sqrt(eps))."""
    exp = golden_model()
    P, So, Ss = O.diff_inchworm(exp, range(0, 4), 2 ** 8)
    for s, name in enumerate(["/diff_inchworm/P/1", "/diff_inchworm/P/2"]):
        g = fx(name)
        assert np.linalg.norm(P[:, s] - g) <= np.sqrt(np.finfo(float).eps) * np.linalg.norm(g)
        assert np.linalg.norm(P[:, s] - g) <= 1e-14 * np.linalg.norm(g)  # bit-identical here
    assert np.all(P.real == 0.0)
    # the self-energy is imaginary; order 1 is exact (no deviation), order 2 and 3 sampled
    for o in (1, 2, 3):
        assert np.all(So[o].real == 0.0)
    assert np.all(Ss[1] == 0) and np.all(np.isnan(Ss[2][1:]))
    assert np.all(So[0] == 0)


def test_first_step_is_the_atomic_euler_step():
    """Step 1 has Σ * P = 0 (one grid point): P(h) = (1 - h (E + lambda)) P(0) exactly."""
    exp = golden_model()
    P, _, _ = O.diff_inchworm(exp, range(0, 4), 2 ** 8, n_steps_max=1)
    h = exp.grid.beta / (exp.grid.n_tau - 1)
    E = np.array([-0.1, 0.0])  # H = -mu n, mu = 0.1; sectors occupied, empty
    lam = np.log(np.exp(-exp.grid.beta * E).sum()) / exp.grid.beta
    np.testing.assert_allclose(P[1], (1 - h * (E + lam)) * (-1j), rtol=1e-15)


def test_order_one_self_energy_closed_form():
    """With only order 1, Σ(tau) = sum over the two pairs of Delta_p(tau) op_f P(tau) op_i:
    here Σ_occ(tau) = Delta_fwd(tau) P_empty(tau) and Σ_empty(tau) = Delta_bwd(tau) P_occ(tau)
    times the diagram's sign, with the propagator values of the same run."""
    exp = golden_model()
    P, So, _ = O.diff_inchworm(exp, [1], 2 ** 4, n_steps_max=4)
    # Σ of step a (tau_f_prev = grid[a]) is stored at tau_f = grid[a + 1]
    d_fwd = exp.pairs[0].propagator.values
    d_bwd = exp.pairs[1].propagator.values
    for a in range(1, 4):
        s_occ, s_emp = So[1][a + 1]
        r1 = s_occ / (d_fwd[a] * P[a, 1]) if abs(d_fwd[a] * P[a, 1]) > 0 else None
        r2 = s_emp / (d_bwd[a] * P[a, 0]) if abs(d_bwd[a] * P[a, 0]) > 0 else None
        assert r1 is not None and r2 is not None
        assert abs(r1 - r2) < 1e-12 and abs(abs(r1) - 1.0) < 1e-12


def test_order_one_self_energy_closed_form_grid_container():
    """The same closed form with the bold P on the grid (interpolate_ppgf = false, so Σ =
    zero(expansion.P) is a grid container too): Σ_orders holds the self-energy of step a
    (tau_f_prev = grid[a]) at tau_f = grid[a + 1], as set_ppgf!(Σ_orders[order], τ_i, τ_f, ...)
    writes it (src/inchworm.jl:766-769), for either container; a row shift pairs Delta and P
    of other times and fails the unit ratio."""
    e0 = golden_model()
    exp = model.Expansion(e0.ed, e0.grid, e0.pairs, interpolate_ppgf=False)
    P, So, _ = O.diff_inchworm(exp, [1], 2 ** 4, n_steps_max=5)
    d_fwd = exp.pairs[0].propagator.values
    d_bwd = exp.pairs[1].propagator.values
    assert np.all(So[1][0] == 0) and np.all(So[1][6:] == 0)  # rows of steps not taken
    for a in range(1, 5):
        s_occ, s_emp = So[1][a + 1]
        r1 = s_occ / (d_fwd[a] * P[a, 1])
        r2 = s_emp / (d_bwd[a] * P[a, 0])
        assert abs(r1 - r2) < 1e-12 and abs(abs(r1) - 1.0) < 1e-12
        # and not the values of the neighbouring rows
        assert abs(s_occ / (d_fwd[a + 1] * P[a + 1, 1]) - s_emp / (d_bwd[a + 1] * P[a + 1, 0])) > 1e-6
