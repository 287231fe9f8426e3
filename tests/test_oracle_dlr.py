"""DLR hybridization functions (keldysh_dlr.DLRImaginaryTimeGF, src/keldysh_dlr.jl) on the
oracle, and the reference's physics checks of test/bethe_dlr.jl.

* The DLR construction (frequencies by pivoted QR of the fermionic kernel, coefficients by
  a solve at the DLR time nodes, the algorithm of Lehmann.jl's DLRGrid / tau2dlr, which is
  not in this image) reproduces the semi-elliptic hybridization to its tolerance, and its
  particle-hole conjugate is exact.
* The oracle's pole-sum lookup equals the DLR evaluated directly, on the branch
  tau >= 0 and, through the statistics sign, tau < 0 (src/keldysh_dlr.jl:185-200).
* test/bethe_dlr.jl:104-190: particle-hole symmetry gives rho = 1/4 for every tested pair
  of order ranges, and inchworm! at orders 0:k meets the reference's tolerances against
  the NCA / OCA / TCA density matrices of test/bethe.h5 (order 1: < 2e-3 of NCA and closer
  to it than to the others; orders 2 and 3: < 4e-3 of OCA and TCA)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from qiw import abi, model

FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bethe_h5.json")))
BETA = 10.0


def ref_rho(name):
    return np.array(FX[f"/rho/{name}"]["data"]).reshape(4, 4)


def test_dlr_reproduces_the_semi_elliptic_hybridization():
    grid = model.ImaginaryTimeGrid(BETA, 129)
    f = lambda tau: model.bethe_values(0.5, BETA, tau, eps=0.25)  # noqa: E731
    g = model.dlr_fit(f, 1.25, BETA, grid, rtol=1e-6)
    tt = np.linspace(0.0, BETA, 4001)
    ref = f(tt)
    assert 8 <= len(g.omega) <= 20
    assert np.abs(g(tt) - ref).max() < 1e-6 * np.abs(ref).max()
    ph = model.ph_conj(g)
    np.testing.assert_allclose(ph(tt), g(BETA - tt), rtol=0, atol=1e-15)
    # a tighter tolerance needs more frequencies and is more accurate
    g9 = model.dlr_fit(f, 1.25, BETA, grid, rtol=1e-9)
    assert len(g9.omega) > len(g.omega) and np.abs(g9(tt) - ref).max() < 1e-8 * np.abs(ref).max()


def test_oracle_lookup_is_the_pole_sum():
    grid = model.ImaginaryTimeGrid(BETA, 17)
    g = 0.25 * model.dlr_fit(lambda tau: model.bethe_values(0.5, BETA, tau, eps=0.25), 1.25, BETA, grid)
    for a, b in [(3.0, 1.0), (7.5, 0.5), (10.0, 0.0), (0.0, 0.0), (2.0, 6.0), (0.5, 9.5)]:
        v = O.dlr_eval(g, a, b)
        want = g(np.array([a - b]))[0] if a >= b else g.xi * g(np.array([BETA + a - b]))[0]
        assert abs(v - want) <= 1e-14 * (1 + abs(want))
    exp = model.bethe_two_orbitals_dlr(17)
    ms = exp.to_struct()
    assert ms.struct.dlr_rank == len(g.omega)


def solve(n_tau, orders, N, mu_bethe=0.25, orders_bare=None):
    exp = model.bethe_two_orbitals_dlr(n_tau, mu_bethe)
    P, _, _ = O.inchworm(exp, orders, orders if orders_bare is None else orders_bare, N)
    return model.fock_density_matrix(exp, P)


@pytest.mark.parametrize("orders_bare,orders", [((0,), (0,)), ((0, 1), (0,)), ((0,), (0, 1)), ((0, 1, 2), (0,)),
                                                ((0,), (0, 1, 2)), ((0, 1, 2, 3), (0,)), ((0,), (0, 1, 2, 3)),
                                                ((0, 1, 2, 3, 4), (0,)), ((0,), (0, 1, 2, 3, 4))])
def test_particle_hole_symmetry(orders_bare, orders):
    rho = np.real(np.diag(solve(3, list(orders), 2 ** 4, mu_bethe=0.0, orders_bare=list(orders_bare))))
    assert np.linalg.norm(rho - 0.25) <= np.sqrt(np.finfo(float).eps) * np.linalg.norm(np.full(4, 0.25))


@pytest.mark.parametrize("k,name,tol", [(1, "NCA", 2e-3), (2, "OCA", 4e-3), (3, "TCA", 4e-3)])
def test_orders_against_bethe_h5(k, name, tol):
    rho = solve(128, list(range(k + 1)), 8 * 2 ** 5)
    r = np.real(np.diag(rho))
    assert abs(r.sum() - 1) < 1e-12 and abs(r[1] - r[2]) < 1e-12
    d = {n: np.abs(rho - ref_rho(n)).max() for n in ("exact", "NCA", "OCA", "TCA")}
    assert d[name] < tol
    others = [n for n in ("NCA", "OCA", "TCA") if n != name] + (["exact"] if k < 3 else [])
    assert all(d[name] < d[o] for o in others), d
