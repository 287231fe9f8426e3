"""The reference's Bethe-lattice physics checks (test/bethe.jl, test/bethe_gf.jl) on the GPU:
two non-interacting orbitals with a grid-stored semi-elliptic hybridization built by the
higher-level Expansion constructor.  The GPU run equals the float64 oracle to 1e-10 (rho and
G at order 2), and its rho and G meet the reference's own tolerances against the DLR NCA,
OCA and TCA results of test/bethe.h5; particle-hole symmetry gives rho = 1/4 exactly."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import model

pytestmark = pytest.mark.gpu
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bethe_h5.json")))


def ref_rho(name):
    return np.array(FX[f"/rho/{name}"]["data"]).reshape(4, 4)


def ref_g(name):
    return np.array([complex(a, b) for a, b in FX[f"/g/{name}"]["data"]])


def gpu_solve(n_tau, orders, N, mu_bethe=0.25, orders_gf=None, orders_bare=None):
    exp = model.bethe_two_orbitals(n_tau, mu_bethe)
    qiw.inchworm(exp, exp.grid, orders, orders if orders_bare is None else orders_bare, N)
    rho = model.fock_density_matrix(exp, exp.P)
    g = None
    if orders_gf is not None:
        exp.add_corr_operators((exp.fock.c(1), exp.fock.c_dag(1)))
        g = -qiw.correlator_2p(exp, exp.grid, orders_gf, N)[0]
    return exp, rho, g


def test_particle_hole_symmetry():
    for orders_bare, orders in [((0, 1, 2, 3), (0,)), ((0,), (0, 1, 2, 3)), ((0, 1, 2), (0, 1, 2))]:
        _, rho, _ = gpu_solve(3, list(orders), 2 ** 4, mu_bethe=0.0, orders_bare=list(orders_bare))
        r = np.real(np.diag(rho))
        assert np.linalg.norm(r - 0.25) <= np.sqrt(np.finfo(float).eps) * 0.5


def test_order_2_against_oracle_and_oca():
    print("bethe: n_tau=128, orders 0:2, N=512, correlator orders 0:1", flush=True)
    exp, rho, g = gpu_solve(128, [0, 1, 2], 8 * 2 ** 6, orders_gf=[0, 1])
    ec = model.bethe_two_orbitals(128)
    P, _, _ = O.inchworm(ec, [0, 1, 2], [0, 1, 2], 8 * 2 ** 6)
    assert np.linalg.norm(exp.P - P) <= 1e-10 * np.linalg.norm(P)
    ec.P, ec.P_n_known = P, ec.grid.n_tau
    ec.add_corr_operators((ec.fock.c(1), ec.fock.c_dag(1)))
    gc = -O.correlator_2p(ec, [0, 1], 8 * 2 ** 6)[0]
    assert np.linalg.norm(g - gc) <= 1e-10 * np.linalg.norm(gc)
    d = {k: np.abs(rho - ref_rho(k)).max() for k in ("exact", "NCA", "OCA", "TCA")}
    assert d["OCA"] < 2e-3 and d["OCA"] < min(d["NCA"], d["TCA"], d["exact"])
    dg = {k: np.abs(g - ref_g(k)).max() for k in ("NCA", "OCA")}
    assert dg["OCA"] < 1e-3 and dg["OCA"] < dg["NCA"]


@pytest.mark.parametrize("k,name,tol", [(1, "NCA", 2e-3), (3, "TCA", 4e-3)])
def test_orders_1_and_3(k, name, tol):
    _, rho, _ = gpu_solve(128, list(range(k + 1)), 8 * 2 ** 5)
    r = np.real(np.diag(rho))
    assert abs(r.sum() - 1) < 1e-12 and abs(r[1] - r[2]) < 1e-12
    d = {n: np.abs(rho - ref_rho(n)).max() for n in ("exact", "NCA", "OCA", "TCA")}
    assert d[name] < tol
    assert d[name] < min(v for n, v in d.items() if n not in (name, "exact"))
