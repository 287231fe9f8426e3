"""GPU parity of DLR hybridization functions (pair propagators of kind QIW_GF_DLR, exact
pole sums in the arc lookups) against the float64 oracle through the C ABI, and the
reference's test/bethe_dlr.jl physics checks on the GPU path.

* Two non-interacting orbitals with the semi-elliptic bath as a DLR (test/bethe_dlr.jl
  model), scalar walk: the whole inchworm! equals the oracle to 1e-10; particle-hole
  symmetry gives rho = 1/4; at orders 0:2 rho is within the reference's 4e-3 of OCA.
* Two-orbital Kanamori model with the bath as a DLR, vector walk: one bold step at orders
  0:3 in float64 (1e-10) and with float32 chain products (2e-6)."""
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
RTOL = 1e-10
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "bethe_h5.json")))


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


@pytest.fixture(scope="module", autouse=True)
def threads():
    O.set_threads(16)


def test_bethe_dlr_whole_run_matches_oracle():
    print("dlr: bethe two orbitals, n_tau=17, orders 0:3, N=2^6", flush=True)
    eg, ec = model.bethe_two_orbitals_dlr(17), model.bethe_two_orbitals_dlr(17)
    Po_g, _ = qiw.inchworm(eg, eg.grid, range(0, 4), range(0, 4), 2 ** 6)
    P_c, Po_c, _ = O.inchworm(ec, range(0, 4), range(0, 4), 2 ** 6)
    assert rel(eg.P, P_c) < RTOL
    for o in range(4):
        assert np.abs(Po_g[o] - Po_c[o]).max() < RTOL * np.abs(P_c).max()


def test_bethe_dlr_particle_hole_symmetry():
    for orders_bare, orders in [((0, 1, 2, 3), (0,)), ((0,), (0, 1, 2, 3)), ((0, 1, 2), (0, 1, 2))]:
        e = model.bethe_two_orbitals_dlr(3, mu_bethe=0.0)
        qiw.inchworm(e, e.grid, list(orders), list(orders_bare), 2 ** 4)
        r = np.real(np.diag(model.fock_density_matrix(e, e.P)))
        assert np.linalg.norm(r - 0.25) <= np.sqrt(np.finfo(float).eps) * 0.5


def test_bethe_dlr_order_2_against_bethe_h5():
    print("dlr: bethe two orbitals, n_tau=128, orders 0:2, N=256", flush=True)
    e = model.bethe_two_orbitals_dlr(128)
    qiw.inchworm(e, e.grid, [0, 1, 2], [0, 1, 2], 8 * 2 ** 5)
    rho = model.fock_density_matrix(e, e.P)
    d = {k: np.abs(rho - np.array(FX[f"/rho/{k}"]["data"]).reshape(4, 4)).max() for k in ("exact", "NCA", "OCA", "TCA")}
    assert d["OCA"] < 4e-3 and d["OCA"] < min(d["NCA"], d["TCA"], d["exact"])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_kanamori_dlr_bold_step(precision):
    prec = abi.QIW_PREC_FP64 if precision == "fp64" else abi.QIW_PREC_FP32
    kc = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 24, 0.5, dlr=True)
    P, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 5, n_steps_max=8)
    kg = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 24, 0.5, dlr=True)
    for e in (kg, kc):
        e.P, e.P_n_known = P.copy(), e.grid.n_tau
    print(f"dlr: kanamori2 bold step n_w=9, orders 0:3, N=2^5, {precision}", flush=True)
    v_gpu, per_gpu = qiw.inchworm_step(kg, 9, range(0, 4), 2 ** 5, precision=prec)
    v_cpu, per_cpu = O.inchworm_step(kc, range(0, 4), 2 ** 5, 9)
    err = rel(v_gpu, v_cpu)
    assert err < (RTOL if precision == "fp64" else 2e-6)
