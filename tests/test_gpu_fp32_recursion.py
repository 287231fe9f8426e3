"""Float32 chain products over the whole two-orbital recursion (the bench's C2 model:
Hubbard-Kanamori, blocks 1, 4, 6, 4, 1, beta = 10, n_tau = 200, orders 0:4, 198 bold steps)
against the same run in float64 on the GPU, with the correlator G = -<c_0up(tau)
c^dag_0up> at orders 0:3.  The target for the float32 path is 1e-6 relative on P, Z, rho
and G after the full recursion (each step feeds the next, so step errors accumulate)."""
import numpy as np
import pytest

import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
TARGET = 1e-6


def solve(precision, N):
    k = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 200, 0.5)
    qiw.inchworm(k, k.grid, range(0, 5), range(0, 5), N, precision=precision)
    P = k.P.copy()
    Z = qiw.partition_function(k)
    qiw.normalize(k, k.grid.beta)
    rho = model.fock_density_matrix(k, k.P)
    k.add_corr_operators((-k.ed.fock.c((0, "up")), k.ed.fock.c_dag((0, "up"))))
    g = qiw.correlator_2p(k, k.grid, range(0, 4), N, precision=precision)[0]
    return P, Z, rho, g


def test_fp32_recursion_error_over_198_steps():
    N = 2 ** 8
    print("c2 recursion: n_tau=200, orders 0:4, N=2^8, float32 then float64", flush=True)
    P32, Z32, r32, g32 = solve(abi.QIW_PREC_FP32, N)
    P64, Z64, r64, g64 = solve(abi.QIW_PREC_FP64, N)
    errs = {"P": np.linalg.norm(P32 - P64) / np.linalg.norm(P64),
            "P_worst_row": max(np.linalg.norm(P32[j] - P64[j]) / np.linalg.norm(P64[j]) for j in range(P64.shape[0])),
            "Z": abs(Z32 - Z64) / abs(Z64),
            "rho": np.abs(r32 - r64).max() / np.abs(r64).max(),
            "G": np.linalg.norm(g32 - g64) / np.linalg.norm(g64)}
    print("float32 against float64 after the whole recursion:", {k: f"{v:.2e}" for k, v in errs.items()})
    for k, v in errs.items():
        assert 0 < v < TARGET, k
