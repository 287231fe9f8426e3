"""GPU parity at the configured sizes of the benchmark workloads, against the float64
oracle through the C ABI.

* Single-orbital Anderson model at n_tau = 200, orders 0:4, N = 2^14 (the bench's C1
  workload): the initial step and a bold step on the oracle's propagator after two bold
  steps, at 1e-10; then properties of the whole GPU run that the method fixes exactly
  (P(0) = -i, the per-step normalisation max diag iP(beta) = 1, spin symmetry of rho, the
  sum rule of G at 0 and beta, periodicity of the bosonic correlators).
* Two-orbital Kanamori model (blocks 1, 4, 6, 4, 1), orders 0:4: one bold step in float64
  at 1e-10 and with float32 chain products within 2e-6, the order-4 levels of the vector
  walk included."""
import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


def c1(with_correlators=False):
    return model.anderson_single_orbital(beta=10.0, n_tau=200, eps=0.1, U=1.0, V=0.5, D=2.0,
                                         with_correlators=with_correlators)


def test_c1_bare_and_bold_step_at_size():
    N = 2 ** 14
    O.set_threads(16)
    print("c1: bare step, N=2^14, orders 0:4", flush=True)
    v_gpu, per_gpu = qiw.inchworm_step_bare(c1(), range(0, 5), N)
    v_cpu, per_cpu = O.inchworm_step_bare(c1(), range(0, 5), N)
    assert rel(v_gpu, v_cpu) < RTOL and rel(per_gpu, per_cpu) < RTOL
    print("c1: bold step n_w=3 on the oracle's propagator after two bold steps", flush=True)
    P, _, _ = O.inchworm(c1(), range(0, 5), range(0, 5), N, n_steps_max=2)
    eg, ec = c1(), c1()
    for e in (eg, ec):
        e.P, e.P_n_known = P.copy(), e.grid.n_tau
    v_gpu, per_gpu = qiw.inchworm_step(eg, 3, range(0, 5), N)
    v_cpu, per_cpu = O.inchworm_step(ec, range(0, 5), N, 3)
    assert rel(v_gpu, v_cpu) < RTOL
    for o in range(5):
        assert np.abs(per_gpu[o] - per_cpu[o]).max() < RTOL * np.abs(v_cpu).max()


def test_c1_whole_run_properties():
    N = 2 ** 14
    e = c1(with_correlators=True)
    print("c1: whole inchworm! and correlators, N=2^14", flush=True)
    qiw.inchworm(e, e.grid, range(0, 5), range(0, 5), N)
    P = e.P
    np.testing.assert_array_equal(P[0], -1j * np.ones(4))          # P(0) = -i in every sector
    assert np.max(-P[-1].imag) == pytest.approx(1.0, abs=1e-15)     # normalize!(P, beta) of the last step
    assert np.all(P.real == 0.0) and np.all(-P.imag > 0)
    qiw.normalize(e, e.grid.beta)                                   # normalize!(P, beta): Z = 1 (README)
    rho = [b.ravel()[0].real for b in qiw.density_matrix(e)]        # sectors (0,0), (0,1), (1,0), (1,1)
    assert sum(rho) == pytest.approx(1.0, abs=1e-14)
    assert abs(rho[1] - rho[2]) < 1e-12 * rho[1]                    # spin symmetry
    g, d, chi = qiw.correlator_2p(e, e.grid, range(0, 4), N)
    # G = -<c_up(tau) c_up^dag>, stored as i <c(tau) c^dag> (values times -i, as the
    # reference's Green's function containers): G(0) + G(beta) = i (<c c^dag> + <c^dag c>) = i
    assert abs((g[0] + g[-1]) - 1j) < 1e-12
    assert np.all(g.real == 0.0)
    assert abs(d[0] - d[-1]) < 1e-12 and abs(chi[0] - chi[-1]) < 1e-12  # bosonic: periodic


def kanamori2(n_tau=24):
    return model.kanamori(2, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_c2_bold_step_order4(precision):
    prec = abi.QIW_PREC_FP64 if precision == "fp64" else abi.QIW_PREC_FP32
    N = 2 ** 4
    O.set_threads(16)
    kc = kanamori2()
    P, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 5, n_steps_max=8)
    kg = kanamori2()
    for e in (kg, kc):
        e.P, e.P_n_known = P.copy(), e.grid.n_tau
    print(f"c2: bold step n_w=9, orders 0:4, N=2^4, {precision}", flush=True)
    v_gpu, per_gpu = qiw.inchworm_step(kg, 9, range(0, 5), N, precision=prec)
    v_cpu, per_cpu = O.inchworm_step(kc, range(0, 5), N, 9)
    err = rel(v_gpu, v_cpu)
    err4 = np.abs(per_gpu[4] - per_cpu[4]).max() / np.abs(per_cpu[4]).max()
    print(f"c2 order 4 ({precision}): rel err step {err:.2e}, order-4 part {err4:.2e}")
    assert np.abs(per_cpu[4]).max() > 0
    if precision == "fp64":
        assert err < RTOL and err4 < 1e-9
    else:
        assert 0 < err < 2e-6 and err4 < 2e-5


def wide_blocks_one_bath(n_tau=12):
    """Three-orbital Kanamori impurity (blocks 1, 6, 15, 20, 15, 6, 1) with a bath on one
    spin-orbital only (2 pairs): order-4 diagrams stay cheap for the oracle while the
    blocks go to the matrix-core walk (float32) and the LDS block walk (float64)."""
    k3 = model.kanamori(3, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5)
    return model.Expansion(k3.ed, k3.grid, k3.pairs[:2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_wide_blocks_bold_step_order4(precision):
    """k_eval_block (float64) and k_eval_mfma_t<4, 2, 2> (float32) at order 4."""
    prec = abi.QIW_PREC_FP64 if precision == "fp64" else abi.QIW_PREC_FP32
    O.set_threads(16)
    ec = wide_blocks_one_bath()
    P, _, _ = O.inchworm(ec, range(0, 3), range(0, 3), 2 ** 3, n_steps_max=4)
    eg = wide_blocks_one_bath()
    for e in (eg, ec):
        e.P, e.P_n_known = P.copy(), e.grid.n_tau
    print(f"wide blocks: bold step n_w=5, orders 0:4, N=2^3, {precision}", flush=True)
    v_gpu, per_gpu = qiw.inchworm_step(eg, 5, range(0, 5), 2 ** 3, precision=prec)
    v_cpu, per_cpu = O.inchworm_step(ec, range(0, 5), 2 ** 3, 5)
    err, err4 = rel(v_gpu, v_cpu), np.abs(per_gpu[4] - per_cpu[4]).max() / np.abs(per_cpu[4]).max()
    print(f"wide blocks order 4 ({precision}): rel err step {err:.2e}, order-4 part {err4:.2e}")
    assert np.abs(per_cpu[4]).max() > 0
    if precision == "fp64":
        assert err < RTOL and err4 < 1e-9
    else:
        assert 0 < err < 2e-6 and err4 < 2e-5
