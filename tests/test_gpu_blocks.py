"""GPU parity for models with sector blocks larger than 1x1 (two-orbital Kanamori model,
blocks [1, 4, 6, 4, 1]) against the float64 oracle, through the C ABI."""
import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


def kanamori2(n_tau=8, interpolate=False):
    return model.kanamori(2, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5, interpolate=interpolate)


def test_block_bare_step():
    k = kanamori2()
    v_gpu, per_gpu = qiw.inchworm_step_bare(k, [0, 1, 2, 3], 2 ** 5)
    v_cpu, per_cpu = O.inchworm_step_bare(k, [0, 1, 2, 3], 2 ** 5)
    assert rel(v_gpu, v_cpu) < RTOL
    assert rel(per_gpu, per_cpu) < RTOL


@pytest.mark.parametrize("interpolate", [False, True])
def test_block_inchworm_steps(interpolate):
    kg, kc = kanamori2(interpolate=interpolate), kanamori2(interpolate=interpolate)
    O.set_threads(16)
    P_c, Po_c, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 6)
    Po_g, _ = qiw.inchworm(kg, kg.grid, range(0, 3), range(0, 3), 2 ** 6)
    assert rel(kg.P, P_c) < RTOL
    for o in Po_c:
        assert np.abs(Po_g[o] - Po_c[o]).max() < RTOL * np.abs(P_c).max()


def test_block_bold_step_order3():
    kg, kc = kanamori2(), kanamori2()
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 6, n_steps_max=2)
    kg.P, kg.P_n_known = P_c, kg.grid.n_tau
    kc.P, kc.P_n_known = P_c, kc.grid.n_tau
    v_gpu, _ = qiw.inchworm_step(kg, 3, range(0, 4), 2 ** 5)
    v_cpu, _ = O.inchworm_step(kc, range(0, 4), 2 ** 5, 3)
    assert rel(v_gpu, v_cpu) < RTOL


def test_block_correlator():
    kg, kc = kanamori2(n_tau=6), kanamori2(n_tau=6)
    f = kg.ed.fock
    for e in (kg, kc):
        e.add_corr_operators((-f.c((0, "up")), f.c_dag((0, "up"))))
        e.add_corr_operators((f.n((0, "up")), f.n((1, "dn"))))
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 5)
    kg.P, kg.P_n_known = P_c, kg.grid.n_tau
    kc.P, kc.P_n_known = P_c, kc.grid.n_tau
    gg = qiw.correlator_2p(kg, kg.grid, range(0, 3), 2 ** 5)
    gc = O.correlator_2p(kc, range(0, 3), 2 ** 5)
    for a, b in zip(gg, gc):
        assert rel(a, b) < RTOL


def test_block_fp32_chain_products():
    """QIW_PREC_FP32: float32 chain products, float64 column sums and bookkeeping.  The
    error against the float64 oracle is stated here: below 2e-6 relative for P and Z."""
    kg, kc = kanamori2(n_tau=10), kanamori2(n_tau=10)
    P_c, _, _ = O.inchworm(kc, range(0, 4), range(0, 4), 2 ** 6)
    qiw.inchworm(kg, kg.grid, range(0, 4), range(0, 4), 2 ** 6, precision=qiw.abi.QIW_PREC_FP32)
    err_P = rel(kg.P, P_c)
    Zg, Zc = qiw.partition_function(kg), sum(np.trace(kc.block(P_c[-1], s)) * 1j for s in range(5))
    err_Z = abs(Zg - Zc) / abs(Zc)
    print(f"fp32 chain products: rel err P {err_P:.2e}, Z {err_Z:.2e}")
    assert 0 < err_P < 2e-6 and err_Z < 2e-6
    assert rel(kg.P, P_c) > 1e-13  # really float32 (differs from the float64 path)


def test_block_fp32_rho_and_g():
    """QIW_PREC_FP32 errors against the float64 oracle for the density matrix rho (from the
    normalised P(beta)) and the Green's function G(tau) = -<c(tau) c^dag> on the two-orbital
    model; bound stated here: 2e-6 relative and 2e-6 max-abs for both."""
    kg, kc = kanamori2(n_tau=10), kanamori2(n_tau=10)
    f = kg.ed.fock
    for e in (kg, kc):
        e.add_corr_operators((-f.c((0, "up")), f.c_dag((0, "up"))))
    P_c, _, _ = O.inchworm(kc, range(0, 4), range(0, 4), 2 ** 6)
    kc.P, kc.P_n_known = P_c, kc.grid.n_tau
    qiw.inchworm(kg, kg.grid, range(0, 4), range(0, 4), 2 ** 6, precision=qiw.abi.QIW_PREC_FP32)
    rho_g, rho_c = qiw.density_matrix(kg), qiw.density_matrix(kc)
    rg, rc = np.concatenate([b.ravel() for b in rho_g]), np.concatenate([b.ravel() for b in rho_c])
    err_rho, abs_rho = rel(rg, rc), np.abs(rg - rc).max()
    g_g = qiw.correlator_2p(kg, kg.grid, range(0, 4), 2 ** 6, precision=qiw.abi.QIW_PREC_FP32)[0]
    g_c = O.correlator_2p(kc, range(0, 4), 2 ** 6)[0]
    err_g, abs_g = rel(g_g, g_c), np.abs(g_g - g_c).max()
    print(f"fp32 chain products: rho rel {err_rho:.2e} max-abs {abs_rho:.2e}; G rel {err_g:.2e} max-abs {abs_g:.2e}")
    assert err_rho < 2e-6 and abs_rho < 2e-6
    assert err_g < 2e-6 and abs_g < 2e-6


def kanamori3(n_tau=5):
    # three-orbital Kanamori model (C3 physics): blocks 1, 6, 15, 20, 15, 6, 1, 12 pairs
    return model.kanamori(3, 3.0, 0.3, 0.0, 20.0, n_tau, 0.4)


def test_three_orbital_blocks():
    """Blocks wider than the vector walk (up to 20) run the LDS block walk."""
    kg, kc = kanamori3(), kanamori3()
    v_gpu, per_gpu = qiw.inchworm_step_bare(kg, [0, 1, 2], 2 ** 3)
    v_cpu, per_cpu = O.inchworm_step_bare(kc, [0, 1, 2], 2 ** 3)
    assert rel(v_gpu, v_cpu) < RTOL
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 3)
    qiw.inchworm(kg, kg.grid, range(0, 3), range(0, 3), 2 ** 3)
    assert rel(kg.P, P_c) < RTOL


def test_three_orbital_matrix_cores_fp32():
    """Wide blocks with float32 chain products run on the matrix cores (k_eval_mfma_t); the
    error against the float64 oracle is stated here: below 2e-6 relative."""
    kg, kc = kanamori3(n_tau=4), kanamori3(n_tau=4)
    fp32 = qiw.abi.QIW_PREC_FP32
    O.set_threads(16)
    P_c, _, _ = O.inchworm(kc, range(0, 4), range(0, 4), 2 ** 2)
    qiw.inchworm(kg, kg.grid, range(0, 4), range(0, 4), 2 ** 2, precision=fp32)
    err = rel(kg.P, P_c)
    print(f"three-orbital, matrix cores (float32): rel err P {err:.2e}")
    assert 0 < err < 2e-6
    kc.P, kc.P_n_known = P_c, kc.grid.n_tau
    kg.P, kg.P_n_known = P_c.copy(), kg.grid.n_tau
    v_c, _ = O.inchworm_step(kc, range(0, 4), 2 ** 3, 2)
    v_g, _ = qiw.inchworm_step(kg, 2, range(0, 4), 2 ** 3, precision=fp32)
    assert rel(v_g, v_c) < 2e-6


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_block_one_table(precision):
    """Two-orbital model whose 4 pairs share one hybridization table (W slot stride 2)."""
    def build():
        k = kanamori2(n_tau=8)
        f = k.ed.fock
        labels = [(a, s) for a in range(2) for s in ("up", "dn")]
        delta = k.pairs[0].propagator
        return model.Expansion(k.ed, k.grid, [model.InteractionPair(f.c_dag(l), f.c(l), delta) for l in labels])
    kg, kc = build(), build()
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 5, n_steps_max=4)
    for e in (kg, kc):
        e.P, e.P_n_known = P_c, e.grid.n_tau
    prec = qiw.abi.QIW_PREC_FP32 if precision == "fp32" else qiw.abi.QIW_PREC_FP64
    v_gpu, _ = qiw.inchworm_step(kg, 4, range(0, 4), 2 ** 5, precision=prec)
    v_cpu, _ = O.inchworm_step(kc, range(0, 4), 2 ** 5, 4)
    assert rel(v_gpu, v_cpu) < (RTOL if precision == "fp64" else 2e-6)


def test_block_work_mappings():
    """Vector programs cut into one-topology chunks, and the matrix-core walk at 1 and 16
    samples per workgroup, keep parity."""
    kg, kc = kanamori2(), kanamori2()
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 6, n_steps_max=2)
    kg.P, kg.P_n_known = P_c, kg.grid.n_tau
    kc.P, kc.P_n_known = P_c, kc.grid.n_tau
    v_cpu, _ = O.inchworm_step(kc, range(0, 4), 2 ** 5, 3)
    g3, c3 = kanamori3(n_tau=4), kanamori3(n_tau=4)
    P3, _, _ = O.inchworm(c3, range(0, 3), range(0, 3), 2 ** 2, n_steps_max=2)
    for e in (g3, c3):
        e.P, e.P_n_known = P3, e.grid.n_tau
    v3_cpu, _ = O.inchworm_step(c3, range(0, 3), 2 ** 3, 2)
    try:
        qiw.set_tuning(chunk_cost=256)
        v_gpu, _ = qiw.inchworm_step(kg, 3, range(0, 4), 2 ** 5)
        assert rel(v_gpu, v_cpu) < RTOL
        for spt in (1, 16):
            qiw.set_tuning(samples_per_task=spt)
            v3, _ = qiw.inchworm_step(g3, 2, range(0, 3), 2 ** 3, precision=qiw.abi.QIW_PREC_FP32)
            assert rel(v3, v3_cpu) < 2e-6
    finally:
        qiw.set_tuning()


def test_autopartitioned_model_from_hamiltonian():
    """Two-orbital Kanamori model built by Expansion(H, soi, grid; hybridization)
    (src/expansion.jl:222-291): the autopartition gives blocks finer than the particle
    number; GPU against the oracle for a short inchworm!."""
    from test_model_construction import kanamori_H
    print("kanamori2 autopartitioned n_tau=6 orders 0:2 N=2^5", flush=True)

    def build():
        f, H, labels = kanamori_H()
        grid = model.ImaginaryTimeGrid(10.0, 6)
        d = 0.25 * model.bethe_gf(1.0, 10.0, grid).values
        hyb = np.zeros((len(labels), len(labels), grid.n_tau), complex)
        for i in range(len(labels)):
            hyb[i, i] = d
        return model.Expansion.from_hamiltonian(H, f, grid, hyb)

    kg, kc = build(), build()
    assert len(kg.ed.dims) > 5
    P_c, _, _ = O.inchworm(kc, range(0, 3), range(0, 3), 2 ** 5)
    qiw.inchworm(kg, kg.grid, range(0, 3), range(0, 3), 2 ** 5)
    assert rel(kg.P, P_c) < RTOL


def test_real_nn_interaction_is_rejected():
    """A real-valued nn-interaction U (the reference's ImaginaryTimeGF{Float64}) makes i*U
    imaginary, which the real-arithmetic kernels do not take: QIW_E_ARG, no evaluation."""
    from test_model_construction import anderson_H
    f = model.FockSpace(["up", "dn"])
    grid = model.ImaginaryTimeGrid(10.0, 8)
    hyb = np.zeros((2, 2, 8), complex)
    hyb[0, 0] = hyb[1, 1] = 0.25 * model.bethe_gf(1.0, 10.0, grid).values
    U = np.zeros((2, 2, 8))
    U[0, 1] = 0.3 * np.cosh(grid.tau - 5.0) / np.cosh(5.0)
    e = model.Expansion.from_hamiltonian(anderson_H(f), f, grid, hyb, nn_interaction=U)
    with pytest.raises(qiw.abi.QiwError) as err:
        qiw.inchworm(e, e.grid, range(0, 2), range(0, 2), 2 ** 4)
    assert err.value.code == qiw.abi.QIW_E_ARG


@pytest.mark.parametrize("walk", ["vector", "matrix-core"])
def test_graph_replay_of_block_walks(walk):
    """The bold-step loop of the vector and matrix-core walks is replayed from a hipGraph
    (from the second call on, once the walk's tables are staged); the replay gives the eager
    loop's result bit for bit."""
    if walk == "vector":
        make = lambda: model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 10, 0.5)  # noqa: E731
        prec = abi.QIW_PREC_FP32
    else:
        def make():
            k3 = model.kanamori(3, 2.0, 0.2, -1.0, 10.0, 8, 0.5)
            return model.Expansion(k3.ed, k3.grid, k3.pairs[:2])
        prec = abi.QIW_PREC_FP32
    print(f"graphs: {walk} walk, whole inchworm!, orders 0:2", flush=True)
    runs, replays = [], []
    for graphs in (True, True, False):
        qiw.set_graphs(graphs)
        e = make()
        qiw.inchworm(e, e.grid, range(0, 3), range(0, 3), 2 ** 5, precision=prec)
        runs.append(e.P.copy())
        replays.append(qiw.stats()["graph_replays"])
    qiw.set_graphs(True)
    assert replays[1] == 1 and replays[2] == 0
    np.testing.assert_array_equal(runs[1], runs[2])
    np.testing.assert_array_equal(runs[0], runs[2])
