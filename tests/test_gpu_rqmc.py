"""GPU parity of randomized qMC (N_seqs sequences per group, mean, standard deviation and
the target_std stopping rule, src/randomization.jl:86-100) against the float64 oracle,
through the C ABI.  Both sides draw the same scrambled sequences (qiw_rand bits), so the
results agree to the float64 tolerance of the plain path: 1e-10 relative.  A deviation is a
difference of samples, so its error is bounded by 1e-10 of the samples' scale (mean plus
deviation), not of the deviation itself."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
RTOL = 1e-10
SEED = 20241


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


def std_close(s_gpu, s_cpu, mean_cpu):
    scale = np.abs(mean_cpu).max() + np.abs(s_cpu).max()
    return np.abs(np.asarray(s_gpu) - np.asarray(s_cpu)).max() <= RTOL * scale


@pytest.fixture(scope="module", autouse=True)
def threads():
    O.set_threads(16)


def run_both(make, orders, orders_bare, N, rand, n_tau=12):
    eg, ec = make(n_tau), make(n_tau)
    n_seqs, scramble, seed, target = rand
    rp = qiw.RandomizationParams(seed if scramble else None, n_seqs, target)
    Po_g, Ps_g = qiw.inchworm(eg, eg.grid, orders, orders_bare, N, rand_params=rp)
    P_c, Po_c, Ps_c = O.inchworm(ec, orders, orders_bare, N, rand=rand)
    return eg, (Po_g, Ps_g), (P_c, Po_c, Ps_c)


def anderson(n_tau):
    return model.anderson_single_orbital(n_tau=n_tau)


def kanamori2(n_tau):
    return model.kanamori(2, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5)


@pytest.mark.parametrize("target", [0.0, 1e3, 3e-4])
def test_inchworm_sequences_and_stopping(target):
    """Four scrambled sequences per group and step; target 0 uses all of them, 1e3 stops
    every group after two, 3e-4 stops some groups earlier than others."""
    print(f"anderson n_tau=12 orders 0:3 N=2^8 N_seqs=4 target={target}", flush=True)
    eg, (Po_g, Ps_g), (P_c, Po_c, Ps_c) = run_both(anderson, range(0, 4), range(0, 4), 2 ** 8,
                                                    (4, 1, SEED, target))
    assert rel(eg.P, P_c) < RTOL
    for o in Po_c:
        assert np.abs(Po_g[o] - Po_c[o]).max() <= RTOL * np.abs(Po_c[o]).max() + 1e-300
        assert std_close(Ps_g[o], Ps_c[o], Po_c[o])
    assert np.all(Ps_g[0] == 0) and np.all(Ps_g[2][2:].real > 0)


def test_block_model_sequences():
    """Two-orbital Kanamori model (blocks up to 6x6, vector walk): three sequences."""
    print("kanamori2 n_tau=6 orders 0:2 N=2^5 N_seqs=3", flush=True)
    eg, (Po_g, Ps_g), (P_c, Po_c, Ps_c) = run_both(kanamori2, range(0, 3), range(0, 3), 2 ** 5,
                                                    (3, 1, SEED + 1, 0.0), n_tau=6)
    assert rel(eg.P, P_c) < RTOL
    for o in Po_c:
        assert std_close(Ps_g[o], Ps_c[o], Po_c[o])


def test_unscrambled_sequences_equal_the_plain_run():
    print("anderson n_tau=10 unscrambled N_seqs=3", flush=True)
    e1, e3 = anderson(10), anderson(10)
    Po1, Ps1 = qiw.inchworm(e1, e1.grid, range(0, 3), range(0, 3), 2 ** 7)
    Po3, Ps3 = qiw.inchworm(e3, e3.grid, range(0, 3), range(0, 3), 2 ** 7,
                            rand_params=qiw.RandomizationParams(None, 3, 0.0))
    # the same values; the sequence bookkeeping rounds the group sums separately
    np.testing.assert_allclose(e3.P, e1.P, rtol=1e-14, atol=1e-16)
    for o in (1, 2):
        np.testing.assert_allclose(Po3[o], Po1[o], rtol=1e-14, atol=1e-16)
        assert np.all(np.isnan(Ps1[o][1:])) and np.all(Ps3[o] == 0)


def test_one_scrambled_sequence():
    print("anderson n_tau=10 one scrambled sequence", flush=True)
    eg, (Po_g, Ps_g), (P_c, Po_c, Ps_c) = run_both(anderson, range(0, 3), range(0, 3), 2 ** 7,
                                                    (1, 1, SEED + 2, 0.0), n_tau=10)
    assert rel(eg.P, P_c) < RTOL
    for o in (1, 2):
        np.testing.assert_array_equal(np.isnan(Ps_g[o]), np.isnan(Ps_c[o]))


def test_bold_step_sequences():
    print("anderson n_tau=16 bold step n_w=7 N_seqs=5", flush=True)
    eg, ec = anderson(16), anderson(16)
    P, _, _ = O.inchworm(ec, range(0, 3), range(0, 3), 2 ** 7, n_steps_max=8)
    for e in (eg, ec):
        e.P, e.P_n_known = P, e.grid.n_tau
    rand = (5, 1, SEED + 3, 0.0)
    v_g, per_g = qiw.inchworm_step(eg, 7, range(0, 4), 2 ** 7, rand_params=qiw.RandomizationParams(SEED + 3, 5, 0.0))
    v_c, per_c = O.inchworm_step(ec, range(0, 4), 2 ** 7, 7, rand=rand)
    assert rel(v_g, v_c) < RTOL
    v_1, _ = O.inchworm_step(ec, range(0, 4), 2 ** 7, 7)
    assert rel(v_1, v_c) > 1e-8  # the sequences are not the staged one


def test_correlator_std():
    print("anderson correlators n_tau=10 orders 0:2 N=2^7 N_seqs=3 RequestStdDev", flush=True)
    eg = model.anderson_single_orbital(n_tau=10, with_correlators=True)
    ec = model.anderson_single_orbital(n_tau=10, with_correlators=True)
    P, _, _ = O.inchworm(ec, range(0, 3), range(0, 3), 2 ** 7)
    for e in (eg, ec):
        e.P, e.P_n_known = P, e.grid.n_tau
    rp = qiw.RandomizationParams(SEED + 4, 3, 0.0)
    gg, sg = qiw.correlator_2p(eg, eg.grid, range(0, 3), 2 ** 7, qiw.RequestStdDev(), rand_params=rp)
    gc, sc = O.correlator_2p(ec, range(0, 3), 2 ** 7, rand=(3, 1, SEED + 4, 0.0), with_std=True)
    for i in range(len(gg)):
        assert rel(gg[i], gc[i]) < RTOL
        assert std_close(sg[i], sc[i], gc[i])
        assert sg[i][0] == 0
    # without RequestStdDev the same call returns the values only
    g_only = qiw.correlator_2p(eg, eg.grid, range(0, 3), 2 ** 7, rand_params=qiw.RandomizationParams(SEED + 4, 3, 0.0))
    assert rel(g_only[0], gc[0]) < RTOL


def test_rand_params_rejected():
    ctx = qiw.context()
    e = anderson(8)
    qiw.inchworm_step_bare(e, [0, 1], 2 ** 4)  # stages the model and the sequence
    for rand, code in [((0, 1, 1, 0.0), abi.QIW_E_ARG), ((65, 1, 1, 0.0), abi.QIW_E_LIMIT),
                       ((2, 1, 1, -1.0), abi.QIW_E_ARG), ((2, 2, 1, 0.0), abi.QIW_E_ARG)]:
        run = abi.RunStruct([0, 1], [0, 1], 2 ** 4, rand=rand)
        assert ctx.lib.qiw_inchworm(ctx.h, run.ref(), None, None, None) == code
    run = abi.RunStruct([0, 1], [], 2 ** 4, rand=(2, 1, 1, 0.0))
    out = np.zeros(2 * 8 * 4)
    assert ctx.lib.qiw_correlator_2p_std(ctx.h, run.ref(), out.ctypes.data_as(abi._f64p), None) == abi.QIW_E_ARG
    # the learned warp takes one unscrambled sequence
    qiw.warp_init(seed=5)
    qiw.warp_enable(True)
    try:
        run = abi.RunStruct([0, 1], [0, 1], 2 ** 4, rand=(2, 1, 1, 0.0))
        assert ctx.lib.qiw_inchworm(ctx.h, run.ref(), None, None, None) == abi.QIW_E_ARG
    finally:
        qiw.warp_enable(False)


def test_no_samples_with_scrambled_sequences():
    """N_samples = 0: only the exact order 0 contributes (src/inchworm.jl:158), with or
    without randomization."""
    e1, e2 = anderson(8), anderson(8)
    Po1, Ps1 = qiw.inchworm(e1, e1.grid, range(0, 3), range(0, 3), 0)
    Po2, Ps2 = qiw.inchworm(e2, e2.grid, range(0, 3), range(0, 3), 0,
                            rand_params=qiw.RandomizationParams(SEED, 3, 0.0))
    np.testing.assert_array_equal(e1.P, e2.P)
    assert np.all(Po2[1] == 0) and np.all(Ps2[1] == 0)
    ec = model.anderson_single_orbital(n_tau=8, with_correlators=True)
    ec.P, ec.P_n_known = e1.P, ec.grid.n_tau
    g, gs = qiw.correlator_2p(ec, ec.grid, range(0, 3), 0, qiw.RequestStdDev(),
                              rand_params=qiw.RandomizationParams(SEED, 3, 0.0))
    assert np.all(np.isfinite(g[0])) and np.all(gs[0] == 0)
