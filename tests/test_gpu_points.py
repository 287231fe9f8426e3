"""GPU parity of caller-supplied point sequences (qiw_rand.scramble = 3, qiw_set_points):
the sequence plug point seq_type of inchworm!/correlator_2p (src/inchworm.jl:129,233,339,
src/randomization.jl:82-99), against the float64 oracle, through the C ABI.

* Points equal to the keyed mode's scrambled Sobol points: the GPU equals its own keyed
  mode bit for bit, and the oracle to 1e-10 (inchworm!).
* RandomSeq points (src/utility.jl:336-414): the GPU equals the oracle to 1e-10 for the two
  steps, correlator_2p with deviations (vanishing groups take two sequences) and
  diff_inchworm!.
* The Python API draws seq_type(2k, scramble_rng=rng) in qiw_scramble_dims order
  (RandomizationParams(..., seq_type=RandomSeq))."""
import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import model
from qiw.randomization import RandomSeq, draw_points
from qiw.sobol import ScrambledSobolSeq

pytestmark = pytest.mark.gpu
RTOL = 1e-10
SEED = 0x5EED1234ABCDEF
N_SEQS = 3
N = 2 ** 5


class Staged:
    """rand_params for point sequences the test stages itself (scramble = 3)."""
    caller_drawn = False

    def __init__(self, n_seqs=N_SEQS):
        self.N_seqs = n_seqs

    def struct_args(self):
        return (self.N_seqs, 3, 0, 0.0)


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def threads():
    O.set_threads(16)


def anderson(n_tau=5, corr=False):
    return model.anderson_single_orbital(beta=2.0, n_tau=n_tau, with_correlators=corr)


def kanamori2(n_tau=5):
    return model.kanamori(2, 2.0, 0.2, -1.0, 2.0, n_tau, 0.5)


def oracle_log(run_oracle, stride):
    rng = np.random.default_rng(3)
    O.set_scrambles(rng.integers(0, 2 ** 32, size=(4000, stride, 32), dtype=np.uint32),
                    rng.integers(0, 2 ** 32, size=(4000, stride), dtype=np.uint32))
    run_oracle((N_SEQS, 2, 0, 0.0))
    return O.scrambles_log()


def stage_both(pts):
    O.set_points(pts)
    qiw.set_points(pts)


def random_points(log, stride, seed=11):
    return draw_points(np.random.default_rng(seed), RandomSeq, [2 * k for (_, _, k, _, _) in log], stride, N)


def test_keyed_points_reproduce_keyed_mode():
    orders, orders_bare = [0, 1, 2], [0, 1, 2, 3]
    print(f"points: inchworm anderson n_tau=5 orders 0:2 bare 0:3 N={N}, keyed Sobol points", flush=True)
    log = oracle_log(lambda r: O.inchworm(anderson(), orders, orders_bare, N, rand=r), 6)
    staged, _ = O.sobol_tables(6)
    pts = np.zeros((len(log), N, 6))
    for i, (dom, key, k, na, s) in enumerate(log):
        m, x = O.rq_scramble(SEED, O.rq_stream(dom, key, k, na, s), staged[:2 * k])
        seq = ScrambledSobolSeq.__new__(ScrambledSobolSeq)
        seq.D, seq.m, seq.x0 = 2 * k, m, x
        pts[i, :, :2 * k] = seq.points(N)
    stage_both(pts)
    eg = anderson()
    Po_g, Ps_g = qiw.inchworm(eg, eg.grid, orders, orders_bare, N, rand_params=Staged())
    P_c, Po_c, Ps_c = O.inchworm(anderson(), orders, orders_bare, N, rand=(N_SEQS, 3, 0, 0.0))
    assert rel(eg.P, P_c) < RTOL
    ek = anderson()
    Po_k, Ps_k = qiw.inchworm(ek, ek.grid, orders, orders_bare, N,
                              rand_params=qiw.RandomizationParams(SEED, N_SEQS, 0.0))
    np.testing.assert_array_equal(ek.P, eg.P)
    for o in Po_k:
        np.testing.assert_array_equal(Ps_k[o], Ps_g[o])


@pytest.mark.parametrize("make", [anderson, kanamori2], ids=["anderson", "kanamori2"])
def test_steps_with_random_seq(make):
    orders = [0, 1, 2, 3]
    e = make()
    P, _, _ = O.inchworm(e, [0, 1], [0, 1], N)
    print(f"points: RandomSeq bold step n_w=2 and bare step, {make.__name__}, orders 0:3 N={N}", flush=True)
    log = oracle_log(lambda r: O.inchworm_step(e, orders, N, 2, P=P, n_known=e.grid.n_tau, rand=r), 6)
    stage_both(random_points(log, 6))
    eg = make()
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    v_g, per_g = qiw.inchworm_step(eg, 2, orders, N, rand_params=Staged())
    v_c, per_c = O.inchworm_step(e, orders, N, 2, P=P, n_known=e.grid.n_tau, rand=(N_SEQS, 3, 0, 0.0))
    assert rel(v_g, v_c) < RTOL and rel(per_g, per_c) < RTOL
    log = oracle_log(lambda r: O.inchworm_step_bare(e, orders, N, rand=r), 6)
    stage_both(random_points(log, 6, seed=12))
    v_g, _ = qiw.inchworm_step_bare(eg, orders, N, rand_params=Staged())
    v_c, _ = O.inchworm_step_bare(e, orders, N, rand=(N_SEQS, 3, 0, 0.0))
    assert rel(v_g, v_c) < RTOL


def test_correlators_and_diff_with_random_seq():
    orders = [0, 1, 2]
    e = anderson(corr=True)
    P, _, _ = O.inchworm(e, orders, orders, N)
    print(f"points: RandomSeq correlators and diff_inchworm, anderson n_tau=5 N={N}", flush=True)
    log = oracle_log(lambda r: O.correlator_2p(e, orders, N, P=P, n_known=e.grid.n_tau, rand=r, with_std=True), 4)
    stage_both(random_points(log, 4))
    eg = anderson(corr=True)
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    g_g, s_g = qiw.correlator_2p(eg, eg.grid, orders, N, qiw.RequestStdDev(), rand_params=Staged())
    g_c, s_c = O.correlator_2p(e, orders, N, P=P, n_known=e.grid.n_tau, rand=(N_SEQS, 3, 0, 0.0), with_std=True)
    for a, b, sa, sb in zip(g_g, g_c, s_g, s_c):
        assert rel(a, b) < RTOL
        assert np.abs(np.asarray(sa) - np.asarray(sb)).max() <= RTOL * (np.abs(b).max() + np.abs(sb).max())
    log = oracle_log(lambda r: O.diff_inchworm(anderson(), [1, 2, 3], N, rand=r), 6)
    stage_both(random_points(log, 6))
    ed = anderson()
    S_g, _ = qiw.diff_inchworm(ed, ed.grid, [1, 2, 3], N, rand_params=Staged())
    P_c, S_c, _ = O.diff_inchworm(anderson(), [1, 2, 3], N, rand=(N_SEQS, 3, 0, 0.0))
    assert rel(ed.P, P_c) < RTOL
    for o in [1, 2, 3]:
        assert rel(S_g[o], S_c[o]) < RTOL


def test_python_api_seq_type():
    orders = [0, 1, 2]
    eg = anderson()
    P, _, _ = O.inchworm(eg, orders, orders, N)
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    print(f"points: Python API seq_type=RandomSeq, bold step anderson orders 0:2 N={N}", flush=True)
    v_g, _ = qiw.inchworm_step(eg, 2, orders, N, rand_params=qiw.RandomizationParams(
        np.random.default_rng(21), N_SEQS, 0.0, seq_type=RandomSeq))
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 3, 0, 0.0))
    dims = qiw.scramble_dims(run, "step")
    O.set_points(draw_points(np.random.default_rng(21), RandomSeq, dims, max(dims), N))
    v_c, _ = O.inchworm_step(eg, orders, N, 2, P=P, n_known=eg.grid.n_tau, rand=(N_SEQS, 3, 0, 0.0))
    assert O.scrambles_used() == len(dims)
    assert rel(v_g, v_c) < RTOL
    with pytest.raises(qiw.abi.QiwError):  # fewer points than N_samples
        qiw.set_points(np.full((len(dims), N - 1, max(dims)), 0.5))
        qiw.inchworm_step(eg, 2, orders, N, rand_params=Staged())
