"""GPU parity of caller-drawn scrambled sequences (qiw_rand.scramble = 2,
qiw_set_scrambles, qiw_scramble_dims) against the float64 oracle, through the C ABI.

* qiw_scramble_dims lists exactly the sequences the oracle -- following the reference's
  loops and its stopping rule (src/randomization.jl:93-99) -- constructs, groups that
  vanish identically included (zero-volume simplices, groups without a contributing
  diagram: the spin-flip correlator).
* With the same staged sequences the GPU equals the oracle to 1e-10 (inchworm!, the two
  steps, correlator_2p with deviations, diff_inchworm!), and equals its own keyed mode bit
  for bit when the staged sequences are the keyed ones.
* The Python API draws ScrambledSobolSeq(2k, scramble_rng=rng) in that order from a numpy
  Generator (RandomizationParams(..., caller_drawn=True))."""
import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import model
from qiw.randomization import draw_sequences

pytestmark = pytest.mark.gpu
RTOL = 1e-10
SEED = 0x5EED1234ABCDEF
N_SEQS = 3
N = 2 ** 5


class Staged:
    """rand_params for sequences the test stages itself (scramble = 2)."""
    caller_drawn = False

    def __init__(self, n_seqs=N_SEQS):
        self.N_seqs = n_seqs

    def struct_args(self):
        return (self.N_seqs, 2, 0, 0.0)


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
    """The slots (domain, key, order, n_pts_after, s) the oracle consumes for a call."""
    rng = np.random.default_rng(3)
    O.set_scrambles(rng.integers(0, 2 ** 32, size=(4000, stride, 32), dtype=np.uint32),
                    rng.integers(0, 2 ** 32, size=(4000, stride), dtype=np.uint32))
    run_oracle((N_SEQS, 2, 0, 0.0))
    return O.scrambles_log()


def keyed_arrays(log, stride):
    staged, _ = O.sobol_tables(stride)
    m = np.zeros((len(log), stride, 32), np.uint32)
    x = np.zeros((len(log), stride), np.uint32)
    for i, (dom, key, k, na, s) in enumerate(log):
        m[i, :2 * k], x[i, :2 * k] = O.rq_scramble(SEED, O.rq_stream(dom, key, k, na, s), staged[:2 * k])
    return m, x


def stage_both(m, x):
    O.set_scrambles(m, x)
    qiw.set_scrambles(m, x)


def dims_of(log):
    return [2 * k for (_, _, k, _, _) in log]


def test_inchworm_matches_oracle_and_keyed_mode():
    orders, orders_bare = [0, 1, 2], [0, 1, 2, 3]
    print(f"given: inchworm anderson n_tau=5 orders 0:2 bare 0:3 N={N}", flush=True)
    log = oracle_log(lambda r: O.inchworm(anderson(), orders, orders_bare, N, rand=r), 6)
    eg = anderson()
    run = qiw.abi.RunStruct(orders, orders_bare, N, 0, rand=(N_SEQS, 2, 0, 0.0))
    qiw.inchworm_step_bare(eg, [0], 0)  # stages the model
    assert qiw.scramble_dims(run, "inchworm") == dims_of(log)
    m, x = keyed_arrays(log, 6)
    stage_both(m, x)
    Po_g, Ps_g = qiw.inchworm(eg, eg.grid, orders, orders_bare, N, rand_params=Staged())
    P_c, Po_c, Ps_c = O.inchworm(anderson(), orders, orders_bare, N, rand=(N_SEQS, 2, 0, 0.0))
    assert rel(eg.P, P_c) < RTOL
    for o in Po_c:
        assert rel(Po_g[o], Po_c[o]) < RTOL
    ek = anderson()
    Po_k, Ps_k = qiw.inchworm(ek, ek.grid, orders, orders_bare, N,
                              rand_params=qiw.RandomizationParams(SEED, N_SEQS, 0.0))
    np.testing.assert_array_equal(ek.P, eg.P)
    for o in Po_k:
        np.testing.assert_array_equal(Ps_k[o], Ps_g[o])


@pytest.mark.parametrize("make", [anderson, kanamori2], ids=["anderson", "kanamori2"])
def test_steps_match_oracle(make):
    orders = [0, 1, 2, 3]
    e = make()
    P, _, _ = O.inchworm(e, [0, 1], [0, 1], N)
    print(f"given: bold step n_w=2 and bare step, {make.__name__}, orders 0:3 N={N}", flush=True)
    log = oracle_log(lambda r: O.inchworm_step(e, orders, N, 2, P=P, n_known=e.grid.n_tau, rand=r), 6)
    m, x = keyed_arrays(log, 6)
    stage_both(m, x)
    eg = make()
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    v_g, per_g = qiw.inchworm_step(eg, 2, orders, N, rand_params=Staged())
    v_c, per_c = O.inchworm_step(e, orders, N, 2, P=P, n_known=e.grid.n_tau, rand=(N_SEQS, 2, 0, 0.0))
    assert rel(v_g, v_c) < RTOL and rel(per_g, per_c) < RTOL
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 2, 0, 0.0))
    assert qiw.scramble_dims(run, "step") == dims_of(log)
    log = oracle_log(lambda r: O.inchworm_step_bare(e, orders, N, rand=r), 6)
    stage_both(*keyed_arrays(log, 6))
    v_g, _ = qiw.inchworm_step_bare(eg, orders, N, rand_params=Staged())
    v_c, _ = O.inchworm_step_bare(e, orders, N, rand=(N_SEQS, 2, 0, 0.0))
    assert rel(v_g, v_c) < RTOL


def test_correlators_with_vanishing_groups():
    """Every tau row draws its own sequences; the row tau_A = beta and the spin-flip pair's
    groups with an odd number of operators between B and A vanish and take two."""
    orders = [0, 1, 2]
    e = anderson(corr=True)
    P, _, _ = O.inchworm(e, orders, orders, N)
    print(f"given: correlators anderson n_tau=5 orders 0:2 N={N}", flush=True)
    log = oracle_log(lambda r: O.correlator_2p(e, orders, N, P=P, n_known=e.grid.n_tau, rand=r, with_std=True), 4)
    m, x = keyed_arrays(log, 4)
    stage_both(m, x)
    eg = anderson(corr=True)
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 2, 0, 0.0))
    qiw.correlator_2p(eg, eg.grid, [0], 0)  # stages the model
    assert qiw.scramble_dims(run, "correlator") == dims_of(log)
    g_g, s_g = qiw.correlator_2p(eg, eg.grid, orders, N, qiw.RequestStdDev(), rand_params=Staged())
    g_c, s_c = O.correlator_2p(e, orders, N, P=P, n_known=e.grid.n_tau, rand=(N_SEQS, 2, 0, 0.0), with_std=True)
    for a, b, sa, sb in zip(g_g, g_c, s_g, s_c):
        assert rel(a, b) < RTOL
        assert np.abs(np.asarray(sa) - np.asarray(sb)).max() <= RTOL * (np.abs(b).max() + np.abs(sb).max())


def test_diff_matches_oracle():
    orders = [1, 2, 3]
    print(f"given: diff_inchworm anderson n_tau=5 orders 1:3 N={N}", flush=True)
    log = oracle_log(lambda r: O.diff_inchworm(anderson(), orders, N, rand=r), 6)
    m, x = keyed_arrays(log, 6)
    stage_both(m, x)
    eg = anderson()
    S_g, Ss_g = qiw.diff_inchworm(eg, eg.grid, orders, N, rand_params=Staged())
    P_c, S_c, Ss_c = O.diff_inchworm(anderson(), orders, N, rand=(N_SEQS, 2, 0, 0.0))
    assert rel(eg.P, P_c) < RTOL
    for o in orders:
        assert rel(S_g[o], S_c[o]) < RTOL
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 2, 0, 0.0))
    assert qiw.scramble_dims(run, "diff") == dims_of(log)


def test_python_api_draws_in_reference_order():
    """RandomizationParams(rng, N_seqs, 0, caller_drawn=True): the sequences drawn from the
    Generator, in qiw_scramble_dims order, are the ones the oracle consumes."""
    orders = [0, 1, 2]
    eg = anderson()
    P, _, _ = O.inchworm(eg, orders, orders, N)
    eg.P, eg.P_n_known = P.copy(), eg.grid.n_tau
    print(f"given: Python API, bold step anderson orders 0:2 N={N}", flush=True)
    v_g, _ = qiw.inchworm_step(eg, 2, orders, N, rand_params=qiw.RandomizationParams(
        np.random.default_rng(11), N_SEQS, 0.0, caller_drawn=True))
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 2, 0, 0.0))
    dims = qiw.scramble_dims(run, "step")
    m, x = draw_sequences(np.random.default_rng(11), dims, max(dims))
    O.set_scrambles(m, x)
    v_c, _ = O.inchworm_step(eg, orders, N, 2, P=P, n_known=eg.grid.n_tau, rand=(N_SEQS, 2, 0, 0.0))
    assert O.scrambles_used() == len(dims)
    assert rel(v_g, v_c) < RTOL


def test_errors():
    orders = [0, 1, 2]
    e = anderson()
    P, _, _ = O.inchworm(e, orders, orders, N)
    e.P, e.P_n_known = P.copy(), e.grid.n_tau
    run = qiw.abi.RunStruct(orders, [], N, 0, rand=(N_SEQS, 2, 0, 0.0))
    qiw.inchworm_step(e, 2, [0], 0)  # stages the model
    n = len(qiw.scramble_dims(run, "step"))
    m = np.zeros((n - 1, 4, 32), np.uint32)
    qiw.set_scrambles(m, None)
    with pytest.raises(qiw.abi.QiwError):  # one sequence short
        qiw.inchworm_step(e, 2, orders, N, rand_params=Staged())
    qiw.set_scrambles(np.zeros((n, 2, 32), np.uint32), None)
    with pytest.raises(qiw.abi.QiwError):  # order 2 needs 4 dimensions
        qiw.inchworm_step(e, 2, orders, N, rand_params=Staged())
    with pytest.raises(ValueError):
        qiw.RandomizationParams(None, 2, 0.0, caller_drawn=True)
