"""Caller-drawn scrambled sequences (qiw_rand.scramble = 2) on the oracle.

The reference constructs one ScrambledSobolSeq(2k; scramble_rng = rng) per sequence of a
sampled group, in the order of its loops (src/inchworm.jl:142-195 bold step, 246-300
initial step, 552-600 differential step, 825-890 correlator per tau), and stops a group
after sequence s >= 2 once the deviation is <= target_std (src/randomization.jl:93-99).
The oracle follows those loops and consumes the caller's sequences as it constructs them,
logging the slot (domain, key, order, n_pts_after, s) of each.  The checks:

* the log has the loops' structure: groups in the reference's order, N_seqs sequences
  each, two where a group's integral vanishes identically (a zero-volume simplex: the
  correlator row tau_A = beta, the first differential step; or no contributing diagram);
* staging, in the logged order, exactly the sequences the keyed mode (scramble = 1) uses
  for those slots reproduces the keyed result bit for bit, and consumes all of them."""
import numpy as np
import pytest

import oracle_lib as O
from qiw import model
from qiw.diagrammatics import get_topologies_at_order

SEED = 0x5EED1234ABCDEF
RQ_BARE, RQ_BOLD, RQ_CORR, RQ_DIFF = 1, 2, 3, 4
N_SEQS = 3


def stage_arbitrary(n, stride):
    """n unrelated sequences (the log does not depend on the points)."""
    rng = np.random.default_rng(7)
    O.set_scrambles(rng.integers(0, 2 ** 32, size=(n, stride, 32), dtype=np.uint32),
                    rng.integers(0, 2 ** 32, size=(n, stride), dtype=np.uint32))


def stage_keyed(keys, stride):
    """The caller-drawn list equal to the keyed sequences of the slots: direction numbers of
    2k dimensions scrambled with the slot stream's bits (dimensions past 2k zero)."""
    staged, _ = O.sobol_tables(stride)
    m = np.zeros((len(keys), stride, 32), np.uint32)
    x = np.zeros((len(keys), stride), np.uint32)
    for i, (dom, key, k, na, s) in enumerate(keys):
        mo, xo = O.rq_scramble(SEED, O.rq_stream(dom, key, k, na, s), staged[:2 * k])
        m[i, :2 * k], x[i, :2 * k] = mo, xo
    O.set_scrambles(m, x)


def record_then_replay(run, stride):
    """run(rand) -> results; returns (log, keyed results, given results)."""
    stage_arbitrary(20000, stride)
    run((N_SEQS, 2, 0, 0.0))
    log = O.scrambles_log()
    stage_keyed(log, stride)
    given = run((N_SEQS, 2, 0, 0.0))
    assert O.scrambles_used() == len(log)
    keyed = run((N_SEQS, 1, SEED, 0.0))
    return log, keyed, given


def assert_same(a, b):
    if isinstance(a, dict):
        for k in a:
            assert_same(a[k], b[k])
    elif isinstance(a, (list, tuple)):
        for x, y in zip(a, b):
            assert_same(x, y)
    else:
        np.testing.assert_array_equal(a, b)


def structure(log):
    """[(domain, key, order, n_pts_after, number of sequences)] in log order; checks that
    each group's sequences are s = 0, 1, ... in turn."""
    out = []
    for dom, key, k, na, s in log:
        if s == 0:
            out.append([dom, key, k, na, 1])
        else:
            assert out and out[-1][:4] == [dom, key, k, na] and out[-1][4] == s
            out[-1][4] += 1
    return [tuple(g) for g in out]


def bold_groups(orders):
    return [(k, na) for k in orders if k > 0 for na in range(1, 2 * k) if get_topologies_at_order(k, na)]


def anderson(n_tau=5, corr=False):
    return model.anderson_single_orbital(beta=2.0, n_tau=n_tau, with_correlators=corr)


def test_inchworm():
    orders, orders_bare = [0, 1, 2], [0, 1, 2, 3]
    log, keyed, given = record_then_replay(lambda r: O.inchworm(anderson(), orders, orders_bare, 2 ** 4, rand=r), 6)
    assert_same(keyed, given)
    want = [(RQ_BARE, 0, k, 2 * k, N_SEQS) for k in orders_bare if k > 0]
    for w in range(1, 4):
        want += [(RQ_BOLD, w, k, na, N_SEQS) for k, na in bold_groups(orders)]
    assert structure(log) == want


def test_steps():
    e = anderson()
    orders = [0, 1, 2, 3]
    P, _, _ = O.inchworm(e, [0, 1], [0, 1], 2 ** 4)
    log, keyed, given = record_then_replay(
        lambda r: O.inchworm_step(e, orders, 2 ** 4, 2, P=P, n_known=e.grid.n_tau, rand=r), 6)
    assert_same(keyed, given)
    assert structure(log) == [(RQ_BOLD, 2, k, na, N_SEQS) for k, na in bold_groups(orders)]
    log, keyed, given = record_then_replay(lambda r: O.inchworm_step_bare(e, orders, 2 ** 4, rand=r), 6)
    assert_same(keyed, given)
    assert structure(log) == [(RQ_BARE, 0, k, 2 * k, N_SEQS) for k in orders if k > 0]


def test_correlator_rows_draw_their_own_sequences():
    """Per tau row the reference constructs new sequences; the keyed mode shares one set
    per pair, which staged once per row reproduces it.  At tau_A = beta the groups with
    points after tau_A vanish and stop after two sequences."""
    e = anderson(corr=True)
    P, _, _ = O.inchworm(e, [0, 1, 2], [0, 1, 2], 2 ** 4)
    orders = [0, 1, 2]
    log, keyed, given = record_then_replay(
        lambda r: O.correlator_2p(e, orders, 2 ** 4, P=P, n_known=e.grid.n_tau, rand=r, with_std=True), 4)
    assert_same(keyed, given)
    st = structure(log)
    groups = [(k, na) for k in orders if k > 0 for na in range(1, 2 * k)
              if get_topologies_at_order(k, na, with_external_arc=True)]
    nc, nt = len(e.corr_operators), e.grid.n_tau
    assert len(st) == nc * (nt - 1) * len(groups)
    i = 0
    for c in range(nc):
        for j in range(1, nt):
            for k, na in groups:
                dom, key, kk, nna, n = st[i]
                assert (dom, key, kk, nna) == (RQ_CORR, c, k, na)
                assert n == (2 if j == nt - 1 else n), "every group vanishes at tau_A = beta"
                assert n in (2, N_SEQS)
                i += 1
    # Inside the interval only the spin-flip pair (S+(tau) S-(0), pair 2) has vanishing
    # groups: those with an odd number of hybridization operators between B and A
    # (d_before = 2k - n_pts_after odd) change the particle number there, which no sector
    # path allows.
    inner = {(c, k, na) for i, (_, c, k, na, n) in enumerate(st) if n == 2 and (i // len(groups)) % (nt - 1) != nt - 2}
    assert inner == {(2, k, na) for k, na in groups if (2 * k - na) % 2 == 1}


def test_diff():
    orders = [1, 2, 3]
    log, keyed, given = record_then_replay(lambda r: O.diff_inchworm(anderson(), orders, 2 ** 4, rand=r), 6)
    assert_same(keyed, given)
    want = []
    for a in range(4):  # tau_f_prev = grid[a]; at a = 0 the simplices have zero volume
        want += [(RQ_DIFF, a, k, 1, 2 if a == 0 else N_SEQS) for k in orders
                 if k >= 2 and get_topologies_at_order(k, 1)]
    assert structure(log) == want


def test_too_few_sequences_fail():
    e = anderson()
    stage_arbitrary(20000, 4)
    O.inchworm_step_bare(e, [0, 1, 2], 2 ** 4, rand=(N_SEQS, 2, 0, 0.0))
    log = O.scrambles_log()
    stage_keyed(log[:-1], 4)
    with pytest.raises(Exception):
        O.inchworm_step_bare(e, [0, 1, 2], 2 ** 4, rand=(N_SEQS, 2, 0, 0.0))
