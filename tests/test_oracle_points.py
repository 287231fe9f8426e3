"""Caller-supplied point sequences (qiw_rand.scramble = 3) on the oracle: the sequence plug
point seq_type of mean_std_from_randomization (src/randomization.jl:82-99; inchworm!'s
seq_type keyword, src/inchworm.jl:129,233,339).

* Staging, slot for slot, the points of the keyed mode's scrambled Sobol sequences (their
  first N points, x_i = x_0 xor the direction numbers of gray(i)) reproduces the keyed
  result bit for bit for every call kind, and consumes exactly the staged sequences: the
  points mode reads the caller's points where the Sobol mode generates them, nothing else.
* RandomSeq (src/utility.jl:336-414) at the plug point: its points are the generator's
  uniform draws in order, skip! skips 2^m points (exact=False) or n, and a randomized
  result built from RandomSeq points agrees with the unscrambled reference within its own
  standard deviation.
* Too few points per sequence, or too few sequences, fail."""
import numpy as np
import pytest

import oracle_lib as O
from qiw import model
from qiw.randomization import RandomSeq, draw_points
from qiw.sobol import ScrambledSobolSeq

SEED = 0x5EED1234ABCDEF
N_SEQS = 3
N = 2 ** 4


def anderson(n_tau=5, corr=False):
    return model.anderson_single_orbital(beta=2.0, n_tau=n_tau, with_correlators=corr)


def sobol_points(m, x, n):
    s = ScrambledSobolSeq.__new__(ScrambledSobolSeq)
    s.D, s.m, s.x0 = m.shape[0], m, x
    return s.points(n)


def keyed_points(log, stride, n):
    staged, _ = O.sobol_tables(stride)
    pts = np.zeros((len(log), n, stride))
    for i, (dom, key, k, na, s) in enumerate(log):
        m, x = O.rq_scramble(SEED, O.rq_stream(dom, key, k, na, s), staged[:2 * k])
        pts[i, :, :2 * k] = sobol_points(m, x, n)
    return pts


def log_of(run, stride):
    rng = np.random.default_rng(3)
    O.set_scrambles(rng.integers(0, 2 ** 32, size=(20000, stride, 32), dtype=np.uint32),
                    rng.integers(0, 2 ** 32, size=(20000, stride), dtype=np.uint32))
    run((N_SEQS, 2, 0, 0.0))
    return O.scrambles_log()


def assert_same(a, b):
    if isinstance(a, dict):
        for k in a:
            assert_same(a[k], b[k])
    elif isinstance(a, (list, tuple)):
        for x, y in zip(a, b):
            assert_same(x, y)
    else:
        np.testing.assert_array_equal(a, b)


def check_keyed_equivalence(run, stride):
    log = log_of(run, stride)
    O.set_points(keyed_points(log, stride, N))
    given = run((N_SEQS, 3, 0, 0.0))
    assert O.scrambles_used() == len(log)
    keyed = run((N_SEQS, 1, SEED, 0.0))
    assert_same(keyed, given)


def test_inchworm_and_steps():
    check_keyed_equivalence(lambda r: O.inchworm(anderson(), [0, 1, 2], [0, 1, 2, 3], N, rand=r), 6)
    e = anderson()
    P, _, _ = O.inchworm(e, [0, 1], [0, 1], N)
    check_keyed_equivalence(lambda r: O.inchworm_step(e, [0, 1, 2, 3], N, 2, P=P, n_known=e.grid.n_tau, rand=r), 6)
    check_keyed_equivalence(lambda r: O.inchworm_step_bare(e, [0, 1, 2], N, rand=r), 4)


def test_correlator_and_diff():
    e = anderson(corr=True)
    P, _, _ = O.inchworm(e, [0, 1, 2], [0, 1, 2], N)
    check_keyed_equivalence(
        lambda r: O.correlator_2p(e, [0, 1, 2], N, P=P, n_known=e.grid.n_tau, rand=r, with_std=True), 4)
    check_keyed_equivalence(lambda r: O.diff_inchworm(anderson(), [1, 2, 3], N, rand=r), 6)


def test_random_seq():
    s = RandomSeq(3, scramble_rng=np.random.default_rng(5))
    seed = int(np.random.default_rng(5).integers(0, 2 ** 32, dtype=np.uint64))
    ref = np.random.Generator(np.random.PCG64(seed)).random((9, 3))
    np.testing.assert_array_equal(s.next(), ref[0])
    s.skip(2)  # 2 < 2 + 1 <= 4: skips 2 points
    np.testing.assert_array_equal(s.next(), ref[3])
    s.skip(3, exact=True)
    np.testing.assert_array_equal(s.points(2), ref[7:9])
    np.testing.assert_array_equal(RandomSeq(2).points(4), np.random.Generator(np.random.PCG64(0)).random((4, 2)))
    pts = draw_points(np.random.default_rng(1), RandomSeq, [2, 4], 4, 8)
    assert pts.shape == (2, 8, 4) and np.all(pts[0, :, 2:] == 0) and np.all((pts >= 0) & (pts < 1))


def test_random_seq_estimate_within_its_deviation():
    """One bold step with RandomSeq points (8 sequences of 2^9 points) against the
    unscrambled step at 2^14: the difference is within 5 standard deviations of the mean
    (the deviation is the reference's, over the sequences, divided by sqrt(N_seqs))."""
    e = anderson()
    P, _, _ = O.inchworm(e, [0, 1, 2], [0, 1, 2], 2 ** 8)
    orders, n_seqs, n = [0, 1, 2], 8, 2 ** 9
    ref, _ = O.inchworm_step(e, orders, 2 ** 14, 2, P=P, n_known=e.grid.n_tau)
    log = log_of(lambda r: O.inchworm_step(e, orders, n, 2, P=P, n_known=e.grid.n_tau,
                                           rand=(n_seqs, 2, 0, 0.0)), 4)
    assert len(log) == n_seqs * 4  # order 1: one group; order 2: n_pts_after 1, 2, 3
    O.set_points(draw_points(np.random.default_rng(9), RandomSeq, [2 * k for (_, _, k, _, _) in log], 4, n))
    v, _ = O.inchworm_step(e, orders, n, 2, P=P, n_known=e.grid.n_tau, rand=(n_seqs, 3, 0, 0.0))
    assert O.scrambles_used() == len(log)
    err = np.abs(np.asarray(v) - np.asarray(ref)).max()
    assert 0 < err < 5e-2 * np.abs(ref).max()


def test_too_few_points_or_sequences():
    e = anderson()
    log = log_of(lambda r: O.inchworm_step_bare(e, [0, 1, 2], N, rand=r), 4)
    rng = np.random.default_rng(4)  # distinct sequences: no group stops after its second
    O.set_points(rng.random((len(log), N - 1, 4)))
    with pytest.raises(Exception):
        O.inchworm_step_bare(e, [0, 1, 2], N, rand=(N_SEQS, 3, 0, 0.0))
    O.set_points(rng.random((len(log) - 1, N, 4)))
    with pytest.raises(Exception):
        O.inchworm_step_bare(e, [0, 1, 2], N, rand=(N_SEQS, 3, 0, 0.0))
