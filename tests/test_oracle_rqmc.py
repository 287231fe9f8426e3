"""Oracle randomized qMC (src/randomization.jl:86-100) and its scramble bits.

The scramble bits are the ABI's own generator (qiw_rand in include/qiw.h); the LMS +
shift scramble they feed is pinned by the reference's MockRNG vectors through
qiw.sobol.lms_scramble (tests/test_oracle_sobol.py).  The mean and the corrected standard
deviation over the sequences, and the stopping rule, are checked against numpy on
single-sequence oracle runs with the same scrambled sequences."""
import numpy as np
import pytest

import oracle_lib as O
from qiw import model
from qiw.sobol import direction_numbers, lms_scramble

M64 = 2 ** 64 - 1


def mix(z):  # SplitMix64's output step after its increment
    z = (z + 0x9E3779B97F4A7C15) & M64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def word(seed, stream, d, w):
    return mix((mix(seed ^ mix(stream)) + 33 * d + w) & M64) & 0xFFFFFFFF


def stream(domain, key, order, n_after, seq):
    return (domain << 60) | (key << 32) | (order << 24) | (n_after << 16) | seq


def bits(seed, st, D):
    """The shift and matrix bits as rand(rng, [0,1], D, 32) / (D, 32, 32) would lay them out."""
    shift = np.zeros((D, 32), np.uint8)
    ltm = np.zeros((D, 32, 32), np.uint8)
    for d in range(D):
        w0 = word(seed, st, d, 0)
        shift[d] = [(w0 >> j) & 1 for j in range(32)]
        for p in range(32):
            wp = word(seed, st, d, 1 + p)
            for q in range(p):
                ltm[d, p, q] = (wp >> (31 - q)) & 1
            ltm[d, p, p + 1:] = 1  # unused by the scramble (only the strict lower triangle is)
    return ltm, shift


def test_stream_layout():
    for args in [(1, 0, 2, 4, 0), (2, 17, 3, 5, 7), (3, 1, 6, 11, 63)]:
        assert O.rq_stream(*args) == stream(*args)


@pytest.mark.parametrize("seed,st", [(0, stream(1, 0, 2, 4, 0)), (0xDEADBEEF12345678, stream(2, 9, 3, 2, 5))])
def test_scramble_matches_lms_with_generated_bits(seed, st):
    D = 4
    m = direction_numbers(D)
    mo, xo = O.rq_scramble(seed, st, m)
    ltm, shift = bits(seed, st, D)
    m_ref, x_ref = lms_scramble(m, ltm, shift)
    assert np.array_equal(mo, m_ref) and np.array_equal(xo, x_ref)


def test_scrambles_differ_per_sequence_and_keep_sobol_structure():
    m = direction_numbers(6)
    a, xa = O.rq_scramble(7, stream(2, 3, 3, 2, 0), m)
    b, xb = O.rq_scramble(7, stream(2, 3, 3, 2, 1), m)
    assert not np.array_equal(a, b) and not np.array_equal(xa, xb)
    # a linear matrix scramble keeps the direction numbers' leading bits: column j of a
    # Sobol dimension has its highest set bit at 31 - j, and so has the scrambled one
    for d in range(6):
        for j in range(32):
            assert int(a[d, j]).bit_length() == int(m[d, j]).bit_length()


def rqmc(n_seqs, scramble, seed, target):
    return (n_seqs, scramble, seed, target)


def bare_model():
    return model.anderson_single_orbital(n_tau=8)


def test_bare_step_mean_std_against_numpy():
    """Initial step, orders 0:2: each sampled order is one group, so its contribution over
    N_seqs sequences is numpy's mean and std(ddof=1) of single-sequence runs."""
    exp = bare_model()
    N, seed = 2 ** 7, 12345
    P, Po, Ps = O.inchworm(exp, [], [0, 1, 2], N, n_steps_max=0, rand=rqmc(4, 1, seed, 0.0))
    m = direction_numbers(4)
    for k in (1, 2):
        ys = []
        for s in range(4):
            ms, xs = O.rq_scramble(seed, stream(1, 0, k, 2 * k, s), m[:2 * k])
            _, per = O.inchworm_step_bare(exp, [k], N, sobol=(ms, xs))
            ys.append(per[0])
        ys = np.array(ys)
        np.testing.assert_allclose(Po[k][1], ys.mean(axis=0), rtol=1e-13, atol=0)
        sd = np.std(ys, axis=0, ddof=1)
        np.testing.assert_allclose(Ps[k][1].real, sd, rtol=1e-12, atol=0)
        assert np.all(Ps[k][1].imag == 0) and np.all(sd > 0)
    assert np.all(Ps[0] == 0)  # order 0 is exact


def test_stopping_rule():
    """target_std above every deviation stops after two sequences; 0 uses all of them."""
    exp = bare_model()
    N, seed = 2 ** 6, 99
    _, Po4, _ = O.inchworm(exp, [], [0, 1], N, n_steps_max=0, rand=rqmc(4, 1, seed, 0.0))
    _, Po2, Ps2 = O.inchworm(exp, [], [0, 1], N, n_steps_max=0, rand=rqmc(2, 1, seed, 0.0))
    _, Pob, Psb = O.inchworm(exp, [], [0, 1], N, n_steps_max=0, rand=rqmc(4, 1, seed, 1e3))
    np.testing.assert_array_equal(Pob[1], Po2[1])
    np.testing.assert_array_equal(Psb[1], Ps2[1])
    assert not np.allclose(Po4[1][1], Po2[1][1], rtol=1e-14, atol=0)


def test_unscrambled_sequences():
    """rng = nothing: every sequence is the staged one, so the mean is the one-sequence
    value and the deviation 0 (one sequence: NaN)."""
    exp = bare_model()
    P1, Po1, Ps1 = O.inchworm(exp, [0, 1, 2], [0, 1, 2], 2 ** 6, n_steps_max=2)
    P3, Po3, Ps3 = O.inchworm(exp, [0, 1, 2], [0, 1, 2], 2 ** 6, n_steps_max=2, rand=rqmc(3, 0, 0, 0.0))
    np.testing.assert_array_equal(P1, P3)
    for o in (1, 2):
        assert np.all(np.isnan(Ps1[o][1:4])) and np.all(Ps3[o][1:4] == 0)
        np.testing.assert_array_equal(Po1[o], Po3[o])


def test_correlator_std_against_numpy():
    """Correlator, orders 0:1: one sampled group per tau row; the value and deviation over
    three sequences are numpy's mean and std(ddof=1) of one-sequence correlators."""
    exp = model.anderson_single_orbital(n_tau=6, with_correlators=True)
    N, seed = 2 ** 6, 4242
    g, gs = O.correlator_2p(exp, [0, 1], N, rand=rqmc(3, 1, seed, 0.0), with_std=True)
    m = direction_numbers(2)
    for c in range(g.shape[0]):
        ys = []
        for s in range(3):
            ms, xs = O.rq_scramble(seed, stream(3, c, 1, 1, s), m)
            ys.append(O.correlator_2p(exp, [0, 1], N, sobol=(ms, xs))[c])
        ys = np.array(ys)
        np.testing.assert_allclose(g[c], ys.mean(axis=0), rtol=1e-12, atol=1e-15)
        # where the sampled group vanishes (zero volume at tau = beta, or no diagrams for the
        # pair) numpy's deviation is the rounding of the mean of three equal values
        np.testing.assert_allclose(gs[c, 1:].real, np.std(ys, axis=0, ddof=1)[1:], rtol=1e-10, atol=1e-15)
        assert np.max(gs[0, 1:-1].real) > 1e-5
        assert gs[c, 0] == 0
