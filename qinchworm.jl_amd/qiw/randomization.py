"""Randomized quasi Monte Carlo parameters (src/randomization.jl:46-69).

RandomizationParams(rng, N_seqs, target_std) selects how many scrambled Sobol sequences
integrate each (order, n_pts_after) group and when to stop early; RequestStdDev selects
the correlator method that also returns standard deviations.  By default the scramble
bits of the sequences come from libqiw's counter-based generator keyed by a 64-bit seed;
with a numpy Generator as rng every call draws a fresh seed from it.  With
caller_drawn=True (target_std = 0) the sequences are drawn from rng itself, one
ScrambledSobolSeq(2k, scramble_rng=rng) per sequence in the reference's order
(src/randomization.jl:94, src/scrambled_sobol.jl:104-109), and staged with
qiw_set_scrambles.  With seq_type (the reference's sequence plug point,
src/randomization.jl:82-94, src/inchworm.jl:129,233,339) any sequence type takes the place
of ScrambledSobolSeq: seq_type(D, scramble_rng=rng) is constructed in the same order, its
first N_samples points (its `points(n)` method, n successive next! of the reference) are
staged with qiw_set_points and integrated on the GPU (qiw_rand.scramble = 3).  RandomSeq
is the reference's pseudo-random one (src/utility.jl:336-414).
"""
import numpy as np

from .sobol import ScrambledSobolSeq


class RandomizationParams:
    """rng: None (unscrambled: every sequence is the staged one), a numpy Generator (a seed
    is drawn from it per call) or an int (the seed itself)."""

    def __init__(self, rng=None, N_seqs=1, target_std=0.0, caller_drawn=False, seq_type=None):
        self.rng = rng
        self.N_seqs = int(N_seqs)
        self.target_std = float(target_std)
        self.seq_type = None if seq_type is ScrambledSobolSeq else seq_type
        self.caller_drawn = bool(caller_drawn) or seq_type is not None
        if self.caller_drawn and (rng is None or isinstance(rng, (int, np.integer))):
            raise ValueError("caller_drawn sequences need a numpy Generator as rng")

    def __repr__(self):
        return f"RandomizationParams(rng={self.rng}, N_seqs={self.N_seqs}, target_std={self.target_std})"

    def struct_args(self):
        """(N_seqs, scramble, seed, target_std) for one libqiw call."""
        if self.rng is None:
            return (self.N_seqs, 0, 0, self.target_std)
        if self.caller_drawn:
            return (self.N_seqs, 2 if self.seq_type is None else 3, 0, self.target_std)
        if isinstance(self.rng, (int, np.integer)):
            seed = int(self.rng)
        else:
            seed = int(self.rng.integers(0, 2 ** 64, dtype=np.uint64))
        return (self.N_seqs, 1, seed, self.target_std)


class RequestStdDev:
    """Selects the correlator_2p method that returns standard deviations too
    (src/randomization.jl:64-69)."""


def draw_sequences(rng, dims, stride):
    """The caller's sequences of one call: ScrambledSobolSeq(D, scramble_rng=rng) for each
    D of dims in turn, as (m [n][stride][32], shift [n][stride]) uint32 arrays (dimensions
    past D zero)."""
    n = len(dims)
    m = np.zeros((max(n, 1), stride, 32), np.uint32)
    x = np.zeros((max(n, 1), stride), np.uint32)
    for i, D in enumerate(dims):
        s = ScrambledSobolSeq(int(D), scramble_rng=rng)
        m[i, :D] = s.m
        x[i, :D] = s.x0
    return m[:n], x[:n]


def draw_points(rng, seq_type, dims, stride, n_points):
    """The caller's point sequences of one call: seq_type(D, scramble_rng=rng) for each D of
    dims in turn, its first n_points points, as an array [n][n_points][stride] (dimensions
    past D zero)."""
    n = len(dims)
    x = np.zeros((max(n, 1), max(int(n_points), 1), stride))
    for i, D in enumerate(dims):
        s = seq_type(int(D), scramble_rng=rng)
        x[i, :n_points, :D] = s.points(int(n_points))
    return x[:n]


class RandomSeq:
    """RandomSeq{RNG} (src/utility.jl:336-414): a D-dimensional pseudo-random sequence, the
    reference's stand-in at the sequence plug point.  Its generator is seeded with one
    32-bit integer drawn from scramble_rng (0 without one); here the generator is numpy's
    PCG64 and a point is D successive uniform doubles in [0, 1) (rand!)."""

    def __init__(self, D, scramble_rng=None):
        self.D = int(D)
        seed = 0 if scramble_rng is None else int(scramble_rng.integers(0, 2 ** 32, dtype=np.uint64))
        self.rng = np.random.Generator(np.random.PCG64(seed))
        self.n = 0  # points drawn so far

    def next(self):
        """next! (:375-385)."""
        self.n += 1
        return self.rng.random(self.D)

    def skip(self, n, exact=False):
        """skip! (:395-414): n points, or with exact=False the 2^m points with
        2^m < n + 1 <= 2^(m+1)."""
        if n < 0:
            raise ValueError(f"{n} is not non-negative")
        if n == 0:
            return self
        k = n if exact else 1 << int(np.floor(np.log2(n + 1)))
        self.rng.random((k, self.D))
        self.n += k
        return self

    def points(self, n, first=None):
        """The next n points (first: the index of the first, >= the points drawn so far)."""
        if first is not None:
            if first < self.n:
                raise ValueError("RandomSeq points are drawn in order")
            self.skip(first - self.n, exact=True)
        self.n += n
        return self.rng.random((n, self.D))
