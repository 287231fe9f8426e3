"""Randomized quasi Monte Carlo parameters (src/randomization.jl:46-69).

RandomizationParams(rng, N_seqs, target_std) selects how many scrambled Sobol sequences
integrate each (order, n_pts_after) group and when to stop early; RequestStdDev selects
the correlator method that also returns standard deviations.  By default the scramble
bits of the sequences come from libqiw's counter-based generator keyed by a 64-bit seed;
with a numpy Generator as rng every call draws a fresh seed from it.  With
caller_drawn=True (target_std = 0) the sequences are drawn from rng itself, one
ScrambledSobolSeq(2k, scramble_rng=rng) per sequence in the reference's order
(src/randomization.jl:94, src/scrambled_sobol.jl:104-109), and staged with
qiw_set_scrambles.
"""
import numpy as np

from .sobol import ScrambledSobolSeq


class RandomizationParams:
    """rng: None (unscrambled: every sequence is the staged one), a numpy Generator (a seed
    is drawn from it per call) or an int (the seed itself)."""

    def __init__(self, rng=None, N_seqs=1, target_std=0.0, caller_drawn=False):
        self.rng = rng
        self.N_seqs = int(N_seqs)
        self.target_std = float(target_std)
        self.caller_drawn = bool(caller_drawn)
        if self.caller_drawn and (rng is None or isinstance(rng, (int, np.integer))):
            raise ValueError("caller_drawn sequences need a numpy Generator as rng")

    def __repr__(self):
        return f"RandomizationParams(rng={self.rng}, N_seqs={self.N_seqs}, target_std={self.target_std})"

    def struct_args(self):
        """(N_seqs, scramble, seed, target_std) for one libqiw call."""
        if self.rng is None:
            return (self.N_seqs, 0, 0, self.target_std)
        if self.caller_drawn:
            return (self.N_seqs, 2, 0, self.target_std)
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
