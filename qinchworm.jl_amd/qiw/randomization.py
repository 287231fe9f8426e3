"""Randomized quasi Monte Carlo parameters (src/randomization.jl:46-69).

RandomizationParams(rng, N_seqs, target_std) selects how many scrambled Sobol sequences
integrate each (order, n_pts_after) group and when to stop early; RequestStdDev selects
the correlator method that also returns standard deviations.  The scramble bits of the
sequences come from libqiw's counter-based generator keyed by a 64-bit seed; with a numpy
Generator as rng every call draws a fresh seed from it (the reference draws the bits
themselves from its rng, src/scrambled_sobol.jl:104-109).
"""
import numpy as np


class RandomizationParams:
    """rng: None (unscrambled: every sequence is the staged one), a numpy Generator (a seed
    is drawn from it per call) or an int (the seed itself)."""

    def __init__(self, rng=None, N_seqs=1, target_std=0.0):
        self.rng = rng
        self.N_seqs = int(N_seqs)
        self.target_std = float(target_std)

    def __repr__(self):
        return f"RandomizationParams(rng={self.rng}, N_seqs={self.N_seqs}, target_std={self.target_std})"

    def struct_args(self):
        """(N_seqs, scramble, seed, target_std) for one libqiw call."""
        if self.rng is None:
            return (self.N_seqs, 0, 0, self.target_std)
        if isinstance(self.rng, (int, np.integer)):
            seed = int(self.rng)
        else:
            seed = int(self.rng.integers(0, 2 ** 64, dtype=np.uint64))
        return (self.N_seqs, 1, seed, self.target_std)


class RequestStdDev:
    """Selects the correlator_2p method that returns standard deviations too
    (src/randomization.jl:64-69)."""
