"""Host side of the Sobol sequence: direction numbers and LMS+shift scrambling parameters.

Mirrors the constructor of ScrambledSobolSeq (src/scrambled_sobol.jl:66-143), whose state
(m, x) is what the C ABI takes (qiw_set_sobol); point generation itself runs on the GPU.
"""
import json
import os

import numpy as np

_TABLE = None


def _table():
    global _TABLE
    if _TABLE is None:
        with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "sobol_joe_kuo.json")) as f:
            t = json.load(f)
        _TABLE = (t["poly"], t["vinit"])
    return _TABLE


def direction_numbers(D):
    """m[d, j] of the reference after :78-100 (unscrambled), shape (D, 32), uint32.

    Dimension 1 is van der Corput (all initial m = 1); dimension d >= 2 uses primitive
    polynomial a = poly[d-1] (leading bit included, degree floor(log2 a)) and initial
    values vinit[d-1, :deg], then the recurrence of src/scrambled_sobol.jl:85-93."""
    poly, vinit = _table()
    if D < 0 or D > len(poly):
        raise ValueError(f"Invalid Sobol dimension {D}")
    m = np.ones((D, 32), dtype=np.uint64)
    for d in range(2, D + 1):
        a = int(poly[d - 1])
        deg = a.bit_length() - 1
        m[d - 1, :deg] = vinit[d - 1][:deg]
        for j in range(deg + 1, 33):
            ac = a
            v = int(m[d - 1, j - deg - 1])
            for k in range(deg):
                v ^= ((ac & 1) * int(m[d - 1, j - deg + k - 1])) << (deg - k)
                ac >>= 1
            m[d - 1, j - 1] = v & 0xFFFFFFFF
    for j in range(1, 33):
        m[:, j - 1] = (m[:, j - 1] * (1 << (32 - j))) & 0xFFFFFFFF
    return m.astype(np.uint32)


def lms_scramble(m, ltm_bits, shift_bits):
    """Linear matrix scramble + digital shift (src/scrambled_sobol.jl:104-140).

    ltm_bits: (D, 32, 32) 0/1 array as drawn by rand(rng, [0,1], D, 32, 32) (only the strict
    lower triangle is used, the diagonal is 1); shift_bits: (D, 32) as drawn by
    rand(rng, [0,1], D, 32).  Returns (m_scrambled, x0)."""
    D = m.shape[0]
    x0 = np.zeros(D, dtype=np.uint32)
    for d in range(D):
        x0[d] = sum(int(shift_bits[d, j]) << j for j in range(32)) & 0xFFFFFFFF
    out = np.zeros_like(m)
    for d in range(D):
        rows = []
        for p in range(32):
            row = 0
            for q in range(32):
                bit = 1 if q == p else (int(ltm_bits[d, p, q]) if q < p else 0)
                row |= bit << (31 - q)
            rows.append(row)
        for j in range(32):
            mdj = int(m[d, j])
            t2 = 0
            for p in range(32):  # p = 32:-1:1 in the reference, result bit 32-p
                t1 = bin(rows[p] & mdj).count("1") & 1
                t2 |= t1 << (31 - p)
            out[d, j] = t2
    return out, x0


def mock_rng_bits(*dims):
    """The reference test's MockRNG draw (test/scrambled_sobol.jl:97-105): element with
    0-based column-major linear index i is popcount(i) % 2."""
    n = int(np.prod(dims))
    flat = np.array([bin(i).count("1") % 2 for i in range(n)], dtype=np.uint8)
    return flat.reshape(dims, order="F")


class ScrambledSobolSeq:
    """State of the reference's ScrambledSobolSeq after construction.

    scramble_rng: None (unscrambled, the reference default), "mock" (the reference test's
    MockRNG) or a numpy Generator (documented project RNG: bits drawn as integers(0, 2))."""

    def __init__(self, D, scramble_rng=None):
        self.D = D
        self.m = direction_numbers(D)
        self.x0 = np.zeros(D, dtype=np.uint32)
        if scramble_rng is not None:
            if isinstance(scramble_rng, str) and scramble_rng == "mock":
                shift = mock_rng_bits(D, 32)
                ltm = mock_rng_bits(D, 32, 32)
            else:
                shift = scramble_rng.integers(0, 2, size=(D, 32), dtype=np.uint8)
                ltm = scramble_rng.integers(0, 2, size=(D, 32, 32), dtype=np.uint8)
            self.m, self.x0 = lms_scramble(self.m, ltm, shift)

    def points(self, n, first=0):
        """Host reference of next!: points first..first+n-1 via the Gray-code form
        x_i = x_0 xor (xor of m[:, j] over the set bits j of gray(i))."""
        out = np.empty((n, self.D))
        for r, i in enumerate(range(first, first + n)):
            g = i ^ (i >> 1)
            x = self.x0.astype(np.uint64).copy()
            j = 0
            while g:
                if g & 1:
                    x ^= self.m[:, j].astype(np.uint64)
                g >>= 1
                j += 1
            out[r] = np.ldexp(x.astype(np.float64), -32)
        return out
