"""Oracle Sobol generator (src/scrambled_sobol.jl) against SciPy and the reference tests
(test/scrambled_sobol.jl)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
from qiw import sobol

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "sobol_scipy.json")))


def orc_direction_numbers(D, ltm=None, shift=None):
    poly, vinit = sobol._table()
    a = np.array(poly[1:D], dtype=np.uint32)
    w = max(len(v) for v in vinit)
    minit = np.zeros((max(D - 1, 1), w), dtype=np.uint32)
    for d in range(1, D):
        minit[d - 1, :len(vinit[d])] = vinit[d]
    m = np.zeros((D, 32), np.uint32)
    x = np.zeros(D, np.uint32)
    u8 = C.POINTER(C.c_uint8)
    O.lib().orc_sobol_direction_numbers(
        D, O._p(a if D > 1 else np.zeros(1, np.uint32), C.c_uint32), O._p(minit, C.c_uint32), w,
        None if ltm is None else ltm.ctypes.data_as(u8), None if shift is None else shift.ctypes.data_as(u8),
        O._p(m, C.c_uint32), O._p(x, C.c_uint32))
    return m, x


def orc_points(m, x, n, skip=0, exact=True):
    D = m.shape[0]
    out = np.zeros((n, D))
    O.lib().orc_sobol_points(D, O._p(np.ascontiguousarray(m), C.c_uint32), O._p(x, C.c_uint32), skip,
                             int(exact), n, O._p(out, C.c_double))
    return out


@pytest.mark.parametrize("D,mexp", [(1, 3), (5, 3), (8, 6)])
def test_unscrambled_matches_scipy_bit_exact(D, mexp):
    m, x = orc_direction_numbers(D)
    assert np.array_equal(m, sobol.direction_numbers(D))  # host mirror agrees bit for bit
    pts = orc_points(m, x, 2 ** mexp)
    assert np.array_equal(pts, np.array(GOLD[f"unscrambled_d{D}"]))


def test_reference_listed_vectors_d5():
    # test/scrambled_sobol.jl:73-80 (SciPy values listed in the reference test)
    ref = [[0, 0, 0, 0, 0], [.5, .5, .5, .5, .5], [.75, .25, .25, .25, .75], [.25, .75, .75, .75, .25],
           [.375, .375, .625, .875, .375], [.875, .875, .125, .375, .875], [.625, .125, .875, .625, .625],
           [.125, .625, .375, .125, .125]]
    m, x = orc_direction_numbers(5)
    assert np.array_equal(orc_points(m, x, 8), np.array(ref))


def test_skip_semantics():
    # skip!(s, 3, exact=true) skips 3; skip!(s, 3) skips 4 = 2^floor(log2(3+1))
    m, x = orc_direction_numbers(5)
    full = orc_points(m, x, 8)
    assert np.array_equal(orc_points(m, x, 5, skip=3, exact=True), full[3:])
    assert np.array_equal(orc_points(m, x, 4, skip=3, exact=False), full[4:])


@pytest.mark.parametrize("D", [1, 5])
def test_mock_scrambled_matches_scipy(D):
    shift = np.ascontiguousarray(sobol.mock_rng_bits(D, 32))
    ltm = np.ascontiguousarray(sobol.mock_rng_bits(D, 32, 32))
    m, x = orc_direction_numbers(D, ltm, shift)
    pts = orc_points(m, x, 8)
    np.testing.assert_allclose(pts, np.array(GOLD[f"mock_scrambled_d{D}"]), atol=1e-10, rtol=0)
    # the host-side scrambler produces the same state the oracle does
    seq = sobol.ScrambledSobolSeq(D, scramble_rng="mock")
    assert np.array_equal(seq.m, m) and np.array_equal(seq.x0, x)
    np.testing.assert_array_equal(seq.points(8), pts)


def test_gray_code_jump_equals_sequential():
    # x_i = x_0 xor (xor of m_j over bits j of gray(i)) -- the form the GPU uses
    seq = sobol.ScrambledSobolSeq(6)
    seq_pts = orc_points(seq.m, seq.x0, 300)
    np.testing.assert_array_equal(seq.points(100, first=200), seq_pts[200:300])
