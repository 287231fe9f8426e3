"""Oracle simplex maps (src/qmc_integrate.jl:198-258, :363-463) against closed-form
integrals over time-ordered simplices (restating test/qmc_integrate.jl:352-512 on the
imaginary branch)."""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as O
from qiw import sobol


def simplex(d_before, d_after, ti, tw, tf, n):
    d = d_before + d_after
    x = np.ascontiguousarray(sobol.ScrambledSobolSeq(d).points(n))
    out = np.zeros((n, d))
    jac = C.c_double()
    O.lib().orc_simplex_points(d_before, d_after, ti, tw, tf, O._p(x, C.c_double), n, O._p(out, C.c_double),
                               C.byref(jac))
    return out, jac.value


@pytest.mark.parametrize("d", [1, 2, 3, 4])
def test_root_transform_ordering_and_moments(d):
    ti, tf = 0.7, 3.2
    u, J = simplex(d, 0, ti, 0, tf, 2 ** 14)
    assert J == pytest.approx((tf - ti) ** d / math.factorial(d), rel=1e-14)
    assert np.all(u[:, :-1] >= u[:, 1:]) and np.all(u >= ti) and np.all(u <= tf)
    # int over ordered simplex of prod t_j = (1/d!) ((tf^2 - ti^2)/2)^d
    val = J * np.mean(np.prod(u, axis=1))
    exact = ((tf ** 2 - ti ** 2) / 2) ** d / math.factorial(d)
    assert val == pytest.approx(exact, rel=2e-4)


@pytest.mark.parametrize("db,da", [(0, 3), (3, 2), (1, 1), (2, 5), (5, 3)])
def test_double_simplex_closed_forms(db, da):
    ti, tw, tf = 1.1, 5.0, 5.5   # the bounds of test/qmc_integrate.jl:423-510
    u, J = simplex(db, da, ti, tw, tf, 2 ** 16)
    vol = (tw - ti) ** db / math.factorial(db) * (tf - tw) ** da / math.factorial(da)
    assert J == pytest.approx(vol, rel=1e-14)
    assert np.all(u[:, :da] >= tw) and np.all(u[:, :da] <= tf)
    assert np.all(u[:, da:] <= tw) and np.all(u[:, da:] >= ti)
    assert np.all(u[:, :-1] >= u[:, 1:])
    val = J * np.mean(np.prod(u, axis=1))
    exact = (((tf ** 2 - tw ** 2) / 2) ** da / math.factorial(da) *
             ((tw ** 2 - ti ** 2) / 2) ** db / math.factorial(db))
    assert val == pytest.approx(exact, rel=5e-4)


def test_reference_constant_and_multilinear_values():
    # test/qmc_integrate.jl:440, :461, :485, :509 on the imaginary branch: the contour
    # integral multiplies by (-i)^d and integrates prod(t.val) = prod(-i tau).
    u, J = simplex(0, 3, 1.1, 5.0, 5.5, 2 ** 19)
    assert ((-1j) ** 3 * J) == pytest.approx(0.0208333j, rel=1e-5)
    u, J = simplex(3, 2, 1.1, 5.0, 5.5, 2 ** 18)
    assert ((-1j) ** 5 * J) == pytest.approx(-1.23581j, rel=1e-5)
