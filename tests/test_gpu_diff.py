"""GPU parity of diff_inchworm! (src/inchworm.jl:528-778) against the float64 oracle and
the reference's golden vector, through the C ABI.  Tolerance: 1e-10 relative, as for
inchworm!; self-energy deviations within 1e-10 of the samples' scale."""
import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu
RTOL = 1e-10


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.fixture(scope="module", autouse=True)
def threads():
    O.set_threads(16)


def golden():
    return model.single_state_golden_model(("occupied", "empty"))


def check(eg, ec, orders, N, rand=None, rp=None):
    So_g, Ss_g = qiw.diff_inchworm(eg, eg.grid, orders, N, rand_params=rp)
    P_c, So_c, Ss_c = O.diff_inchworm(ec, orders, N, rand=rand)
    assert rel(eg.P, P_c) < RTOL
    for o in So_c:
        assert np.abs(So_g[o] - So_c[o]).max() <= RTOL * max(np.abs(So_c[o]).max(), 1e-300)
        np.testing.assert_array_equal(np.isnan(Ss_g[o]), np.isnan(Ss_c[o]))
        fin = ~np.isnan(Ss_c[o])
        scale = np.abs(So_c[o]).max() + np.abs(Ss_c[o][fin]).max(initial=0.0)
        assert np.abs(Ss_g[o][fin] - Ss_c[o][fin]).max(initial=0.0) <= RTOL * max(scale, 1e-300)
    return So_g, Ss_g


def test_diff_golden():
    print("golden model diff_inchworm orders 0:3 N=2^8", flush=True)
    eg, ec = golden(), golden()
    check(eg, ec, range(0, 4), 2 ** 8)
    from test_oracle_diff import fx
    for s, name in enumerate(["/diff_inchworm/P/1", "/diff_inchworm/P/2"]):
        g = fx(name)
        assert np.linalg.norm(eg.P[:, s] - g) <= np.sqrt(np.finfo(float).eps) * np.linalg.norm(g)


def test_diff_grid_propagator_flat_walk():
    """Single-orbital Anderson model, grid-stored P (the self-energy row is tau_f_prev's)."""
    print("anderson n_tau=16 diff_inchworm orders 0:3 N=2^8", flush=True)
    eg, ec = model.anderson_single_orbital(n_tau=16), model.anderson_single_orbital(n_tau=16)
    check(eg, ec, range(0, 4), 2 ** 8)


def test_diff_block_model():
    print("kanamori2 n_tau=6 diff_inchworm orders 0:2 N=2^5", flush=True)
    mk = lambda: model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 6, 0.5)  # noqa: E731
    check(mk(), mk(), range(0, 3), 2 ** 5)


def test_diff_randomized():
    print("golden model diff_inchworm N_seqs=3 scrambled", flush=True)
    eg, ec = golden(), golden()
    So, Ss = check(eg, ec, range(0, 4), 2 ** 7, rand=(3, 1, 777, 0.0), rp=qiw.RandomizationParams(777, 3, 0.0))
    # order 1 is exact; in this one-level model the order-2 self-energy vanishes
    assert np.all(Ss[1] == 0) and np.all(Ss[2] == 0) and np.all(Ss[3][3:].real > 0)


def test_diff_needs_a_self_energy_order():
    ctx = qiw.context()
    e = golden()
    with pytest.raises(abi.QiwError) as err:
        qiw.diff_inchworm(e, e.grid, [0], 2 ** 4)
    assert err.value.code == abi.QIW_E_ARG
    del ctx
