"""qiw_inchworm_step_noalloc through the C ABI against qiw_inchworm_step: the same step with
nothing allocated, its per-call buffers in the device's static scratch array.  Two no-alloc
calls in a row, the second with a smaller input (fewer samples), each equal to the
allocating entry point's result; then the refusal when no plan for the run is resident."""
import numpy as np
import pytest

import qiw
from qiw import abi, model

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm((np.asarray(a) - np.asarray(b)).ravel()) / np.linalg.norm(np.asarray(b).ravel())


def noalloc_step(ctx, exp, n_w, orders, N, precision):
    run = abi.RunStruct(list(orders), [], N, 0, precision)
    val = np.zeros(exp.bt_size, np.complex128)
    per = np.zeros((len(list(orders)), exp.bt_size), np.complex128)
    rc = ctx.lib.qiw_inchworm_step_noalloc(ctx.h, run.ref(), int(n_w), val.ctypes.data_as(abi._f64p),
                                          per.ctypes.data_as(abi._f64p))
    return rc, val, per


@pytest.mark.parametrize("name", ["anderson", "two_orbital"])
def test_noalloc_step_matches_the_allocating_step(name):
    ctx = qiw.context()
    if name == "anderson":  # flat scalar walk (1x1 sectors), float64
        exp, orders, n_w, prec = model.anderson_single_orbital(n_tau=40), range(0, 5), 17, abi.QIW_PREC_FP64
        big, small = 2 ** 10, 2 ** 8
    else:  # vector walk (blocks 1, 4, 6, 4, 1), float32 chain products
        exp, orders, n_w, prec = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 24, 0.5), range(0, 4), 11, abi.QIW_PREC_FP32
        big, small = 2 ** 6, 2 ** 4
    print(f"noalloc: {name}, n_w={n_w}, N={big} then {small}", flush=True)
    # the allocating calls build and keep the plans of both sample counts
    ref = {N: qiw.inchworm_step(exp, n_w, orders, N, precision=prec, ctx=ctx) for N in (big, small)}
    # two no-alloc calls in a row, the second with fewer samples
    for N in (big, small):
        rc, val, per = noalloc_step(ctx, exp, n_w, orders, N, prec)
        assert rc == abi.QIW_OK, ctx.lib.qiw_last_error(ctx.h).decode()
        v_ref, per_ref = ref[N]
        assert rel(val, v_ref) < 1e-13          # same kernels, same fixed summation order
        assert rel(per, per_ref) < 1e-13


def test_noalloc_step_needs_a_resident_plan():
    ctx = qiw.context()
    exp = model.anderson_single_orbital(n_tau=40)
    qiw.inchworm_step(exp, 17, range(0, 3), 2 ** 8, ctx=ctx)
    rc, _, _ = noalloc_step(ctx, exp, 17, range(0, 3), 2 ** 12, abi.QIW_PREC_FP64)  # never run: no plan
    assert rc == abi.QIW_E_STATE
    assert "resident" in ctx.lib.qiw_last_error(ctx.h).decode()
