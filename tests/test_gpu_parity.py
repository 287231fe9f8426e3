"""GPU parity: libqiw on gfx950 against the CPU float64 oracle, through the C ABI.

Tolerances: Sobol points and their Gray-code jump are bit-exact; float64 results agree
with the oracle to 1e-10 relative (the float64 target of the hot path)."""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as O
import qiw
from qiw import abi, model
from qiw.sobol import ScrambledSobolSeq

pytestmark = pytest.mark.gpu
FX = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "inchworm_h5.json")))
RTOL = 1e-10


def fx(name):
    return np.array([complex(a, b) for a, b in FX[name]["data"]])


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.linalg.norm((a - b).ravel()) / max(np.linalg.norm(b.ravel()), 1e-300)


@pytest.fixture(scope="module")
def ctx():
    return qiw.context()


def sobol_on_gpu(ctx, seq, first, n):
    D = seq.D
    ctx.check(ctx.lib.qiw_set_sobol(ctx.h, D, abi.ptr(np.ascontiguousarray(seq.m), C.c_uint32),
                                    abi.ptr(np.ascontiguousarray(seq.x0), C.c_uint32)))
    out = np.zeros((n, D))
    ctx.check(ctx.lib.qiw_sobol_points(ctx.h, D, first, n, out.ctypes.data_as(abi._f64p)))
    return out


@pytest.mark.parametrize("scr", [None, "mock"])
def test_sobol_points_bit_exact(ctx, scr):
    seq = ScrambledSobolSeq(8 if scr is None else 5, scramble_rng=scr)
    gpu = sobol_on_gpu(ctx, seq, 0, 4096)
    cpu = np.zeros_like(gpu)
    O.lib().orc_sobol_points(seq.D, O._p(seq.m, C.c_uint32), O._p(seq.x0, C.c_uint32), 0, 1, 4096,
                             O._p(cpu, C.c_double))
    assert np.array_equal(gpu, cpu)
    # random access (the jump a sharded run would use) equals the sequential stream
    assert np.array_equal(sobol_on_gpu(ctx, seq, 1000, 100), cpu[1000:1100])


@pytest.mark.parametrize("db,da", [(0, 4), (3, 2), (5, 3), (1, 1), (0, 8)])
def test_simplex_points(ctx, db, da):
    d = db + da
    seq = ScrambledSobolSeq(d)
    sobol_on_gpu(ctx, seq, 0, 1)
    n = 2048
    gpu = np.zeros((n, d))
    ti, tw, tf = 0.3, 4.0, 4.7
    ctx.check(ctx.lib.qiw_simplex_points(ctx.h, db, da, ti, tw, tf, n, gpu.ctypes.data_as(abi._f64p)))
    x = np.ascontiguousarray(seq.points(n))
    cpu = np.zeros((n, d))
    O.lib().orc_simplex_points(db if da else d, da, ti, tw, tf, O._p(x, C.c_double), n, O._p(cpu, C.c_double), None)
    np.testing.assert_allclose(gpu, cpu, rtol=1e-14, atol=1e-14)


def golden():
    return model.single_state_golden_model(("occupied", "empty"))


def test_bare_step_golden(ctx):
    exp = golden()
    v_gpu, per_gpu = qiw.inchworm_step_bare(exp, [0, 1, 2], 2 ** 8)
    v_cpu, per_cpu = O.inchworm_step_bare(exp, [0, 1, 2], 2 ** 8)
    assert rel(v_gpu, v_cpu) < 1e-13 and rel(per_gpu, per_cpu) < 1e-13
    P1, P2 = fx("/inchworm/P/1"), fx("/inchworm/P/2")
    assert abs((v_gpu[1] / v_gpu[0]).real - (P2[1] / P1[1]).real) < 1e-13


def test_bold_step_on_evolved_propagator(ctx):
    exp = model.anderson_single_orbital(n_tau=24)
    P, _, _ = O.inchworm(exp, range(0, 4), range(0, 4), 2 ** 8, n_steps_max=10)
    exp.P, exp.P_n_known = P, exp.grid.n_tau
    for n_w in (3, 11):
        v_gpu, per_gpu = qiw.inchworm_step(exp, n_w, range(0, 5), 2 ** 9)
        v_cpu, per_cpu = O.inchworm_step(exp, range(0, 5), 2 ** 9, n_w)
        assert rel(v_gpu, v_cpu) < RTOL
        for o in range(5):  # per-order parts to 1e-10 of the step value
            assert np.abs(per_gpu[o] - per_cpu[o]).max() < RTOL * np.abs(v_cpu).max()


def test_inchworm_golden_and_correlator(ctx):
    exp_g, exp_c = golden(), golden()
    Po_g, Ps_g = qiw.inchworm(exp_g, exp_g.grid, range(0, 4), range(0, 3), 2 ** 8)
    P_c, Po_c, Ps_c = O.inchworm(exp_c, range(0, 4), range(0, 3), 2 ** 8)
    assert rel(exp_g.P, P_c) < RTOL
    for o in Po_c:
        assert rel(Po_g[o], Po_c[o]) < 1e-9
        np.testing.assert_array_equal(np.isnan(Ps_g[o]), np.isnan(Ps_c[o]))
    # the reference's own check: isapprox per sector, rtol = sqrt(eps)
    for s, name in enumerate(["/inchworm/P/1", "/inchworm/P/2"]):
        assert rel(exp_g.P[:, s], fx(name)) < np.sqrt(np.finfo(float).eps)
    exp_g.add_corr_operators((exp_g.fock.c("0"), exp_g.fock.c_dag("0")))
    exp_c.add_corr_operators((exp_c.fock.c("0"), exp_c.fock.c_dag("0")))
    exp_c.P, exp_c.P_n_known = P_c, exp_c.grid.n_tau
    g_gpu = -qiw.correlator_2p(exp_g, exp_g.grid, range(0, 4), 2 ** 8)[0]
    g_cpu = -O.correlator_2p(exp_c, range(0, 4), 2 ** 8)[0]
    assert rel(g_gpu, g_cpu) < RTOL
    assert rel(g_gpu, fx("/inchworm/g")) < np.sqrt(np.finfo(float).eps)


def test_inchworm_single_orbital_c0(ctx):
    """Single-orbital Anderson model: beta=10, n_tau=64, orders 0:3, N=2^10, float64."""
    exp_g = model.anderson_single_orbital(n_tau=64)
    exp_c = model.anderson_single_orbital(n_tau=64)
    Po_g, _ = qiw.inchworm(exp_g, exp_g.grid, range(0, 4), range(0, 4), 2 ** 10)
    O.set_threads(16)
    P_c, Po_c, _ = O.inchworm(exp_c, range(0, 4), range(0, 4), 2 ** 10)
    assert rel(exp_g.P, P_c) < RTOL
    assert np.max(np.abs(exp_g.P[-1].imag)) == pytest.approx(1.0, abs=1e-15)  # normalize!(P, tau)
    Z = qiw.partition_function(exp_g)
    assert Z.real > 0 and abs(Z.imag) < 1e-14


def test_correlators_single_orbital(ctx):
    exp_g = model.anderson_single_orbital(n_tau=32, with_correlators=True)
    exp_c = model.anderson_single_orbital(n_tau=32, with_correlators=True)
    qiw.inchworm(exp_g, exp_g.grid, range(0, 4), range(0, 4), 2 ** 9)
    P_c, _, _ = O.inchworm(exp_c, range(0, 4), range(0, 4), 2 ** 9)
    exp_c.P, exp_c.P_n_known = P_c, exp_c.grid.n_tau
    gg = qiw.correlator_2p(exp_g, exp_g.grid, range(0, 4), 2 ** 9)
    gc = O.correlator_2p(exp_c, range(0, 4), 2 ** 9)
    for a, b in zip(gg, gc):
        assert rel(a, b) < RTOL


def test_error_codes(ctx):
    exp = golden()
    with pytest.raises(abi.QiwError) as e:
        qiw.inchworm(exp, exp.grid, [0, 1], [0, 1], 100)
    assert e.value.code == abi.QIW_E_NOT_POW2
    k = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 8, 0.5)
    with pytest.raises(abi.QiwError) as e:
        qiw.inchworm(k, k.grid, [0, 7], [0, 1], 2 ** 4)   # order beyond the compiled limit
    assert e.value.code in (abi.QIW_E_LIMIT, abi.QIW_E_DIM)


def anderson_tables(n_tables, n_tau=16):
    """Single-orbital Anderson model whose pairs use 1 or 3 distinct hybridization tables,
    so the W slots of the walk (arc * stride + table) differ from arc * n_tables."""
    fock = model.FockSpace(["up", "dn"])
    H = 0.1 * (fock.n("up") + fock.n("dn")) + 1.0 * fock.n("up") @ fock.n("dn")
    ed = model.EDCore(H, fock, lambda b: (fock.occupation(b, "up"), fock.occupation(b, "dn")))
    grid = model.ImaginaryTimeGrid(10.0, n_tau)
    delta = 0.25 * model.bethe_gf(1.0, 10.0, grid)
    if n_tables == 1:
        pairs = [model.InteractionPair(fock.c_dag(s), fock.c(s), delta) for s in ("up", "dn")]
    else:
        pairs = [model.InteractionPair(fock.c_dag("up"), fock.c("up"), delta),
                 model.InteractionPair(fock.c("up"), fock.c_dag("up"), model.ph_conj(delta)),
                 model.InteractionPair(fock.c_dag("dn"), fock.c("dn"), 0.5 * delta)]
    return model.Expansion(ed, grid, pairs)


@pytest.mark.parametrize("n_tables", [1, 3])
def test_table_count_other_than_stride(ctx, n_tables):
    """One and three distinct tables: W slots are laid out with the kernel's table stride."""
    exp_g, exp_c = anderson_tables(n_tables), anderson_tables(n_tables)
    P, _, _ = O.inchworm(exp_c, range(0, 4), range(0, 4), 2 ** 7, n_steps_max=6)
    for e in (exp_g, exp_c):
        e.P, e.P_n_known = P, e.grid.n_tau
    v_gpu, per_gpu = qiw.inchworm_step(exp_g, 5, range(0, 5), 2 ** 8)
    v_cpu, per_cpu = O.inchworm_step(exp_c, range(0, 5), 2 ** 8, 5)
    assert rel(v_gpu, v_cpu) < RTOL
    for o in range(1, 5):
        assert np.abs(per_cpu[o]).max() > 0
        assert np.abs(per_gpu[o] - per_cpu[o]).max() < RTOL * np.abs(v_cpu).max()


def test_plans_reused_and_rebuilt(ctx):
    """Compiled batches stay on the device between calls while the model structure, the
    Sobol tables and the run parameters are unchanged, and are rebuilt when any changes:
    interleaved calls on two models and two sample counts all match the oracle."""
    models = {nt: (anderson_tables(nt), anderson_tables(nt)) for nt in (1, 3)}
    for eg, ec in models.values():
        P, _, _ = O.inchworm(ec, range(0, 4), range(0, 4), 2 ** 6, n_steps_max=6)
        for e in (eg, ec):
            e.P, e.P_n_known = P, e.grid.n_tau
    ref = {(nt, N): O.inchworm_step(models[nt][1], range(0, 5), N, 5)[0] for nt in (1, 3) for N in (2 ** 7, 2 ** 8)}
    first = {}
    for nt, N in [(1, 2 ** 8), (3, 2 ** 8), (1, 2 ** 8), (1, 2 ** 7), (3, 2 ** 8), (1, 2 ** 8)]:
        v, _ = qiw.inchworm_step(models[nt][0], 5, range(0, 5), N)
        assert rel(v, ref[(nt, N)]) < RTOL
        if (nt, N) in first:  # a reused plan reproduces the value bit for bit
            np.testing.assert_array_equal(v, first[(nt, N)])
        first.setdefault((nt, N), v)


def test_correlator_repeated_call(ctx):
    exp = model.anderson_single_orbital(n_tau=16, with_correlators=True)
    qiw.inchworm(exp, exp.grid, range(0, 3), range(0, 3), 2 ** 7)
    g1 = qiw.correlator_2p(exp, exp.grid, range(0, 3), 2 ** 7)
    g2 = qiw.correlator_2p(exp, exp.grid, range(0, 3), 2 ** 7)
    for a, b in zip(g1, g2):
        np.testing.assert_array_equal(a, b)


def test_work_mappings_keep_parity(ctx):
    """Every work mapping (topology chunk size) gives the oracle's values, and the tuner
    keeps one of its timed candidates."""
    exp_g, exp_c = anderson_tables(3), anderson_tables(3)
    P, _, _ = O.inchworm(exp_c, range(0, 4), range(0, 4), 2 ** 6, n_steps_max=6)
    for e in (exp_g, exp_c):
        e.P, e.P_n_known = P, e.grid.n_tau
    ref, _ = O.inchworm_step(exp_c, range(0, 5), 2 ** 8, 5)
    try:
        for cw in (64, 256, 2048):
            qiw.set_tuning(chunk_words=cw)
            assert qiw.get_tuning()["chunk_words"] == cw
            v, _ = qiw.inchworm_step(exp_g, 5, range(0, 5), 2 ** 8)
            assert rel(v, ref) < RTOL
        t = qiw.autotune(exp_g, range(0, 5), 2 ** 8)
        assert t["chunk_words"] in (256, 384, 512, 640, 896, 1024, 2048) and t["n_tried"] == 7 and t["best_ms"] > 0
        v, _ = qiw.inchworm_step(exp_g, 5, range(0, 5), 2 ** 8)
        assert rel(v, ref) < RTOL
        with pytest.raises(qiw.abi.QiwError):
            qiw.set_tuning(samples_per_task=3)
    finally:
        qiw.set_tuning()


def spinless3(n_tau=8):
    """Three spinless orbitals with density-density interactions: 8 one-dimensional
    sectors (one per Fock state), pairs c^dag_i/c_i and their conjugates, 2 tables."""
    labels = ["a", "b", "c"]
    fock = model.FockSpace(labels)
    eps = {"a": -0.3, "b": 0.1, "c": 0.4}
    H = sum(eps[l] * fock.n(l) for l in labels)
    H = H + 0.8 * (fock.n("a") @ fock.n("b") + fock.n("b") @ fock.n("c") + fock.n("a") @ fock.n("c"))
    ed = model.EDCore(H, fock, lambda b: b)
    grid = model.ImaginaryTimeGrid(5.0, n_tau)
    delta = 0.25 * model.bethe_gf(1.0, 5.0, grid)
    pairs = []
    for l in labels:
        pairs.append(model.InteractionPair(fock.c_dag(l), fock.c(l), delta))
        pairs.append(model.InteractionPair(fock.c(l), fock.c_dag(l), model.ph_conj(delta)))
    return model.Expansion(ed, grid, pairs)


@pytest.mark.parametrize("K,N", [(4, 2 ** 6), (5, 2 ** 4)])
def test_eight_sectors_and_order_five(ctx, K, N):
    """Flat walk with 8 sectors (orders up to 4) and its order-6 instantiation (order 5)."""
    exp_g, exp_c = spinless3(), spinless3()
    P, _, _ = O.inchworm(exp_c, range(0, 3), range(0, 3), 2 ** 4, n_steps_max=3)
    for e in (exp_g, exp_c):
        e.P, e.P_n_known = P, e.grid.n_tau
    v_gpu, per_gpu = qiw.inchworm_step(exp_g, 3, range(0, K + 1), N)
    v_cpu, per_cpu = O.inchworm_step(exp_c, range(0, K + 1), N, 3)
    assert rel(v_gpu, v_cpu) < RTOL
    for o in range(K + 1):
        assert np.abs(per_gpu[o] - per_cpu[o]).max() < RTOL * np.abs(v_cpu).max()
    assert np.abs(per_cpu[K]).max() > 0


def test_step_graph_replays_match_eager():
    """The bold-step loop replayed from its hipGraph gives the eager run's results bit for
    bit, on every replay, with other GPU work (a correlator, Sobol points) and host
    synchronisation between the replays; a change of the run (another N) is captured anew."""
    print("step graph: anderson n_tau=24 orders 0:3 N=2^8, 3 replays", flush=True)
    ctx = qiw.context()
    try:
        qiw.set_graphs(False)
        e = model.anderson_single_orbital(n_tau=24, with_correlators=True)
        Po_e, _ = qiw.inchworm(e, e.grid, range(0, 4), range(0, 4), 2 ** 8)
        P_e = e.P.copy()
        s_e = qiw.stats()
        qiw.set_graphs(True)
        for rep in range(3):
            g = model.anderson_single_orbital(n_tau=24, with_correlators=True)
            Po_g, _ = qiw.inchworm(g, g.grid, range(0, 4), range(0, 4), 2 ** 8)
            np.testing.assert_array_equal(g.P, P_e)
            for o in Po_e:
                np.testing.assert_array_equal(Po_g[o], Po_e[o])
            s_g = qiw.stats()
            assert s_g["n_kernel_launches"] == s_e["n_kernel_launches"] and s_g["kernel_ms"] > 0
            qiw.correlator_2p(g, g.grid, range(0, 3), 2 ** 6)  # eager work between replays
            sobol_on_gpu(ctx, ScrambledSobolSeq(4), 0, 64)
        h = model.anderson_single_orbital(n_tau=24, with_correlators=True)
        qiw.inchworm(h, h.grid, range(0, 4), range(0, 4), 2 ** 7)
        hc = model.anderson_single_orbital(n_tau=24, with_correlators=True)
        qiw.set_graphs(False)
        qiw.inchworm(hc, hc.grid, range(0, 4), range(0, 4), 2 ** 7)
        np.testing.assert_array_equal(h.P, hc.P)
    finally:
        qiw.set_graphs(True)
