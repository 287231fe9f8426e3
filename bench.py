"""Benchmark of the qMC diagram-evaluation path on one MI355X.

Headline workload: the two-orbital Hubbard-Kanamori impurity (U = 2, J = 0.2, eps = -1
relative to half filling; 16 states in blocks 1, 4, 6, 4, 1 by particle number; 4
semi-elliptic baths, D = 2, V = 0.5, 8 pairs; beta = 10, n_tau = 200) solved by inchworm!
at orders 0:4 with N = 2^16 unscrambled Sobol points per (order, n_pts_after) group, the
chain products in float32 with float64 (compensated) sums and bookkeeping.  One bench step
is one bold step (all 16 sampled groups and the exact order-0 group in one evaluator
launch, then the on-device step bookkeeping); K steps are timed at tau_w = grid[100],
grid[101], ..., and the time to solution of the whole inchworm! is extrapolated as the
initial step plus 198 bold steps (every bold step does the same work).

Prints one JSON line: diagram evaluations per second (one evaluation = one (topology, qMC
sample) integrand evaluation summed over initial sectors and pair assignments), with the
roofline of the evaluator kernel, the float64 oracle timed on a bounded sample of the same
step on the host cores, the float32 error against float64, and secondary lines for the
single-orbital solve (C1), the three-orbital step (C3), the randomized-qMC study and the
differential formulation.
"""
import argparse
import contextlib
import gc
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "qinchworm.jl_amd"))
sys.path.insert(0, os.path.join(REPO, "tests"))

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 (vector and matrix)
FP32_PEAK_TFLOPS = 157.3  # MI355X dense FP32, published (vector with packed FP32, and matrix)
FP32_VALU_UNPACKED_TFLOPS = 78.6  # FP32 VALU without packed math (disabled in this build)
C1_TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "k_eval_flat_c1_hbm_traffic.md")
C2_TRAFFIC_PROFILE = os.path.join(REPO, "profiles", "k_eval_vec_c2_hbm_traffic.md")


_T0 = time.perf_counter()


def progress(msg):
    """One progress line on stderr (the JSON result is the only stdout line)."""
    print(f"[bench {time.perf_counter() - _T0:7.1f} s] {msg}", file=sys.stderr, flush=True)


def measured_traffic(path):
    """HBM bytes per evaluator launch from a committed rocprofv3 FETCH_SIZE/WRITE_SIZE
    summary (copy-calibrated), or None."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":")[1])
    except OSError:
        pass
    return None


def host_cpu():
    """Host CPU model, logical CPUs and the OpenMP thread setting of this process."""
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(), "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def cpu_threads():
    """All the host cores this process is meant to use: OMP_NUM_THREADS when set (the GPU
    host's CPU share), else every logical CPU."""
    env = os.environ.get("OMP_NUM_THREADS")
    try:
        n = int(env) if env else 0
    except ValueError:
        n = 0
    return max(1, min(n, os.cpu_count() or 1) if n > 0 else (os.cpu_count() or 1))


def c2_expansion():
    from qiw import model
    return model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 200, 0.5)


def two_orbital(log2n=14, cpu_n=16, with_cpu=True):
    """One bold step of the two-orbital Kanamori model (blocks 1,4,6,4,1; 8 pairs; beta=10,
    n_tau=200; tau_w = grid[100]; orders 0:4) on the GPU in float64 and with float32 chain
    products, and the float64 oracle on a bounded sample of the same step."""
    import numpy as np
    import qiw
    from qiw import abi
    k = c2_expansion()
    N = 2 ** log2n
    out = {"workload": f"two-orbital Kanamori (C2 model), one bold step tau_w=grid[100], orders 0:4, N=2^{log2n}"}
    vals = {}
    for name, prec in (("fp64", abi.QIW_PREC_FP64), ("fp32", abi.QIW_PREC_FP32)):
        with no_gc():
            t0 = time.perf_counter()
            v, _ = qiw.inchworm_step(k, 100, range(0, 5), N, precision=prec)
            dt = time.perf_counter() - t0
        st = qiw.stats()
        vals[name] = v
        out[f"gpu_{name}_s"] = dt
        out[f"gpu_{name}_evals_per_s"] = st["diagram_evals"] / dt
    out["fp32_rel_err_vs_fp64"] = float(np.linalg.norm(vals["fp32"] - vals["fp64"]) / np.linalg.norm(vals["fp64"]))
    if with_cpu:
        import oracle_lib as O
        threads = cpu_threads()
        O.set_threads(threads)
        O.reset_counters()
        kc = c2_expansion()
        t0 = time.perf_counter()
        O.inchworm_step(kc, range(0, 5), cpu_n, 100)
        dt = time.perf_counter() - t0
        _, evals, _ = O.counters()
        out["cpu_evals_per_s"] = evals / dt
        out["cpu_cores"] = threads
        out["cpu_sample"] = f"same step at N={cpu_n}, {evals} diagram evals in {dt:.1f} s, float64 oracle"
        out["speedup_fp64"] = out["gpu_fp64_evals_per_s"] / out["cpu_evals_per_s"]
        out["speedup_fp32"] = out["gpu_fp32_evals_per_s"] / out["cpu_evals_per_s"]
    return out



@contextlib.contextmanager
def no_gc():
    """A timed region without Python's cyclic collector: collect first, then keep it off, as
    timeit does (a full collection of the interpreter's objects takes tens of milliseconds
    and would otherwise land inside whichever step happens to cross its threshold)."""
    gc.collect()
    was = gc.isenabled()
    gc.disable()
    try:
        yield
    finally:
        if was:
            gc.enable()


def c1_expansion():
    from qiw import model
    return model.anderson_single_orbital(beta=10.0, n_tau=200, eps=0.1, U=1.0, V=0.5, D=2.0,
                                         with_correlators=True)


def one_step(exp, N):
    import qiw
    qiw.inchworm(exp, exp.grid, range(0, 5), range(0, 5), N)
    s1 = qiw.stats()
    qiw.correlator_2p(exp, exp.grid, range(0, 4), N)
    s2 = qiw.stats()
    return s1, s2


def cpu_baseline(N, steps=2):
    """The float64 oracle on a bounded sample of the same workload: the bare step and the
    first `steps` bold steps of the C1 inchworm! (same orders and N), on the host cores."""
    import oracle_lib as O
    threads = cpu_threads()
    O.set_threads(threads)
    O.reset_counters()
    exp = c1_expansion()
    t0 = time.perf_counter()
    O.inchworm(exp, range(0, 5), range(0, 5), N, n_steps_max=steps)
    dt = time.perf_counter() - t0
    _, evals, _ = O.counters()
    return {"value": evals / dt, "unit": "diagram evals/s", "cores": threads, "kind": "port",
            "sample": f"C1 inchworm! bare step + {steps} bold steps (orders 0:4, N=2^14), "
                      f"{evals} diagram evals in {dt:.2f} s, OMP threads={threads}"}


def c3_expansion():
    from qiw import model
    return model.kanamori(3, 3.0, 0.3, 0.0, 20.0, 400, 0.4)


def three_orbital(log2n=12, cpu_n=16, with_cpu=True):
    """One bold step of the three-orbital Kanamori model (blocks 1,6,15,20,15,6,1; 12 pairs;
    beta=20, n_tau=400; tau_w = grid[200]; orders 0:3) with float32 chain products on the
    matrix cores, and the float64 oracle on a bounded sample of the same step."""
    import qiw
    from qiw import abi
    k = c3_expansion()
    N = 2 ** log2n
    out = {"workload": f"three-orbital Kanamori (C3 model), one bold step tau_w=grid[200], orders 0:3, "
                       f"N=2^{log2n}, float32 chain products on MFMA"}
    qiw.inchworm_step(k, 200, range(0, 4), 2 ** 6, precision=abi.QIW_PREC_FP32)  # stages the model
    with no_gc():
        t0 = time.perf_counter()
        qiw.inchworm_step(k, 200, range(0, 4), N, precision=abi.QIW_PREC_FP32)
        dt = time.perf_counter() - t0
    st = qiw.stats()
    out["gpu_fp32_s"] = dt
    out["kernel_ms"] = st["kernel_ms"]
    out["gpu_fp32_evals_per_s"] = st["diagram_evals"] / dt
    out["gpu_leaves_per_s"] = st["leaves"] / dt
    if with_cpu:
        import oracle_lib as O
        threads = cpu_threads()
        O.set_threads(threads)
        O.reset_counters()
        kc = c3_expansion()
        t0 = time.perf_counter()
        O.inchworm_step(kc, range(0, 4), cpu_n, 200)
        dt = time.perf_counter() - t0
        _, evals, _ = O.counters()
        out["cpu_evals_per_s"] = evals / dt
        out["cpu_cores"] = threads
        out["cpu_sample"] = f"same step at N={cpu_n}, {evals} diagram evals in {dt:.1f} s, float64 oracle"
        out["speedup_fp32"] = out["gpu_fp32_evals_per_s"] / out["cpu_evals_per_s"]
    return out


def warp_study(n_tau=50, orders=(0, 1, 2), ref_log2n=20, log2ns=tuple(range(10, 17)), steps=(10, 25, 45),
               iters=4000, n_train=2 ** 14):
    """Error versus samples of the learned warp against plain qMC on the two-orbital model
    (C2 physics at n_tau=50, orders 0:2).  The propagator comes from a plain float64 run;
    the flow is trained on |integrand| samples of that run; for three bold steps the reference is
    the float64 step at 2^ref_log2n plain points (the GPU float64 path, equal to the
    oracle to 1e-10 in the parity tests)."""
    import numpy as np
    import qiw
    from qiw import model
    k = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, n_tau, 0.5)
    orders = list(orders)
    qiw.inchworm(k, k.grid, orders, orders, 2 ** 12)
    P = k.P.copy()
    qiw.warp_init(iters=iters, batch=8192, n_train_steps=8, n_train=n_train, lr=2e-3, seed=1)
    t0 = time.perf_counter()
    hist = qiw.warp_train(k, orders)
    train_s = time.perf_counter() - t0
    err = {"plain": np.zeros((len(steps), len(log2ns))), "warped": np.zeros((len(steps), len(log2ns)))}
    for si, jw in enumerate(steps):
        k.P = P.copy()
        qiw.warp_enable(False)
        ref, _ = qiw.inchworm_step(k, jw, orders, 2 ** ref_log2n)
        for ni, ln in enumerate(log2ns):
            for name, on in (("plain", False), ("warped", True)):
                qiw.warp_enable(on)
                v, _ = qiw.inchworm_step(k, jw, orders, 2 ** ln)
                err[name][si, ni] = np.linalg.norm(v - ref) / np.linalg.norm(ref)
    qiw.warp_enable(False)
    e_plain, e_warp = err["plain"].mean(axis=0), err["warped"].mean(axis=0)
    # plain samples needed to reach the warped error, by log-log interpolation
    lp = np.log2(e_plain)
    gains = []
    for ni, ln in enumerate(log2ns):
        target = np.log2(e_warp[ni])
        if target >= lp[0]:
            gains.append(None)
            continue
        if target <= lp[-1]:
            gains.append(float(2 ** (log2ns[-1] + (lp[-1] - target) / max(lp[-2] - lp[-1], 1e-9))
                               / 2 ** ln))
            continue
        j = int(np.nonzero(lp <= target)[0][0])
        x = log2ns[j - 1] + (lp[j - 1] - target) / (lp[j - 1] - lp[j])
        gains.append(float(2 ** x / 2 ** ln))
    return {"workload": f"two-orbital Kanamori, n_tau={n_tau}, orders {orders[0]}:{orders[-1]}, bold steps "
                        f"{list(steps)}; reference: plain float64 at 2^{ref_log2n}",
            "N": [2 ** ln for ln in log2ns], "rel_err_plain": e_plain.tolist(), "rel_err_warped": e_warp.tolist(),
            "sample_gain": gains, "train_iters": iters, "train_samples_per_step_group": n_train,
            "train_s": train_s,
            "loss_first": float(np.mean(hist[:20])), "loss_last": float(np.mean(hist[-20:]))}


def reference_step(k, jw, orders, log2n, block_log2n, precision):
    """The plain qMC step at N = 2^log2n as the mean of its 2^(log2n - block_log2n) blocks of
    2^block_log2n consecutive Sobol points, one launch each.  Block b of a Sobol sequence is
    its first 2^m points (m = block_log2n) digitally shifted: the point index b 2^m + r has
    the Gray code gray(r) xor (gray(b) << m) xor ((b & 1) << (m - 1)), so the block's shift
    is the xor of the direction numbers of the last two terms."""
    import numpy as np
    import qiw
    from qiw.sobol import ScrambledSobolSeq
    D = 2 * max(orders)
    base = ScrambledSobolSeq(D)
    m = block_log2n

    class Shifted:
        pass

    acc = None
    nb = 2 ** (log2n - m)
    for b in range(nb):
        bits = ((b ^ (b >> 1)) << m) ^ ((b & 1) << (m - 1))
        shift = base.x0.copy()
        for j in range(32):
            if (bits >> j) & 1:
                shift ^= base.m[:, j]
        seq = Shifted()
        seq.D, seq.m, seq.x0 = D, base.m, shift.astype(np.uint32)
        v, _ = qiw.inchworm_step(k, jw, orders, 2 ** m, precision=precision, seq=seq)
        acc = v if acc is None else acc + v
        progress(f"reference block {b + 1} of {nb} (2^{m} points each)")
    return acc / nb


def warp_time_to_error(log2ns=tuple(range(10, 17)), ref_log2n=20, jw=100, iters=4000, n_train=2 ** 14,
                       p_log2n=8):
    """The metric's time-to-error leg on the learned-warp workload: the two-orbital model of
    the headline (n_tau = 200, orders 0:4), one bold step at tau_w = grid[jw] in float32
    chain products, plain and warped qMC at N = 2^10 .. 2^16, each compared with the plain
    step at N = 2^ref_log2n (the reference estimate).  The propagator is that of a plain float32 solve at
    N = 2^p_log2n (the same P for every estimate); the flow (4 rational-quadratic layers,
    width 64) is trained on |integrand| samples of 8 bold steps, 2^14 per group.  The
    reference estimate runs the float32 chain products (1.8e-8 from float64 on one step, far
    below the errors it measures); a float64 one at 2^20 would take about 17 minutes.  Time per estimate is
    the evaluator time of the step (qiw_stats.kernel_ms); time to error is interpolated in
    log-log between the measured N."""
    import numpy as np
    import qiw
    from qiw import abi, model
    fp32 = abi.QIW_PREC_FP32
    k = model.kanamori(2, 2.0, 0.2, -1.0, 10.0, 200, 0.5)
    orders = list(range(0, 5))
    qiw.inchworm(k, k.grid, orders, orders, 2 ** p_log2n, precision=fp32)
    progress(f"warp leg: propagator (plain solve at 2^{p_log2n})")
    P = k.P.copy()
    qiw.warp_init(iters=iters, batch=8192, n_train_steps=8, n_train=n_train, lr=2e-3, seed=1)
    t0 = time.perf_counter()
    hist = qiw.warp_train(k, orders)
    train_s = time.perf_counter() - t0
    progress(f"warp leg: trained ({train_s:.1f} s)")
    k.P = P.copy()
    qiw.warp_enable(False)
    ref = reference_step(k, jw, orders, ref_log2n, 16, fp32)
    res = {"plain": [], "warped": []}
    for ln in log2ns:
        for name, on in (("plain", False), ("warped", True)):
            qiw.warp_enable(on)
            k.P = P.copy()
            v, _ = qiw.inchworm_step(k, jw, orders, 2 ** ln, precision=fp32)
            st = qiw.stats()
            res[name].append((2 ** ln, float(np.linalg.norm(v - ref) / np.linalg.norm(ref)), st["kernel_ms"] / 1e3))
        progress(f"warp leg: N=2^{ln}")
    qiw.warp_enable(False)

    def time_to(err, rows):  # log-log interpolation of time against error
        e = np.log([r[1] for r in rows])
        t = np.log([r[2] for r in rows])
        x = np.log(err)
        if x > e.max() or x < e.min():
            return None
        order = np.argsort(e)
        return float(np.exp(np.interp(x, e[order], t[order])))

    targets = [r[1] for r in res["warped"]]
    speedups = []
    for err in targets:
        tp, tw = time_to(err, res["plain"]), time_to(err, res["warped"])
        speedups.append(None if (tp is None or tw is None) else tp / tw)
    return {"workload": f"two-orbital Hubbard-Kanamori (the headline model), n_tau=200, orders 0:4, bold step "
                        f"tau_w=grid[{jw}], float32 chains; reference: plain at 2^{ref_log2n}",
            "N": [r[0] for r in res["plain"]], "rel_err_plain": [r[1] for r in res["plain"]],
            "rel_err_warped": [r[1] for r in res["warped"]], "time_s_plain": [r[2] for r in res["plain"]],
            "time_s_warped": [r[2] for r in res["warped"]],
            "time_to_error_speedup_at_warped_errors": speedups, "train_s": train_s,
            "train_iters": iters, "train_samples_per_step_group": n_train,
            "loss_first": float(np.mean(hist[:20])), "loss_last": float(np.mean(hist[-20:]))}


def rqmc_study(order=4, n_tau=200, log2n=14, n_seqs=8, seed=12345678, with_cpu=True, cpu_steps=2):
    """The reference's randomized-qMC study (bench/rqmc/rqmc.jl:37-117) on one GPU: the
    Bethe single-level model, inchworm! orders 0:order (bare and bold) and then
    correlator_2p(RequestStdDev) orders 0:order-1, N samples per sequence, N_seqs scrambled
    sequences (target_std 0: all of them).  Reports the wall time of both calls, diagram
    evaluations per second, and the largest deviation of G from the exact lattice value
    beside the largest estimated standard deviation of G."""
    import numpy as np
    import qiw
    from qiw import model
    N = 2 ** log2n

    def run():
        exp = model.bethe_single_level(n_tau)
        rp = qiw.RandomizationParams(seed, n_seqs, 0.0)
        t0 = time.perf_counter()
        _, Ps = qiw.inchworm(exp, exp.grid, range(0, order + 1), range(0, order + 1), N, rand_params=rp)
        s1 = qiw.stats()
        g, gs = qiw.correlator_2p(exp, exp.grid, range(0, order), N, qiw.RequestStdDev(), rand_params=rp)
        s2 = qiw.stats()
        return exp, Ps, g, gs, time.perf_counter() - t0, s1, s2

    run()  # plans and buffers
    with no_gc():
        exp, Ps, g, gs, dt, s1, s2 = run()
    evals = s1["diagram_evals"] + s2["diagram_evals"]
    cpu = {}
    if with_cpu:  # the oracle on the bare step and the first bold steps of the same inchworm!
        import oracle_lib as O
        threads = cpu_threads()
        O.set_threads(threads)
        O.reset_counters()
        t0 = time.perf_counter()
        O.inchworm(model.bethe_single_level(n_tau), range(0, order + 1), range(0, order + 1), N,
                   n_steps_max=cpu_steps, rand=(n_seqs, 1, seed, 0.0))
        dc = time.perf_counter() - t0
        _, ce, _ = O.counters()
        cpu = {"cpu_evals_per_s": ce / dc, "cpu_cores": threads,
               "cpu_sample": f"inchworm! bare step + {cpu_steps} bold steps, {ce} diagram evals in {dc:.2f} s, "
                             "float64 oracle, same sequences",
               "speedup": (evals / dt) / (ce / dc)}
    return {"workload": f"bench/rqmc study: Bethe single level, beta=10, n_tau={n_tau}, inchworm! orders 0:{order}, "
                        f"correlator_2p(RequestStdDev) orders 0:{order - 1}, N=2^{log2n} per sequence, "
                        f"N_seqs={n_seqs} scrambled, float64",
            "wall_s": dt, "diagram_evals_per_s": evals / dt,
            "evaluator_ms": s1["kernel_ms"] + s2["kernel_ms"],
            "max_P_std": float(max(np.abs(v).max() for v in Ps.values())),
            "max_G_std": float(np.abs(gs[0]).max()),
            "max_G_error_vs_exact": float(np.abs(g[0] - exp.corr_exact).max()), **cpu}


def diff_study(log2n=14, with_cpu=True, cpu_steps=4):
    """diff_inchworm! (the differential formulation) on the C1 model: orders 0:4, N=2^14,
    all n_tau - 1 = 199 steps on one GPU; the float64 oracle times the first `cpu_steps`
    steps of the same run on the host cores."""
    import qiw
    N = 2 ** log2n
    exp = c1_expansion()
    qiw.diff_inchworm(exp, exp.grid, range(0, 5), N)  # plans and buffers
    exp = c1_expansion()
    with no_gc():
        t0 = time.perf_counter()
        qiw.diff_inchworm(exp, exp.grid, range(0, 5), N)
        dt = time.perf_counter() - t0
    st = qiw.stats()
    out = {"workload": f"diff_inchworm! on the C1 model (single-orbital Anderson, beta=10, n_tau=200), orders 0:4, "
                       f"N=2^{log2n}, float64, 199 steps",
           "gpu_s": dt, "gpu_evals_per_s": st["diagram_evals"] / dt, "evaluator_ms": st["kernel_ms"]}
    if with_cpu:
        import oracle_lib as O
        threads = cpu_threads()
        O.set_threads(threads)
        O.reset_counters()
        t0 = time.perf_counter()
        O.diff_inchworm(c1_expansion(), range(0, 5), N, n_steps_max=cpu_steps)
        dc = time.perf_counter() - t0
        _, evals, _ = O.counters()
        out.update({"cpu_evals_per_s": evals / dc, "cpu_cores": threads,
                    "cpu_sample": f"first {cpu_steps} steps, {evals} diagram evals in {dc:.2f} s, float64 oracle"})
        out["speedup"] = out["gpu_evals_per_s"] / out["cpu_evals_per_s"]
    return out


def device_sync():
    """torch.cuda.synchronize on a GPU host, a no-op where torch sees no GPU."""
    try:
        import torch
        if torch.cuda.is_available():
            return torch.cuda.synchronize
    except Exception:
        pass
    return lambda: None


def copy_probe(mib):
    """A plain copy kernel of known size: 2 x mib MiB of HBM traffic per launch."""
    import torch
    a = torch.ones(mib * 2 ** 20 // 4, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": a.numel() * 4, "copy_probe_bytes_written": b.numel() * 4}))


def single_orbital(log2n, steps, warmup, sync, with_cpu):
    """The C1 workload (one full solve per step): inchworm! on the single-orbital Anderson
    model (n_tau = 200, orders 0:4/0:4, N = 2^log2n, float64) followed by correlator_2p for
    G, d and chi at orders 0:3.  Two timed passes over the same steps: the library default
    (the bold-step loop replayed from a hipGraph) gives the value; the eager pass, with HIP
    events around every evaluator launch on its stream, gives the launch average of the
    roofline."""
    import qiw
    N = 2 ** log2n
    exp = c1_expansion()

    def timed(graphs):
        qiw.set_graphs(graphs)
        for _ in range(max(warmup, 1 if graphs else 0)):  # a graph is captured on its first run
            one_step(exp, N)
        sync()
        acc = {"evals": 0.0, "flops": 0.0, "kernel_ms": 0.0, "launches": 0}
        with no_gc():
            t0 = time.perf_counter()
            for _ in range(steps):
                s1, s2 = one_step(exp, N)
                acc["evals"] += s1["diagram_evals"] + s2["diagram_evals"]
                acc["flops"] += s1["flops"] + s2["flops"]
                acc["kernel_ms"] += s1["kernel_ms"] + s2["kernel_ms"]
                acc["launches"] += s1["n_kernel_launches"] + s2["n_kernel_launches"]
            sync()
            acc["dt"] = time.perf_counter() - t0
        return acc

    g = timed(True)
    e = timed(False)
    qiw.set_graphs(True)
    avg_launch_ms = e["kernel_ms"] / max(e["launches"], 1)
    achieved = e["flops"] / e["launches"] / (avg_launch_ms * 1e-3) / 1e12 if e["launches"] else 0.0
    out = {"workload": "C1 single-orbital Anderson, beta=10, n_tau=200, inchworm! orders 0:4/0:4 + correlator_2p "
                       f"(G, d, chi) orders 0:3, N=2^{log2n}, float64; one step = the whole solve",
           "value": g["evals"] / g["dt"], "unit": "diagram evals/s", "ms_per_step": 1e3 * g["dt"] / steps,
           "time_to_solution_s": g["dt"] / steps,
           "eager": {"ms_per_step": 1e3 * e["dt"] / steps, "value": e["evals"] / e["dt"]},
           "roofline": {"bound": "salu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP64_PEAK_TFLOPS, "traffic": measured_traffic(C1_TRAFFIC_PROFILE),
                        "kernel": "k_eval_flat_t<4,4,2,false>", "avg_launch_ms": avg_launch_ms,
                        "note": "float64 scalar walk: one multiply per trie edge, bound by scalar-unit issue"}}
    if with_cpu:
        out["cpu_baseline"] = cpu_baseline(N)
    return out


def two_orbital_headline(log2n, steps, warmup, sync, with_cpu, cpu_log2n=8, err_log2n=12):
    """The headline: K timed bold steps of the two-orbital inchworm! (float32 chain
    products) at tau_w = grid[100 + i], the initial step timed once, the float32 error
    against float64 on one step at N = 2^err_log2n, and the float64 oracle on a bounded
    sample of the same step (N = 2^cpu_log2n points per group) on all host threads."""
    import numpy as np
    import qiw
    from qiw import abi
    N = 2 ** log2n
    fp32 = abi.QIW_PREC_FP32
    k = c2_expansion()
    orders = range(0, 5)
    for i in range(max(warmup, 1)):  # plans and buffers at this N
        qiw.inchworm_step(k, 100, orders, N, precision=fp32)
        progress(f"two-orbital warmup step {i + 1} (N=2^{log2n})")
    sync()
    acc = {"evals": 0.0, "flops": 0.0, "kernel_ms": 0.0, "launches": 0, "leaves": 0.0}
    with no_gc():
        t0 = time.perf_counter()
        for i in range(steps):
            qiw.inchworm_step(k, 100 + i, orders, N, precision=fp32)
            progress(f"two-orbital timed step {i + 1} of {steps}")
            st = qiw.stats()
            acc["evals"] += st["diagram_evals"]
            acc["flops"] += st["flops"]
            acc["kernel_ms"] += st["kernel_ms"]
            acc["launches"] += st["n_kernel_launches"]
            acc["leaves"] += st["leaves"]
        sync()
        dt = time.perf_counter() - t0
    qiw.inchworm_step_bare(k, orders, N)  # plans of the initial step
    with no_gc():
        t0 = time.perf_counter()
        qiw.inchworm_step_bare(k, orders, N)
        sync()
        t_bare = time.perf_counter() - t0
    progress("two-orbital initial step")
    t_step = dt / steps
    avg_launch_ms = acc["kernel_ms"] / max(acc["launches"], 1)
    achieved = acc["flops"] / acc["launches"] / (avg_launch_ms * 1e-3) / 1e12 if acc["launches"] else 0.0
    # float32 against float64 on one bold step
    ne = 2 ** err_log2n
    v32, _ = qiw.inchworm_step(k, 100, orders, ne, precision=fp32)
    v64, _ = qiw.inchworm_step(k, 100, orders, ne, precision=abi.QIW_PREC_FP64)
    progress("two-orbital float32 against float64")
    out = {"value": acc["evals"] / dt, "ms_per_step": 1e3 * t_step, "t_bare_s": t_bare,
           "time_to_solution_s": t_bare + 198 * t_step,
           "leaves_per_s": acc["leaves"] / dt,
           "fp32_vs_fp64": {"N": ne, "rel_err_step": float(np.linalg.norm(v32 - v64) / np.linalg.norm(v64)),
                            "max_abs_err_step": float(np.abs(v32 - v64).max())},
           "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                        "frac": achieved / FP32_PEAK_TFLOPS,
                        "frac_of_unpacked_fp32_valu": achieved / FP32_VALU_UNPACKED_TFLOPS,
                        "traffic": measured_traffic(C2_TRAFFIC_PROFILE), "kernel": "k_eval_vec_t<float,6,4,2,2>",
                        "avg_launch_ms": avg_launch_ms, "flops_per_launch": acc["flops"] / max(acc["launches"], 1),
                        "note": "float32 chain products on the vector ALUs (blocks <= 6 do not fill MFMA tiles); "
                                "packed FP32 is disabled in this build, so 78.6 TFLOP/s is the reachable ceiling; "
                                "the busiest unit is the vector-memory address path (per-lane propagator-cell "
                                "gathers, 79 % busy, profiles/k_eval_vec_c2_counters.md), VALU issue 39 %"}}
    if with_cpu:
        import oracle_lib as O
        threads = cpu_threads()
        O.set_threads(threads)
        O.reset_counters()
        kc = c2_expansion()
        ncpu = 2 ** cpu_log2n
        t0 = time.perf_counter()
        O.inchworm_step(kc, orders, ncpu, 100)
        dc = time.perf_counter() - t0
        progress("two-orbital CPU baseline")
        _, evals, _ = O.counters()
        out["cpu_baseline"] = {"value": evals / dc, "unit": "diagram evals/s", "cores": threads, "kind": "port",
                               "sample": f"the same bold step (tau_w = grid[100], orders 0:4) at N=2^{cpu_log2n} "
                                         f"points per group: {evals} diagram evals in {dc:.1f} s, float64 oracle, "
                                         f"OpenMP over samples", **host_cpu()}
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--log2n", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--headline-only", action="store_true", help="only the two-orbital headline (profiling runs)")
    ap.add_argument("--no-c3", action="store_true", help="skip the three-orbital step line")
    ap.add_argument("--no-rqmc", action="store_true", help="skip the randomized-qMC and differential lines")
    ap.add_argument("--block-models-only", action="store_true",
                    help="only the two- and three-orbital steps (vector and matrix-core walks), one JSON line")
    ap.add_argument("--c1-only", action="store_true",
                    help="only the single-orbital solve (flat scalar walk), no CPU leg (profiling runs)")
    ap.add_argument("--c3-only", type=int, default=0, metavar="LOG2N",
                    help="only the three-orbital bold step (matrix-core walk) at N=2^LOG2N, no CPU leg "
                         "(profiling runs)")
    ap.add_argument("--warp-study", action="store_true",
                    help="only run the learned-warp error-versus-samples study and print its JSON line")
    ap.add_argument("--warp-time-to-error", action="store_true",
                    help="only run the time-to-error leg of the learned warp on the headline model "
                         "(orders 0:4, n_tau=200; about 10 minutes) and print its JSON line")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="only run 3 device-to-device copies of this many MiB (calibrates the "
                         "FETCH_SIZE / WRITE_SIZE hardware counters) and exit")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); multi-GPU sharding is "
              "designed but not built", file=sys.stderr)
        sys.exit(2)
    if args.copy_probe_mib > 0:
        copy_probe(args.copy_probe_mib)
        return
    if args.warp_study:
        print(json.dumps(warp_study()), flush=True)
        return
    if args.warp_time_to_error:
        print(json.dumps(warp_time_to_error()), flush=True)
        return
    if args.c1_only:
        print(json.dumps({"single_orbital": single_orbital(14, args.steps, args.warmup, device_sync(), False)}),
              flush=True)
        return
    if args.c3_only > 0:
        print(json.dumps({"three_orbital_step": three_orbital(log2n=args.c3_only, with_cpu=False)}), flush=True)
        return
    if args.block_models_only:
        print(json.dumps({"two_orbital_step": two_orbital(with_cpu=not args.no_cpu),
                          "three_orbital_step": three_orbital(with_cpu=not args.no_cpu)}), flush=True)
        return
    sync = device_sync()
    h = two_orbital_headline(args.log2n, args.steps, args.warmup, sync, not args.no_cpu)
    N = 2 ** args.log2n
    line = {
        "metric": "diagram evals/s; inchworm! time-to-error vs FP64 oracle, warped vs plain qMC",
        "value": h["value"],
        "unit": "diagram evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (deterministic model, unscrambled Sobol points)",
        "config": {"workload": "two-orbital Hubbard-Kanamori (U=2, J=0.2, eps=-1; blocks 1,4,6,4,1; 8 pairs; "
                               "beta=10, n_tau=200), inchworm! orders 0:4, N=2^%d per group, float32 chain "
                               "products with float64 compensated sums; one step = one bold step "
                               "(tau_w = grid[100+i]); time_to_solution_s extrapolated as the initial step "
                               "+ 198 bold steps" % args.log2n,
                   "n_tau": 200, "orders": "0:4", "N_samples": N, "precision": "fp32 chains, fp64 accumulation"},
        "time_to_solution_s": h["time_to_solution_s"],
        "time_to_solution_note": "extrapolated: t_bare + 198 x (mean bold step)",
        "t_bare_s": h["t_bare_s"],
        "leaves_per_s": h["leaves_per_s"],
        "fp32_vs_fp64": h["fp32_vs_fp64"],
        "roofline": h["roofline"],
    }
    if "cpu_baseline" in h:
        line["cpu_baseline"] = h["cpu_baseline"]
        line["speedup_vs_cpu_baseline"] = h["value"] / h["cpu_baseline"]["value"]
    if not args.headline_only:
        line["single_orbital"] = single_orbital(14, 3, 1, sync, not args.no_cpu)
        progress("single-orbital solve")
        if not args.no_c3:
            line["three_orbital_step"] = three_orbital(with_cpu=not args.no_cpu)
            progress("three-orbital step")
        if not args.no_rqmc:
            line["rqmc_study"] = rqmc_study(with_cpu=not args.no_cpu)
            progress("randomized qMC study")
            line["diff_inchworm"] = diff_study(with_cpu=not args.no_cpu)
            progress("differential inchworm")
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
