"""Where the host compiler cuts a vector program into topology chunks (qiw_chunk_starts in
qinchworm.jl_amd/csrc/qiw_chunks.h), compiled with g++ and checked on random topology
costs against properties and a brute-force optimum:

* the starts partition the topologies into non-empty contiguous chunks, in order;
* plain rule (one wavefront per task): every chunk but the last has reached chunk_cost, and
  closes at the topology that reached it;
* matrix-core rule (four chunks per workgroup): the chunk count is at most the plain count
  rounded up to a multiple of four (and at most the number of topologies), and the largest
  chunk costs exactly the least any partition into that many contiguous chunks can reach
  (dynamic programming over all partitions), so never more than under the plain rule."""
import os
import shutil
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "qinchworm.jl_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path_factory.mktemp("chunks") / "chunk_starts")
    subprocess.run([gxx, "-O1", "-std=c++17", "-I", CSRC, os.path.join(HERE, "native", "chunk_starts_main.cpp"),
                    "-o", exe], check=True)

    def run(cases):
        text = "".join(f"{cc} {mult} {' '.join(str(int(c)) for c in costs)}\n" for cc, mult, costs in cases)
        out = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout
        return [[int(v) for v in line.split()] for line in out.strip("\n").split("\n")]
    return run


def chunk_costs(costs, starts):
    bounds = list(starts) + [len(costs)]
    return [int(sum(costs[bounds[i]:bounds[i + 1]])) for i in range(len(starts))]


def min_max_partition(costs, k):
    """The least possible largest chunk over all partitions of costs, in order, into at
    most k non-empty contiguous chunks (brute-force dynamic programming)."""
    n = len(costs)
    pre = np.concatenate([[0], np.cumsum(costs)])
    inf = float("inf")
    best = [[inf] * (n + 1) for _ in range(k + 1)]
    best[0][0] = 0
    for parts in range(1, k + 1):
        for i in range(1, n + 1):
            for j in range(parts - 1, i):
                best[parts][i] = min(best[parts][i], max(best[parts - 1][j], pre[i] - pre[j]))
    return min(best[p][n] for p in range(1, k + 1))


def random_cases(rng, n_cases, multiple):
    cases = []
    for _ in range(n_cases):
        n = int(rng.integers(1, 13))
        kind = rng.integers(0, 3)
        if kind == 0:
            costs = rng.integers(1, 100, size=n)
        elif kind == 1:  # a few heavy topologies among light ones (the three-orbital groups)
            costs = np.where(rng.random(n) < 0.3, rng.integers(200, 1000, size=n), rng.integers(1, 20, size=n))
        else:
            costs = rng.integers(0, 5, size=n)  # zero-cost topologies included
        cc = int(rng.integers(1, max(2, int(costs.sum()) + 1)))
        cases.append((cc, multiple, costs))
    return cases


def check_partition(costs, starts):
    assert starts[0] == 0
    assert all(a < b for a, b in zip(starts, starts[1:]))
    assert starts[-1] < len(costs)


def plain_starts_ok(costs, cc, starts):
    cs = chunk_costs(costs, starts)
    bounds = list(starts) + [len(costs)]
    for i in range(len(starts) - 1):
        assert cs[i] >= cc  # closed only once it reached chunk_cost
        assert cs[i] - costs[bounds[i + 1] - 1] < cc  # ... and at the topology that reached it


def test_plain_rule(driver):
    rng = np.random.default_rng(7)
    cases = random_cases(rng, 300, 1)
    for (cc, _, costs), starts in zip(cases, driver(cases)):
        check_partition(costs, starts)
        plain_starts_ok(costs, cc, starts)


def test_one_chunk_without_a_cost(driver):
    costs = np.array([5, 7, 1])
    assert driver([(0, 1, costs), (0, 4, costs)]) == [[0], [0]]


def test_matrix_core_rule_is_the_min_max_partition(driver):
    rng = np.random.default_rng(11)
    cases = random_cases(rng, 300, 4)
    plain = driver([(cc, 1, costs) for cc, _, costs in cases])
    for (cc, _, costs), p, starts in zip(cases, plain, driver(cases)):
        check_partition(costs, starts)
        target = min(-(-len(p) // 4) * 4, len(costs))
        assert len(starts) <= target
        best = min_max_partition(list(costs), target)
        assert max(chunk_costs(costs, starts)) == best
        assert best <= max(chunk_costs(costs, p))


def test_heavy_tail_example(driver):
    """Costs 5, 1, 1, 1, 1, 5 with chunk_cost 1: the plain rule makes six chunks (two
    workgroups, the second holding a 5 beside two empty slots); the matrix-core rule packs
    the light middle into one chunk: 5 | 1 1 1 1 | 5."""
    costs = np.array([5, 1, 1, 1, 1, 5])
    plain, mf = driver([(1, 1, costs), (1, 4, costs)])
    assert plain == [0, 1, 2, 3, 4, 5]
    assert chunk_costs(costs, mf) == [5, 4, 5]
