"""Topology enumeration (src/diagrammatics.jl) against test/diagrammatics.jl and
README.md:178-201.  Both the oracle's and the host mirror's enumerations are checked."""
import ctypes as C
from itertools import permutations

import numpy as np
import pytest

import oracle_lib as O
from qiw import diagrammatics as dg


def dfact(n):
    r = 1
    while n > 1:
        r *= n
        n -= 2
    return r


def orc_tops(order, k=-1, ext=False):
    n = O.lib().orc_topologies(order, k, int(ext), 0, None, None)
    pairs = np.zeros((max(n, 1), 2 * max(order, 1)), np.int32)
    par = np.zeros(max(n, 1), np.int32)
    O.lib().orc_topologies(order, k, int(ext), n, O._p(pairs, C.c_int32), O._p(par, C.c_int32))
    return [([(int(pairs[t, 2 * a]), int(pairs[t, 2 * a + 1])) for a in range(order)], int(par[t])) for t in range(n)]


def perm_sign(seq):
    seq = list(seq)
    s = 1
    for i in range(len(seq)):
        for j in range(i + 1, len(seq)):
            if seq[i] > seq[j]:
                s = -s
    return s


@pytest.mark.parametrize("n", range(1, 6))
def test_counts_and_parity(n):
    tops = orc_tops(n)
    assert len(tops) == dfact(2 * n - 1)
    host = dg.get_topologies_at_order(n)
    assert [(t.pairs, t.parity) for t in host] == tops
    for pairs, parity in tops:
        flat = [x for p in pairs for x in p]
        assert sorted(flat) == list(range(1, 2 * n + 1))
        assert parity == perm_sign(flat)  # levicivita
        for k in range(1, n + 1):        # (-1)^k identity of test/diagrammatics.jl:44-51
            ndk = sum(1 for p in pairs if (p[0] <= k < p[1]) or (p[1] <= k < p[0]))
            assert (-1) ** k == (-1) ** ndk


def test_irreducible_chord_diagrams_A000699():
    assert [len(orc_tops(n, 1)) for n in range(1, 6)] == [1, 1, 4, 27, 248]


def test_readme_group_counts():
    assert [len(orc_tops(3, k)) for k in range(1, 6)] == [4, 6, 7, 6, 4]
    assert [len(orc_tops(4, k)) for k in range(1, 8)] == [27, 36, 40, 42, 40, 36, 27]
    for o in range(1, 5):
        for k in range(1, 2 * o):
            host = dg.get_topologies_at_order(o, k, with_external_arc=True)
            assert [(t.pairs, t.parity) for t in host] == orc_tops(o, k, True)
