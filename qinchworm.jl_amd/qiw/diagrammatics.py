"""Diagram topologies (src/diagrammatics.jl).

A topology of order n pairs the points 1..2n into n arcs (first < second within each
generated pair) and carries the parity of the permutation.  Topology point 1 is the
latest time of a configuration.
"""
from dataclasses import dataclass
from typing import List, Tuple


@dataclass
class Topology:
    order: int
    pairs: List[Tuple[int, int]]
    parity: int


def _generate(partial_pairs, parity, unpaired, out):
    # generate_topologies_impl (src/diagrammatics.jl:288-303): pairing the first unpaired
    # point with the i-th one adds one transposition per skipped point, sign (-1)^i.
    if not unpaired:
        out.append(Topology(len(partial_pairs), list(partial_pairs), parity))
        return
    for i in range(2, len(unpaired) + 1):
        p = (unpaired[0], unpaired[i - 1])
        rest = [unpaired[k] for k in range(len(unpaired)) if k not in (0, i - 1)]
        _generate(partial_pairs + [p], parity * (-1) ** i, rest, out)


def generate_topologies(n):
    out = []
    _generate([], 1, list(range(1, 2 * n + 1)), out)
    return out


def iscrossing(p1, p2):
    a1, b1 = sorted(p1)
    a2, b2 = sorted(p2)
    return (a1 < a2 < b1 < b2) or (a2 < a1 < b2 < b1)


def _doubly_k(p, k):
    return (p[0] <= k < p[1]) or (p[1] <= k < p[0])


def is_doubly_k_connected(top, k):
    """Every connected component of the arc-crossing graph holds an arc with one end
    <= k and the other > k (src/diagrammatics.jl:231-235)."""
    stack = [p for p in top.pairs if _doubly_k(p, k)]
    disconnected = [p for p in top.pairs if not _doubly_k(p, k)]
    while stack:
        pc = stack.pop()
        for i in range(len(disconnected) - 1, -1, -1):
            if iscrossing(pc, disconnected[i]):
                stack.append(disconnected.pop(i))
    return not disconnected


def get_topologies_at_order(order, k=None, with_external_arc=False):
    """src/diagrammatics.jl:322-337."""
    tops = generate_topologies(order)
    if k is None:
        return tops
    tops = [t for t in tops if is_doubly_k_connected(t, k)]
    if with_external_arc:
        tops = [Topology(t.order, t.pairs, (-1) ** k * t.parity) for t in tops]
    return tops
