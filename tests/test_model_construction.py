"""Caller-side construction of an expansion problem (src/expansion.jl:222-291): the
autopartition of the Fock space, the pairs generated from a hybridization matrix and an
nn-interaction, and the hybridization of a semi-elliptic bath.  KeldyshED and Keldysh are
not in the image: the partition is checked by the properties that define it, the bath
function against SciPy quadrature and its sum rule, and the constructed model against the
hand-built one on the oracle."""
import numpy as np
import pytest

import oracle_lib as O
from qiw import model as M


def anderson_H(f, eps=0.1, U=1.0):
    return eps * (f.n("up") + f.n("dn")) + U * f.n("up") @ f.n("dn")


def kanamori_H(n_orb=2, U=2.0, J=0.2):
    labels = [(a, s) for a in range(n_orb) for s in ("up", "dn")]
    f = M.FockSpace(labels)
    n = {l: f.n(l) for l in labels}
    cd = {l: f.c_dag(l) for l in labels}
    c = {l: f.c(l) for l in labels}
    H = np.zeros((f.dim, f.dim))
    for a in range(n_orb):
        H += U * n[(a, "up")] @ n[(a, "dn")]
        for b in range(a + 1, n_orb):
            for s in ("up", "dn"):
                t = "dn" if s == "up" else "up"
                H += (U - 2 * J) * n[(a, s)] @ n[(b, t)] + (U - 3 * J) * n[(a, s)] @ n[(b, s)]
    for a in range(n_orb):
        for b in range(n_orb):
            if a != b:
                H -= J * cd[(a, "up")] @ c[(a, "dn")] @ cd[(b, "dn")] @ c[(b, "up")]
                H += J * cd[(a, "up")] @ cd[(a, "dn")] @ c[(b, "dn")] @ c[(b, "up")]
    return f, H, labels


def check_partition(f, H, subspaces, breakers=()):
    assert sorted(b for s in subspaces for b in s) == list(range(f.dim))
    sec = np.empty(f.dim, dtype=int)
    for i, s in enumerate(subspaces):
        sec[s] = i
    for Mx in [H] + list(breakers):  # block diagonal
        r, c = np.nonzero(np.abs(Mx) > 1e-13)
        assert np.all(sec[r] == sec[c])
    for l in f.labels:  # one block per column for every c and c^dag
        for C in (f.c(l), f.c_dag(l)):
            r, c = np.nonzero(np.abs(C) > 0)
            for s in range(len(subspaces)):
                assert len(set(sec[r[sec[c] == s]])) <= 1


def test_autopartition_anderson_is_occupation_resolved():
    f = M.FockSpace(["up", "dn"])
    H = anderson_H(f)
    sub = M.autopartition(H, f)
    assert sub == [[0], [1], [2], [3]]
    check_partition(f, H, sub)


def test_autopartition_kanamori_refines_particle_number():
    f, H, labels = kanamori_H()
    breakers = []
    for l in labels:
        breakers += [f.c_dag(l) @ f.c(l), f.c(l) @ f.c_dag(l)]
    sub = M.autopartition(H, f, breakers)
    check_partition(f, H, sub, breakers)
    # every subspace has one particle number and one S_z (the Hamiltonian conserves both)
    for s in sub:
        assert len({bin(b).count("1") for b in s}) == 1
        sz = {sum((1 if l[1] == "up" else -1) * f.occupation(b, l) for l in labels) for b in s}
        assert len(sz) == 1
    assert len(sub) > 5 and max(len(s) for s in sub) < 6  # finer than the N blocks (1,4,6,4,1)
    # each subspace is one connected component of H, or is held together by the one-block
    # rule of an operator: the states an operator reaches from one source subspace
    sec = {b: i for i, s_ in enumerate(sub) for b in s_}
    for i, s_ in enumerate(sub):
        comp, todo = {s_[0]}, [s_[0]]
        while todo:  # closure under H, then under co-images of one source subspace
            b = todo.pop()
            linked = {int(x) for x in np.nonzero(np.abs(H[:, b]) > 1e-13)[0]}
            for l in labels:
                for C in (f.c(l), f.c_dag(l)):
                    for src in range(len(sub)):
                        img = {int(r) for c_ in sub[src] for r in np.nonzero(C[:, c_])[0]}
                        if b in img:
                            linked |= {x for x in img if sec[x] == i}
            for x in linked - comp:
                comp.add(x)
                todo.append(x)
        assert comp == set(s_)


def test_bethe_bath_against_quadrature_and_sum_rule():
    from scipy.integrate import quad
    beta, t = 10.0, 1.0
    grid = M.ImaginaryTimeGrid(beta, 9)
    g = M.bethe_gf(t, beta, grid).values
    for k, tau in enumerate(grid.tau):
        def integrand(e):
            rho = np.sqrt(max(4 * t * t - e * e, 0.0)) / (2 * np.pi * t * t)
            return rho * np.exp(-e * tau) / (1 + np.exp(-beta * e))
        ref, _ = quad(integrand, -2 * t, 2 * t, limit=200, epsabs=1e-14, epsrel=1e-13)
        assert abs(g[k] - (-1j * ref)) < 1e-10
    assert abs(g[0] + g[-1] - (-1j)) < 1e-13  # G(0+) + G(beta-) = -i for fermions
    assert abs(g[0] - g[-1]) < 1e-13  # particle-hole symmetric bath


def test_constructed_anderson_matches_hand_built_on_oracle():
    """Expansion(H, soi, grid; hybridization) for the Anderson model gives the hand-built
    model's bare step, sector by sector (sectors matched through their basis state)."""
    beta, n_tau = 10.0, 12
    f = M.FockSpace(["up", "dn"])
    grid = M.ImaginaryTimeGrid(beta, n_tau)
    d = 0.25 * M.bethe_gf(1.0, beta, grid).values
    hyb = np.zeros((2, 2, n_tau), complex)
    hyb[0, 0] = d
    hyb[1, 1] = d
    built = M.Expansion.from_hamiltonian(anderson_H(f), f, grid, hyb)
    hand = M.anderson_single_orbital(beta=beta, n_tau=n_tau)
    assert len(built.pairs) == 4
    v_b, _ = O.inchworm_step_bare(built, [0, 1, 2], 2 ** 6)
    v_h, _ = O.inchworm_step_bare(hand, [0, 1, 2], 2 ** 6)
    for sb, states in enumerate(built.ed.subspaces):
        sh = hand.ed.subspaces.index(states)
        assert abs(v_b[built.bt_offsets[sb]] - v_h[hand.bt_offsets[sh]]) < 1e-14


def test_nn_interaction_pairs():
    f = M.FockSpace(["up", "dn"])
    grid = M.ImaginaryTimeGrid(10.0, 8)
    hyb = np.zeros((2, 2, 8), complex)
    U = np.zeros((2, 2, 8))
    U[0, 1] = 0.3 * np.cosh(grid.tau - 5.0) / np.cosh(5.0)
    e = M.Expansion.from_hamiltonian(anderson_H(f), f, grid, hyb, nn_interaction=U)
    assert len(e.pairs) == 1
    p = e.pair_operator_mat[0]
    assert all(si == sf for si, (sf, _) in p[0].items())  # n_i keeps every sector


@pytest.mark.parametrize("bad", [np.zeros((2, 2, 5))])
def test_shape_checked(bad):
    f = M.FockSpace(["up", "dn"])
    with pytest.raises(AssertionError):
        M.Expansion.from_hamiltonian(anderson_H(f), f, M.ImaginaryTimeGrid(10.0, 8), bad)
