"""Caller-side construction of an expansion problem (the part the reference delegates to
KeldyshED.jl and Keldysh.jl): Fock-space operators, exact diagonalization in symmetry
sectors, operator blocks in the eigenbasis, hybridization tables from a bath density of
states, and the `Expansion` container (src/expansion.jl:107-310).

Only setup work happens here; none of it is on the sampled hot path.
"""
import numpy as np

from . import abi


# ------------------------------------------------------------------------------------
# Fock space with Jordan-Wigner fermions
# ------------------------------------------------------------------------------------
class FockSpace:
    """Fermionic Fock space over an ordered list of mode labels.  Basis state b has bit i
    set when mode i is occupied; c_i carries the sign (-1)^(number of occupied modes
    below i)."""

    def __init__(self, labels):
        self.labels = list(labels)
        self.n_modes = len(self.labels)
        self.dim = 1 << self.n_modes

    def index(self, label):
        return self.labels.index(label)

    def c(self, label):
        i = self.index(label)
        M = np.zeros((self.dim, self.dim))
        for b in range(self.dim):
            if b >> i & 1:
                sign = -1.0 if bin(b & ((1 << i) - 1)).count("1") % 2 else 1.0
                M[b ^ (1 << i), b] = sign
        return M

    def c_dag(self, label):
        return self.c(label).T.copy()

    def n(self, label):
        return self.c_dag(label) @ self.c(label)

    def occupation(self, b, label):
        return b >> self.index(label) & 1


# ------------------------------------------------------------------------------------
# Exact diagonalization in sectors
# ------------------------------------------------------------------------------------
class EDCore:
    """Diagonalizes H inside the sectors given by `sector_key(basis_state)`; sectors are
    ordered by sorted key.  With sector_key None the sectors are autopartition(H, fock,
    symmetry_breakers).  Operators are represented by their blocks between eigenbases
    (the role of KeldyshED's EDCore / operator_blocks)."""

    def __init__(self, H, fock, sector_key=None, symmetry_breakers=()):
        self.fock = fock
        keys = {}
        if sector_key is None:
            for s, states in enumerate(autopartition(H, fock, symmetry_breakers)):
                keys[s] = states
        else:
            for b in range(fock.dim):
                keys.setdefault(sector_key(b), []).append(b)
        self.sector_keys = sorted(keys)
        self.subspaces = [keys[k] for k in self.sector_keys]
        sec_of = np.empty(fock.dim, dtype=int)
        for s, states in enumerate(self.subspaces):
            sec_of[states] = s
        off = H.copy()
        for s, states in enumerate(self.subspaces):
            off[np.ix_(states, states)] = 0.0
        if np.abs(off).max(initial=0.0) > 1e-12:
            raise ValueError("Hamiltonian is not block-diagonal in the requested sectors")
        self.energies, self.U = [], []
        for states in self.subspaces:
            E, U = np.linalg.eigh(H[np.ix_(states, states)])
            self.energies.append(E)
            self.U.append(U)
        self.dims = [len(s) for s in self.subspaces]

    def operator_blocks(self, O, tol=1e-13):
        """SectorBlockMatrix of O: dict s_i -> (s_f, m_f x m_i block in the eigenbases)
        (operator_to_sector_block_matrix, src/sector_block_matrix.jl:52-66)."""
        sbm = {}
        for si, st_i in enumerate(self.subspaces):
            for sf, st_f in enumerate(self.subspaces):
                blk = O[np.ix_(st_f, st_i)]
                if np.abs(blk).max(initial=0.0) <= tol:
                    continue
                if si in sbm:
                    raise ValueError("Operator is not representable by a SectorBlockMatrix "
                                     "(more than one non-zero blocks per column)")
                sbm[si] = (sf, self.U[sf].T.conj() @ blk @ self.U[si])
        return sbm

    def partition_function(self, beta):
        emin = min(e.min() for e in self.energies)
        return sum(np.exp(-beta * (e - emin)).sum() for e in self.energies) * np.exp(-beta * emin)

    def density_matrix(self, beta):
        emin = min(e.min() for e in self.energies)
        w = [np.exp(-beta * (e - emin)) for e in self.energies]
        z = sum(x.sum() for x in w)
        return [np.diag(x / z) for x in w]


def autopartition(H, fock, symmetry_breakers=(), tol=1e-13):
    """The finest partition of the Fock basis into invariant subspaces, the role of
    KeldyshED's autopartition behind EDCore(H, soi; symmetry_breakers)
    (src/expansion.jl:283).  Phase 1: basis states coupled by H or by a symmetry breaker
    share a subspace (union-find).  Phase 2: every c_i and c^dag_i must map each subspace
    into a single subspace (one block per column of its SectorBlockMatrix): the target
    subspaces of one source are merged, for both operators of every mode, until nothing
    changes.  Returns the subspaces as sorted lists of basis states, ordered by their
    smallest state."""
    parent = list(range(fock.dim))

    def find(x):
        while parent[x] != x:
            parent[x] = parent[parent[x]]
            x = parent[x]
        return x

    def union(x, y):
        x, y = find(x), find(y)
        if x != y:
            parent[max(x, y)] = min(x, y)
            return True
        return False

    for M in [H] + list(symmetry_breakers):
        rows, cols = np.nonzero(np.abs(M) > tol)
        for r, c in zip(rows, cols):
            union(int(r), int(c))
    ops = []
    for label in fock.labels:
        C = fock.c(label)
        ops += [C, C.T.copy()]
    changed = True
    while changed:
        changed = False
        for C in ops:
            targets = {}
            rows, cols = np.nonzero(np.abs(C) > tol)
            for r, c in zip(rows, cols):
                targets.setdefault(find(int(c)), set()).add(find(int(r)))
            for tset in targets.values():
                tl = sorted(tset)
                for t in tl[1:]:
                    changed |= union(tl[0], t)
    groups = {}
    for b in range(fock.dim):
        groups.setdefault(find(b), []).append(b)
    return sorted(groups.values(), key=lambda g: g[0])


# ------------------------------------------------------------------------------------
# Scalar imaginary-time Green's functions on the uniform grid
# ------------------------------------------------------------------------------------
class ImaginaryTimeGrid:
    def __init__(self, beta, n_tau):
        self.beta = float(beta)
        self.n_tau = int(n_tau)
        self.tau = np.arange(self.n_tau) * (self.beta / (self.n_tau - 1))

    def __len__(self):
        return self.n_tau


class ScalarGF:
    """g(tau_j) on the grid plus its interpolation kind and statistics sign."""

    def __init__(self, values, kind=abi.QIW_GF_GRID_LINEAR, xi=-1):
        self.values = np.asarray(values, dtype=np.complex128)
        self.kind = kind
        self.xi = xi

    def __mul__(self, a):
        return ScalarGF(self.values * a, self.kind, self.xi)

    __rmul__ = __mul__


def spline_interpolated(g):
    """SplineInterpolatedGF(g) (src/spline_gf.jl:65-72): same values, cubic-spline lookups."""
    return ScalarGF(g.values, abi.QIW_GF_CUBIC_SPLINE, g.xi)


def ph_conj(g):
    """g(beta - tau) (src/utility.jl:210-217, src/spline_gf.jl:250-255).  For a DLR
    function the pole sum's kernel obeys K(beta - tau, w) = K(tau, -w), so the conjugate
    is the same coefficients at the negated frequencies (exact; the reference refits on its
    DLR nodes, src/keldysh_dlr.jl:262-270, which agrees to the DLR tolerance)."""
    if isinstance(g, DLRGF):
        return DLRGF(-g.omega, g.coef, g.beta, g.tau, g.xi)
    return ScalarGF(g.values[::-1].copy(), g.kind, g.xi)


# ------------------------------------------------------------------------------------
# Discrete Lehmann representation (keldysh_dlr.DLRImaginaryTimeGF, src/keldysh_dlr.jl)
# ------------------------------------------------------------------------------------
def dlr_kernel(tau, w, beta):
    """The fermionic kernel e^{-w tau} / (1 + e^{-w beta}), 0 <= tau <= beta, in the form
    whose exponents are never positive; tau and w broadcast."""
    tau, w = np.broadcast_arrays(np.asarray(tau, float), np.asarray(w, float))
    pos = w >= 0
    return np.where(pos, np.exp(-np.abs(w) * tau) / (1.0 + np.exp(-np.abs(w) * beta)),
                    np.exp(-np.abs(w) * (beta - tau)) / (1.0 + np.exp(-np.abs(w) * beta)))


class DLRGF:
    """g(tau) = sum_k coef_k K(tau, omega_k) (scalar DLRImaginaryTimeGF; pair propagators of
    kind QIW_GF_DLR).  `values` holds g on the uniform grid (the qiw_model's grid values,
    not read for this kind)."""
    kind = abi.QIW_GF_DLR

    def __init__(self, omega, coef, beta, tau, xi=-1):
        self.omega = np.asarray(omega, float)
        self.coef = np.asarray(coef, np.complex128)
        self.beta = float(beta)
        self.tau = np.asarray(tau, float)
        self.xi = xi
        self.values = self(self.tau)

    def __call__(self, tau):
        return dlr_kernel(np.asarray(tau, float)[:, None], self.omega[None, :], self.beta) @ self.coef

    def __mul__(self, a):
        return DLRGF(self.omega, self.coef * a, self.beta, self.tau, self.xi)

    __rmul__ = __mul__


def _cheb_panels(edges, n):
    """Chebyshev nodes of the first kind, n per panel [edges[i], edges[i+1]]."""
    x = np.cos((2 * np.arange(n) + 1) * np.pi / (2 * n))[::-1]
    return np.concatenate([0.5 * (a + b) + 0.5 * (b - a) * x for a, b in zip(edges[:-1], edges[1:])])


def dlr_frequencies(Euv, beta, rtol, n_cheb=24):
    """Real frequencies of a fermionic DLR of cutoff Euv at tolerance rtol, and the time
    nodes its coefficients are fitted on, after the construction of Lehmann.jl's DLRGrid
    (Euv, beta, rtol, isFermi) (Kaye, Chen and Parcollet, Phys. Rev. B 105, 235115
    (2022)): the kernel K(tau, w) is sampled on composite Chebyshev grids refined
    dyadically towards w = 0 and towards tau = 0 and beta, and a column-pivoted QR picks the
    frequencies up to relative size rtol.  The frequencies are picked in pairs +-w (a QR
    of K(tau, w) stacked on K(beta - tau, w) = K(tau, -w), w >= 0), so that the particle-
    hole conjugate of a DLR is again one on the same frequencies and a particle-hole
    symmetric function has a symmetric DLR; the fit nodes are the symmetric fine time grid."""
    from scipy.linalg import qr
    lam = Euv * beta
    m = max(int(np.ceil(np.log2(max(lam, 2.0)))), 1)
    wedges = np.concatenate([[0.0], lam * 2.0 ** -np.arange(m, -1, -1.0)])
    xp = _cheb_panels(wedges, n_cheb)                         # w * beta >= 0
    tedges = np.concatenate([[0.0], 0.5 * 2.0 ** -np.arange(m + 4, -1, -1.0)])
    th = _cheb_panels(tedges, n_cheb)
    t = np.concatenate([th, 1.0 - th[::-1]])                  # tau / beta, symmetric
    A = np.vstack([dlr_kernel(t[:, None], xp[None, :], 1.0), dlr_kernel(1.0 - t[:, None], xp[None, :], 1.0)])
    _, R, piv = qr(A, mode="economic", pivoting=True)
    d = np.abs(np.diag(R))
    xs = np.sort(xp[piv[:int(np.sum(d > rtol * d[0]))]])
    return np.concatenate([-xs[::-1], xs]) / beta, t * beta


def dlr_fit(f, Euv, beta, grid, rtol=1e-6, xi=-1):
    """The DLR of the function f(tau) (complex, 0 <= tau <= beta): coefficients by least
    squares on the symmetric fine time grid (Lehmann.jl's tau2dlr fits on its DLR nodes;
    DLRImaginaryTimeGF(f, grid), src/keldysh_dlr.jl:212-236)."""
    omega, tau = dlr_frequencies(Euv, beta, rtol)
    g = np.asarray(f(tau), np.complex128)
    coef = np.linalg.lstsq(dlr_kernel(tau[:, None], omega[None, :], beta), g, rcond=None)[0]
    return DLRGF(omega, coef, beta, grid.tau, xi)


def bethe_values(t, beta, tau, n_quad=400, eps=0.0):
    """G(tau) of the semi-elliptic DOS of half-bandwidth 2t centred at eps, at any taus
    (the quadrature of bethe_gf)."""
    x, w = np.polynomial.legendre.leggauss(n_quad)
    theta = 0.5 * np.pi * (x + 1.0)
    wt = 0.5 * np.pi * w * (2.0 / np.pi) * np.sin(theta) ** 2
    e = eps + 2.0 * t * np.cos(theta)
    return -1j * (wt[:, None] * _fermi_weighted_exp(e, beta, np.asarray(tau, float))).sum(axis=0)


def _fermi_weighted_exp(eps, beta, tau):
    """(1 - n_F(eps)) exp(-eps tau) for 0 <= tau <= beta without overflow."""
    eps = np.asarray(eps, float)[..., None]
    tau = np.asarray(tau, float)[None, ...]
    pos = eps >= 0
    a = np.where(pos, np.exp(-eps * tau) / (1.0 + np.exp(-beta * np.abs(eps))),
                 np.exp(eps * (beta - tau)) / (1.0 + np.exp(-beta * np.abs(eps))))
    return a


def delta_dos_gf(eps, beta, grid, weight=1.0):
    """Free fermion of level eps, G(tau) = -i (1 - n_F(eps)) e^{-eps tau} (kd.DeltaDOS)."""
    return ScalarGF(-1j * weight * _fermi_weighted_exp([eps], beta, grid.tau)[0])


def bethe_gf(t, beta, grid, n_quad=400, eps=0.0):
    """Semi-elliptic density of states of half-bandwidth 2t centred at eps
    (kd.bethe_dos(t=t, ϵ=eps)), G(tau) = -i int d w rho(w) (1 - n_F(w)) e^{-w tau}, by
    Gauss-Legendre on w = eps + 2t cos(theta), where rho d w = (2/pi) sin^2(theta) d theta."""
    x, w = np.polynomial.legendre.leggauss(n_quad)
    theta = 0.5 * np.pi * (x + 1.0)
    wt = 0.5 * np.pi * w * (2.0 / np.pi) * np.sin(theta) ** 2
    eps = eps + 2.0 * t * np.cos(theta)
    vals = (wt[:, None] * _fermi_weighted_exp(eps, beta, grid.tau)).sum(axis=0)
    return ScalarGF(-1j * vals)


# ------------------------------------------------------------------------------------
# Expansion
# ------------------------------------------------------------------------------------
class InteractionPair:
    """InteractionPair(operator_f, operator_i, propagator) (src/expansion.jl:79-90):
    operator_i sits at the earlier time, operator_f at the later one."""

    def __init__(self, operator_f, operator_i, propagator):
        self.operator_f = operator_f
        self.operator_i = operator_i
        self.propagator = propagator


class Expansion:
    """Strong-coupling expansion problem (src/expansion.jl:107-197).  `P` holds the bold
    propagator table in BT layout (n_tau, sum m^2) complex, updated by inchworm()."""

    def __init__(self, ed, grid, interaction_pairs, corr_operators=(), interpolate_ppgf=False):
        self.ed = ed
        self.grid = grid
        self.pairs = list(interaction_pairs)
        self.interpolate_ppgf = bool(interpolate_ppgf)
        self.identity_mat = {s: (s, np.eye(d)) for s, d in enumerate(ed.dims)}
        self.pair_operator_mat = [(ed.operator_blocks(p.operator_i), ed.operator_blocks(p.operator_f))
                                  for p in self.pairs]
        self.corr_operators = []
        self.corr_operators_mat = []
        for A, B in corr_operators:
            self.add_corr_operators((A, B))
        self.subspace_attachable_pairs = [
            [p for p, ops in enumerate(self.pair_operator_mat) if s in ops[0]]
            for s in range(len(ed.dims))]
        self.bt_offsets = np.cumsum([0] + [d * d for d in ed.dims])[:-1]
        self.bt_size = int(sum(d * d for d in ed.dims))
        self.P = self.atomic_ppgf()
        self.P_der0 = self.initial_ppgf_derivative()
        self.P_n_known = 1 if self.interpolate_ppgf else grid.n_tau

    @property
    def ppgf_kind(self):
        return abi.QIW_GF_INC_SPLINE if self.interpolate_ppgf else abi.QIW_GF_GRID_LINEAR

    @classmethod
    def from_hamiltonian(cls, hamiltonian, fock, grid, hybridization, nn_interaction=None, corr_operators=(),
                         interpolate_ppgf=False, tol=0.0):
        """The higher-level constructor Expansion(hamiltonian, soi, grid; hybridization,
        nn_interaction, corr_operators, interpolate_ppgf) (src/expansion.jl:222-291).
        hybridization: complex array [n_modes][n_modes][n_tau] of Delta_ij(tau_k), modes in
        fock.labels order; nn_interaction: real array of U_ij(tau_k) of the same shape.
        For each non-zero Delta_ij: the pairs (c^dag_i, c_j, Delta_ij) and (c_j, c^dag_i,
        Delta_ij(beta - tau)), and c^dag_i c_j, c_j c^dag_i as symmetry breakers of the
        autopartition; for each non-zero U_ij the pair (n_i, n_j, U_ij)."""
        labels = fock.labels
        hyb = np.asarray(hybridization, dtype=np.complex128)
        assert hyb.shape == (len(labels), len(labels), grid.n_tau)
        pairs, breakers = [], []
        for i, li in enumerate(labels):
            for j, lj in enumerate(labels):
                vals = hyb[i, j]
                if np.all(np.abs(vals) <= tol):
                    continue
                delta = ScalarGF(vals)
                cd_i, c_j = fock.c_dag(li), fock.c(lj)
                pairs.append(InteractionPair(cd_i, c_j, delta))
                pairs.append(InteractionPair(c_j, cd_i, ph_conj(delta)))
                breakers += [cd_i @ c_j, c_j @ cd_i]
        if nn_interaction is not None:
            U = np.asarray(nn_interaction, dtype=np.float64)
            assert U.shape == (len(labels), len(labels), grid.n_tau)
            for i, li in enumerate(labels):
                for j, lj in enumerate(labels):
                    if np.all(np.abs(U[i, j]) <= tol):
                        continue
                    pairs.append(InteractionPair(fock.n(li), fock.n(lj), ScalarGF(U[i, j].astype(np.complex128))))
        ed = EDCore(hamiltonian, fock, symmetry_breakers=breakers)
        exp = cls(ed, grid, pairs, corr_operators, interpolate_ppgf=interpolate_ppgf)
        exp.fock = fock
        return exp

    def add_corr_operators(self, ops):
        """add_corr_operators! (src/expansion.jl:303-310)."""
        A, B = ops
        self.corr_operators.append((A, B))
        self.corr_operators_mat.append((self.ed.operator_blocks(A), self.ed.operator_blocks(B)))
        self._struct = None

    def _lambda(self):
        beta = self.grid.beta
        return np.log(self.ed.partition_function(beta)) / beta

    def atomic_ppgf(self):
        """P_s(tau) = -i exp(-tau (E_s + lambda)), Tr iP(beta) = 1 (src/ppgf.jl:144-207)."""
        lam = self._lambda()
        P = np.zeros((self.grid.n_tau, self.bt_size), dtype=np.complex128)
        for s, E in enumerate(self.ed.energies):
            d = len(E)
            o = self.bt_offsets[s]
            for i in range(d):
                P[:, o + i + i * d] = -1j * np.exp(-self.grid.tau * (E[i] + lam))
        return P

    def initial_ppgf_derivative(self):
        """dP/dtau at 0: -i diag(-E - lambda) (src/ppgf.jl:710-720)."""
        lam = self._lambda()
        der = np.zeros(self.bt_size, dtype=np.complex128)
        for s, E in enumerate(self.ed.energies):
            d = len(E)
            o = self.bt_offsets[s]
            for i in range(d):
                der[o + i + i * d] = -1j * (-E[i] - lam)
        return der

    def block(self, table_row, s):
        d = self.ed.dims[s]
        o = self.bt_offsets[s]
        return table_row[o:o + d * d].reshape(d, d, order="F")

    def to_struct(self):
        """The qiw_model this expansion is handed over as."""
        props = np.array([p.propagator.values for p in self.pairs]).reshape(len(self.pairs), self.grid.n_tau)
        dlr = None
        ranks = [len(p.propagator.omega) for p in self.pairs if p.propagator.kind == abi.QIW_GF_DLR]
        if ranks:
            r = max(ranks)
            w = np.zeros((len(self.pairs), r))
            c = np.zeros((len(self.pairs), r), np.complex128)
            for i, p in enumerate(self.pairs):
                if p.propagator.kind == abi.QIW_GF_DLR:
                    k = len(p.propagator.omega)
                    w[i, :k], c[i, :k] = p.propagator.omega, p.propagator.coef
            dlr = (w, c)
        return abi.ModelStruct(
            self.grid.beta, self.grid.n_tau, self.ed.dims, self.ed.energies, self.ppgf_kind,
            self.pair_operator_mat, props, [p.propagator.kind for p in self.pairs],
            [p.propagator.xi for p in self.pairs], self.corr_operators_mat, dlr=dlr)


# ------------------------------------------------------------------------------------
# Model families of the benchmark configurations
# ------------------------------------------------------------------------------------
def anderson_single_orbital(beta=10.0, n_tau=64, eps=0.1, U=1.0, V=0.5, D=2.0,
                            interpolate=False, with_correlators=False):
    """Single-orbital Anderson impurity H = eps (n_up + n_dn) + U n_up n_dn, 4 states in
    1x1 blocks by (N_up, N_dn), semi-elliptic bath of half-bandwidth D, Delta = V^2 G_bethe,
    pairs c^dag_s/c_s and their particle-hole conjugates (README.md:55-137)."""
    fock = FockSpace(["up", "dn"])
    H = eps * (fock.n("up") + fock.n("dn")) + U * fock.n("up") @ fock.n("dn")
    ed = EDCore(H, fock, lambda b: (fock.occupation(b, "up"), fock.occupation(b, "dn")))
    grid = ImaginaryTimeGrid(beta, n_tau)
    delta = V ** 2 * bethe_gf(D / 2.0, beta, grid)
    if interpolate:
        delta = spline_interpolated(delta)
    pairs = []
    for s in ["up", "dn"]:
        pairs.append(InteractionPair(fock.c_dag(s), fock.c(s), delta))
        pairs.append(InteractionPair(fock.c(s), fock.c_dag(s), ph_conj(delta)))
    corr = []
    if with_correlators:
        corr = [(-fock.c("up"), fock.c_dag("up")),
                (fock.n("up"), fock.n("dn")),
                (fock.c_dag("up") @ fock.c("dn"), fock.c_dag("dn") @ fock.c("up"))]
    return Expansion(ed, grid, pairs, corr, interpolate_ppgf=interpolate)


def bethe_two_orbitals(n_tau, mu_bethe=0.25, beta=10.0, V=0.5, t_bethe=0.5):
    """The model of test/bethe.jl:60-88 and test/bethe_gf.jl:62-90: two non-interacting
    orbitals (H = -mu (n_1 + n_2), mu = 0) each coupled to a semi-elliptic bath of half-
    bandwidth 2 t_bethe centred at mu_bethe, Delta = V^2 G_bethe I(2), built by the
    higher-level Expansion constructor (src/expansion.jl:222-291)."""
    fock = FockSpace([1, 2])
    H = 0.0 * (fock.n(1) + fock.n(2))
    grid = ImaginaryTimeGrid(beta, n_tau)
    g = bethe_gf(t_bethe, beta, grid, eps=mu_bethe).values
    hyb = np.zeros((2, 2, n_tau), np.complex128)
    hyb[0, 0] = hyb[1, 1] = V ** 2 * g
    return Expansion.from_hamiltonian(H, fock, grid, hyb)


def bethe_two_orbitals_dlr(n_tau, mu_bethe=0.25, beta=10.0, V=0.5, t_bethe=0.5, Euv=1.25, rtol=1e-6):
    """The model of test/bethe_dlr.jl:55-97: two non-interacting orbitals, each with the
    semi-elliptic hybridization V^2 G_bethe(t = t_bethe, eps = mu_bethe) as a DLR
    (DLRGrid(Euv = 1.25, beta, rtol = 1e-6, isFermi)), pairs c^dag_i/c_i with Delta and
    c_i/c^dag_i with ph_conj(Delta) built by the low-level Expansion constructor."""
    fock = FockSpace([1, 2])
    H = 0.0 * (fock.n(1) + fock.n(2))
    ed = EDCore(H, fock, lambda b: (fock.occupation(b, 1), fock.occupation(b, 2)))
    grid = ImaginaryTimeGrid(beta, n_tau)
    delta = V ** 2 * dlr_fit(lambda tau: bethe_values(t_bethe, beta, tau, eps=mu_bethe), Euv, beta, grid, rtol)
    pairs = []
    for i in (1, 2):
        pairs.append(InteractionPair(fock.c_dag(i), fock.c(i), delta))
        pairs.append(InteractionPair(fock.c(i), fock.c_dag(i), ph_conj(delta)))
    exp = Expansion(ed, grid, pairs)
    exp.fock = fock
    return exp


def fock_density_matrix(expansion, P):
    """The density matrix in the Fock basis (KeldyshED full_hs_matrix(tofockbasis(...)))
    from i P(beta) normalised to unit trace; P is the (n_tau, sum m^2) table."""
    ed = expansion.ed
    rho = np.zeros((ed.fock.dim, ed.fock.dim), np.complex128)
    blocks, off = [], 0
    for s, m in enumerate(ed.dims):
        blocks.append((1j * P[-1, off:off + m * m]).reshape(m, m, order="F"))
        off += m * m
    z = sum(np.trace(b) for b in blocks)
    for st, U, b in zip(ed.subspaces, ed.U, blocks):
        rho[np.ix_(st, st)] = U @ (b / z) @ U.T.conj()
    return rho


def kanamori(n_orb, U, J, eps_shift, beta, n_tau, V, D=2.0, interpolate=False, dlr=False):
    """n_orb-orbital Hubbard-Kanamori impurity (spin-rotation invariant):
    H = mu sum n + U sum_a n_a,up n_a,dn + (U-2J) sum_{a<b,s} n_a,s n_b,-s
        + (U-3J) sum_{a<b,s} n_a,s n_b,s
        - J sum_{a!=b} c+_a,up c_a,dn c+_b,dn c_b,up + J sum_{a!=b} c+_a,up c+_a,dn c_b,dn c_b,up,
    with mu = eps_shift - [U/2 + (n_orb - 1)(2U - 5J)/2] (eps_shift = 0 is half filling).
    Sectors by total particle number; each spin-orbital couples to its own semi-elliptic
    bath (half-bandwidth D, hopping V): pairs c^dag_i/c_i and particle-hole conjugates."""
    labels = [(a, s) for a in range(n_orb) for s in ("up", "dn")]
    fock = FockSpace(labels)
    n = {l: fock.n(l) for l in labels}
    mu = eps_shift - (U / 2.0 + (n_orb - 1) * (2 * U - 5 * J) / 2.0)
    H = mu * sum(n.values())
    for a in range(n_orb):
        H = H + U * n[(a, "up")] @ n[(a, "dn")]
    for a in range(n_orb):
        for b in range(a + 1, n_orb):
            for s, sb in (("up", "dn"), ("dn", "up")):
                H = H + (U - 2 * J) * n[(a, s)] @ n[(b, sb)]
                H = H + (U - 3 * J) * n[(a, s)] @ n[(b, s)]
    cd = {l: fock.c_dag(l) for l in labels}
    c = {l: fock.c(l) for l in labels}
    for a in range(n_orb):
        for b in range(n_orb):
            if a == b:
                continue
            H = H - J * cd[(a, "up")] @ c[(a, "dn")] @ cd[(b, "dn")] @ c[(b, "up")]
            H = H + J * cd[(a, "up")] @ cd[(a, "dn")] @ c[(b, "dn")] @ c[(b, "up")]
    ed = EDCore(H, fock, lambda b: bin(b).count("1"))
    grid = ImaginaryTimeGrid(beta, n_tau)
    if dlr:  # the same bath as a DLR (cutoff the band edge D, tolerance 1e-6)
        delta = V ** 2 * dlr_fit(lambda tau: bethe_values(D / 2.0, beta, tau), D, beta, grid, rtol=1e-6)
    else:
        delta = V ** 2 * bethe_gf(D / 2.0, beta, grid)
    if interpolate:
        delta = spline_interpolated(delta)
    pairs = []
    for l in labels:
        pairs.append(InteractionPair(cd[l], c[l], delta))
        pairs.append(InteractionPair(c[l], cd[l], ph_conj(delta)))
    return Expansion(ed, grid, pairs, interpolate_ppgf=interpolate)


def bethe_single_level(n_tau, beta=10.0, t_bethe=2.0, mu=0.0):
    """The model of the reference's randomized-qMC study (bench/rqmc/rqmc.jl:37-63): one
    spinless level, H = -mu n (mu = 0), semi-elliptic bath kd.bethe_dos(t = t_bethe / 2)
    with V = t_bethe / 2, Delta = V^2 G_bethe on the grid, pairs c^dag/c and their
    particle-hole conjugates, correlator (c, c^dag).  The exact impurity Green's function
    is the lattice one, G = -Delta / V^2 (the study's gf_ref), so the correlator's exact
    value is corr_exact = Delta / V^2."""
    fock = FockSpace(["0"])
    H = -mu * fock.n("0")
    ed = EDCore(H, fock, lambda b: fock.occupation(b, "0"))
    grid = ImaginaryTimeGrid(beta, n_tau)
    V = 0.5 * t_bethe
    delta = V ** 2 * bethe_gf(t_bethe / 2.0, beta, grid)
    pairs = [InteractionPair(fock.c_dag("0"), fock.c("0"), delta),
             InteractionPair(fock.c("0"), fock.c_dag("0"), ph_conj(delta))]
    exp = Expansion(ed, grid, pairs, [(fock.c("0"), fock.c_dag("0"))])
    exp.fock = fock
    exp.corr_exact = delta.values / V ** 2
    return exp


def single_state_golden_model(sector_order=("occupied", "empty")):
    """The model of test/inchworm.jl:53-67, :180-212: H = -mu n (mu = 0.1), one bath level
    eps = 0.1 (kd.DeltaDOS) with V = -0.1, beta = 10, n_tau = 20, cubic-spline Delta and
    incremental-spline P."""
    beta, mu, eps, V, n_tau = 10.0, 0.1, 0.1, -0.1, 20
    fock = FockSpace(["0"])
    H = -mu * fock.n("0")
    order = {"empty": 0, "occupied": 1}
    rank = {order[k]: i for i, k in enumerate(sector_order)}
    ed = EDCore(H, fock, lambda b: rank[fock.occupation(b, "0")])
    grid = ImaginaryTimeGrid(beta, n_tau)
    delta = V ** 2 * delta_dos_gf(eps, beta, grid)
    delta_rev = ph_conj(delta)  # -Delta[t2, t1, false] = Delta(beta - tau) for fermions
    pairs = [InteractionPair(fock.c_dag("0"), fock.c("0"), spline_interpolated(delta)),
             InteractionPair(fock.c("0"), fock.c_dag("0"), spline_interpolated(delta_rev))]
    exp = Expansion(ed, grid, pairs, interpolate_ppgf=True)
    exp.fock = fock
    return exp
