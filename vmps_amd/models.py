"""Symmetries, local bases and matrix-product operators of the models the
workloads use.

The reference builds its MPOs generically (MpoTerms.h:1175-1233 lays out
W[loc][s1][s2][k] as Bipeds of sparse matrices keyed by auxiliary sectors).
Here the nearest-neighbour Hamiltonians are written directly as lower
triangular operator-valued matrices; the auxiliary index a of W carries the
charge accumulated to its left, so the block (q_in(a), q_out(b)) of
W[s1][s2][k] holds <s1|O|s2> for an operator O of charge qOp[k] = q_out - q_in
(the flow AAWWAA matches, DmrgIndexGymnastics.h:218-246).

Models:
  HeisenbergU1   S=1/2, H = J sum Sz Sz + J/2 (S+ S- + S- S+)   (models/HeisenbergU1.h)
  HubbardU1xU1   H = -t sum_sigma (c+ c + h.c.) + U sum n_up n_dn (models/HubbardU1xU1.h)
  HeisenbergSU2  S=1/2, H = J sum S.S in SU(2)-reduced form          (models/HeisenbergSU2.h)
  HeisenbergSU2LR  H = sum_d J_d sum_i S_i.S_{i+d}, range R, D_W = R + 2
  HeisenbergSU2Full  H = sum_{i<j} J_ij S_i.S_j for any coupling matrix,
                   D_W = 2 + rank of the couplings across each bond
  HubbardSU2xU1 / HubbardSU2xU1Ladder  chain / snake-ordered two-leg ladder
"""
from dataclasses import dataclass, field
import numpy as np

U1_KIND = 1
SU2_KIND = 2


class Sym:
    """Runtime symmetry: a tuple of U(1)/SU(2) components (symmetry/U1.h, SU2.h, S1xS2.h)."""

    def __init__(self, kinds):
        self.kinds = tuple(int(k) for k in kinds)
        self.nq = len(self.kinds)

    def reduce(self, a, b):
        comps = []
        for i, k in enumerate(self.kinds):
            if k == U1_KIND:
                comps.append([a[i] + b[i]])
            else:
                lo, hi = abs(a[i] - b[i]) + 1, abs(a[i] + b[i]) - 1
                comps.append(list(range(lo, hi + 1, 2)))
        out = [()]
        for c in comps:
            out = [o + (x,) for o in out for x in c]
        return out

    def validate3(self, a, b, c):
        for i, k in enumerate(self.kinds):
            if k == U1_KIND:
                if a[i] + b[i] != c[i]:
                    return False
            else:
                if not (c[i] - 1 >= abs(a[i] - b[i]) and c[i] - 1 <= a[i] + b[i] - 2):
                    return False
        return True

    def add(self, a, b):
        assert all(k == U1_KIND for k in self.kinds)
        return tuple(x + y for x, y in zip(a, b))

    def preimages(self, qout, b):
        """Every q with qout in q (x) b."""
        comps = []
        for i, k in enumerate(self.kinds):
            if k == U1_KIND:
                comps.append([qout[i] - b[i]])
            else:
                comps.append([x for x in range(abs(qout[i] - b[i]) + 1, qout[i] + b[i], 2) if x >= 1])
        out = [()]
        for c in comps:
            out = [o + (x,) for o in out for x in c]
        return out

    def vacuum(self):
        return tuple(1 if k == SU2_KIND else 0 for k in self.kinds)

    def coeff_dot(self, q):
        w = 1.0
        for i, k in enumerate(self.kinds):
            if k == SU2_KIND:
                w *= q[i]
        return w

    @property
    def non_abelian(self):
        return any(k == SU2_KIND for k in self.kinds)


U1 = Sym([U1_KIND])
U1xU1 = Sym([U1_KIND, U1_KIND])


@dataclass
class MpoSite:
    """One site tensor W[s1][s2][k] in block-sparse form.

    aux_in / aux_out: lists of (q, dim) sectors of the left / right auxiliary
    bond; a global auxiliary index enumerates the sectors in list order.
    blocks: list of (s1, s2, k, q_in, q_out, rows, cols, row[], col[], val[]).
    """
    sym: Sym
    qloc: list
    qop: list
    aux_in: list
    aux_out: list
    blocks: list = field(default_factory=list)
    # global auxiliary index -> index of the transposed operator string
    # (W[a, b]^T = W[conj a, conj b]); None when not provided
    conj_in: list = None
    conj_out: list = None

    @property
    def dw_in(self):
        return sum(d for _, d in self.aux_in)

    @property
    def dw_out(self):
        return sum(d for _, d in self.aux_out)

    def dense(self):
        """W as a dense array [a, b, s1, s2] over global auxiliary indices (abelian only)."""
        off_in = _offsets(self.aux_in)
        off_out = _offsets(self.aux_out)
        d = len(self.qloc)
        W = np.zeros((self.dw_in, self.dw_out, d, d))
        for (s1, s2, k, qi, qo, rows, cols, r, c, v) in self.blocks:
            for rr, cc, vv in zip(r, c, v):
                W[off_in[qi] + rr, off_out[qo] + cc, s1, s2] += vv
        return W


def _offsets(sectors):
    off, o = {}, 0
    for q, d in sectors:
        off[tuple(q)] = o
        o += d
    return off


def _aux_sectors(qs):
    """Group global auxiliary indices by charge, in order of first appearance.
    Returns (sectors, local index of each global index)."""
    order, local, count = [], [], {}
    for q in qs:
        q = tuple(q)
        if q not in count:
            count[q] = 0
            order.append(q)
        local.append(count[q])
        count[q] += 1
    # regroup so that sector blocks are contiguous in global numbering
    sectors = [(q, count[q]) for q in order]
    return sectors, local


def build_mpo_site(sym, qloc, aux_q_in, aux_q_out, entries, tol=0.0):
    """Operator-valued matrix -> block-sparse W.

    entries: list of (a, b, op, qop) with op a dense d x d matrix <s1|op|s2>
    of charge qop, placed at global auxiliary position (a, b)."""
    aux_in, loc_in = _aux_sectors(aux_q_in)
    aux_out, loc_out = _aux_sectors(aux_q_out)
    qops = []
    for (_, _, _, qop) in entries:
        if tuple(qop) not in qops:
            qops.append(tuple(qop))
    acc = {}
    for (a, b, op, qop) in entries:
        qa, qb = tuple(aux_q_in[a]), tuple(aux_q_out[b])
        assert sym.validate3(qa, tuple(qop), qb), "operator charge does not match the auxiliary flow"
        k = qops.index(tuple(qop))
        for s1 in range(len(qloc)):
            for s2 in range(len(qloc)):
                v = op[s1, s2]
                if abs(v) <= tol:
                    continue
                assert sym.validate3(tuple(qloc[s2]), tuple(qop), tuple(qloc[s1]))
                key = (s1, s2, k, qa, qb)
                acc.setdefault(key, {})
                acc[key][(loc_in[a], loc_out[b])] = acc[key].get((loc_in[a], loc_out[b]), 0.0) + v
    dim_in = dict(aux_in)
    dim_out = dict(aux_out)
    blocks = []
    for (s1, s2, k, qa, qb), nz in sorted(acc.items()):
        items = sorted(nz.items(), key=lambda kv: (kv[0][1], kv[0][0]))
        r = np.array([i[0][0] for i in items], dtype=np.int32)
        c = np.array([i[0][1] for i in items], dtype=np.int32)
        v = np.array([i[1] for i in items], dtype=np.float64)
        blocks.append((s1, s2, k, qa, qb, dim_in[qa], dim_out[qb], r, c, v))
    return MpoSite(sym, [tuple(q) for q in qloc], qops, aux_in, aux_out, blocks)


def _select(entries, rows, cols, nrow_new=None, ncol_new=None):
    """Restrict an operator-valued matrix to rows/cols (boundary sites)."""
    rmap = {r: i for i, r in enumerate(rows)} if rows is not None else None
    cmap = {c: i for i, c in enumerate(cols)} if cols is not None else None
    out = []
    for (a, b, op, q) in entries:
        if rmap is not None and a not in rmap:
            continue
        if cmap is not None and b not in cmap:
            continue
        out.append((rmap[a] if rmap is not None else a, cmap[b] if cmap is not None else b, op, q))
    return out


# ---------------------------------------------------------------------------
# HeisenbergU1, S = 1/2.  qloc in units of 2Sz: state 0 = down (-1), 1 = up (+1)
# ---------------------------------------------------------------------------
def heisenberg_u1_ops():
    Sp = np.array([[0.0, 0.0], [1.0, 0.0]])  # <s1|S+|s2>: up <- down
    Sm = Sp.T.copy()
    Sz = np.diag([-0.5, 0.5])
    Id = np.eye(2)
    qloc = [(-1,), (1,)]
    return qloc, Id, Sp, Sm, Sz


def heisenberg_u1_mpo(J=1.0, position="bulk"):
    """D_W = 5 lower-triangular MPO.  Global auxiliary order
    0 done | 1 S+ pending (charge -2) | 2 S- pending (+2) | 3 Sz pending | 4 start."""
    qloc, Id, Sp, Sm, Sz = heisenberg_u1_ops()
    q = [(0,), (-2,), (2,), (0,), (0,)]
    ent = [
        (0, 0, Id, (0,)),
        (1, 0, Sp, (2,)),
        (2, 0, Sm, (-2,)),
        (3, 0, Sz, (0,)),
        (4, 1, 0.5 * J * Sm, (-2,)),
        (4, 2, 0.5 * J * Sp, (2,)),
        (4, 3, J * Sz, (0,)),
        (4, 4, Id, (0,)),
    ]
    return _place(U1, qloc, q, ent, position, conj=[0, 2, 1, 3, 4])


# ---------------------------------------------------------------------------
# HubbardU1xU1.  qloc (N, 2Sz): |0>, |up>, |dn>, |updn> = a+_up a+_dn |0>
# ---------------------------------------------------------------------------
def hubbard_ops():
    a_up = np.zeros((4, 4))
    a_dn = np.zeros((4, 4))
    a_up[0, 1] = 1.0   # |up> -> |0>
    a_up[2, 3] = 1.0   # |updn> -> |dn>
    a_dn[0, 2] = 1.0   # |dn> -> |0>
    a_dn[1, 3] = -1.0  # |updn> -> -|up>
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    Id = np.eye(4)
    n_up = a_up.T @ a_up
    n_dn = a_dn.T @ a_dn
    qloc = [(0, 0), (1, 1), (1, -1), (2, 0)]
    return qloc, Id, F, a_up, a_dn, n_up, n_dn


def hubbard_u1xu1_mpo(t=1.0, U=8.0, position="bulk"):
    """D_W = 6 lower-triangular MPO with Jordan-Wigner strings:
    c+_i c_{i+1} = (a+_i F_i) a_{i+1},  c+_{i+1} c_i = (F_i a_i) a+_{i+1}.
    Global auxiliary order: 0 done | 1 (a+F)_up | 2 (F a)_up | 3 (a+F)_dn | 4 (F a)_dn | 5 start."""
    qloc, Id, F, a_up, a_dn, n_up, n_dn = hubbard_ops()
    q = [(0, 0), (1, 1), (-1, -1), (1, -1), (-1, 1), (0, 0)]
    ent = [
        (0, 0, Id, (0, 0)),
        (1, 0, a_up, (-1, -1)),
        (2, 0, a_up.T.copy(), (1, 1)),
        (3, 0, a_dn, (-1, 1)),
        (4, 0, a_dn.T.copy(), (1, -1)),
        (5, 0, U * (n_up @ n_dn), (0, 0)),
        (5, 1, -t * (a_up.T @ F), (1, 1)),
        (5, 2, -t * (F @ a_up), (-1, -1)),
        (5, 3, -t * (a_dn.T @ F), (1, -1)),
        (5, 4, -t * (F @ a_dn), (-1, 1)),
        (5, 5, Id, (0, 0)),
    ]
    return _place(U1xU1, qloc, q, ent, position, conj=[0, 2, 1, 4, 3, 5])


# ---------------------------------------------------------------------------
# HeisenbergSU2, S = 1/2, in reduced (Wigner-Eckart) form: one local
# multiplet q = 2S+1 = 2.  Reduced elements as sites/SpinSiteSU2.h:57-61:
# <S||S||S> = sqrt(S(S+1)); S^dagger = S^T * coeff_adjoint(2, 2, 3) = -S
# (tensors/SiteOperatorQ.h:292-305, symmetry/SU2.h coeff_adjoint).  The
# nearest-neighbour term is sqrt(3) J (S^dagger_i S_{i+1}) coupled to a
# singlet (models/HeisenbergSU2.h:259-262).
# ---------------------------------------------------------------------------
SU2 = Sym([SU2_KIND])
SU2xU1 = Sym([SU2_KIND, U1_KIND])


def heisenberg_su2_mpo(J=1.0, position="bulk", S=0.5):
    q2 = int(round(2 * S)) + 1
    red = np.array([[np.sqrt(S * (S + 1.0))]])
    Id = np.eye(1)
    qloc = [(q2,)]
    # global auxiliary order: 0 done | 1 S pending (q=3) | 2 start
    q = [(1,), (3,), (1,)]
    ent = [
        (0, 0, Id, (1,)),
        (1, 0, red, (3,)),
        (2, 1, np.sqrt(3.0) * J * (-red), (3,)),
        (2, 2, Id, (1,)),
    ]
    return _place(SU2, qloc, q, ent, position)


# ---------------------------------------------------------------------------
# HubbardSU2xU1 (spin SU(2) x charge U(1)) in reduced form.  Local multiplets
# (q = 2S+1, N): empty (1,0), single (2,1), double (1,2)
# (sites/FermionSiteSU2xU1.h:57-79).  Reduced elements: c(empty, single) =
# sqrt(2), c(single, double) = 1; c^dagger = c.adjoint() (coeff_adjoint:
# cdag(single, empty) = 1, cdag(double, single) = -sqrt(2)); F = diag(1,-1,1);
# d = double occupancy.  Products with the scalar F carry coeff_prod = 1.
# Hopping -sqrt(2) t [(c^dagger F)_i c_{i+1} + (c F)_i c^dagger_{i+1}]
# (models/HubbardSU2xU1.h:404-438), on-site U d.
# ---------------------------------------------------------------------------
def hubbard_su2xu1_ops():
    qloc = [(1, 0), (2, 1), (1, 2)]
    c = np.zeros((3, 3))
    c[0, 1] = np.sqrt(2.0)
    c[1, 2] = 1.0
    cdag = np.zeros((3, 3))
    cdag[1, 0] = 1.0
    cdag[2, 1] = -np.sqrt(2.0)
    F = np.diag([1.0, -1.0, 1.0])
    d = np.diag([0.0, 0.0, 1.0])
    return qloc, np.eye(3), F, c, cdag, d


def hubbard_su2xu1_mpo(t=1.0, U=8.0, position="bulk"):
    """D_W = 4 (in multiplets): 0 done | 1 c pending (2,+1) | 2 c^dagger pending (2,-1) | 3 start."""
    qloc, Id, F, c, cdag, d = hubbard_su2xu1_ops()
    q = [(1, 0), (2, 1), (2, -1), (1, 0)]
    s2 = np.sqrt(2.0)
    ent = [
        (0, 0, Id, (1, 0)),
        (1, 0, c, (2, -1)),
        (2, 0, cdag, (2, 1)),
        (3, 0, U * d, (1, 0)),
        (3, 1, -s2 * t * (cdag @ F), (2, 1)),
        (3, 2, -s2 * t * (c @ F), (2, -1)),
        (3, 3, Id, (1, 0)),
    ]
    return _place(SU2xU1, qloc, q, ent, position)


def heisenberg_su2_lr_mpo(Js, position="bulk", S=0.5):
    """H = sum_i sum_{d=1..R} Js[d-1] S_i . S_{i+d} in SU(2)-reduced form
    (the long-range Heisenberg MPO; HeisenbergSU2 with Jfull,
    models/HeisenbergSU2.h:223-229, has this channel structure).  Auxiliary
    order: 0 done | d = 1..R: an S placed d sites ago (q = 3) | R+1 start;
    D_W = R + 2.  The spin-1 channel passes sites with the identity."""
    R = len(Js)
    q2 = int(round(2 * S)) + 1
    red = np.array([[np.sqrt(S * (S + 1.0))]])
    Id = np.eye(1)
    qloc = [(q2,)]
    q = [(1,)] + [(3,)] * R + [(1,)]
    ent = [(0, 0, Id, (1,)), (R + 1, 1, np.sqrt(3.0) * (-red), (3,)), (R + 1, R + 1, Id, (1,))]
    for d in range(1, R + 1):
        if Js[d - 1] != 0.0:
            ent.append((d, 0, Js[d - 1] * red, (3,)))
        if d < R:
            ent.append((d, d + 1, Id, (1,)))
    return _place(SU2, qloc, q, ent, position)


def coupling_channels(J, tol=1e-12):
    """Operator channels of a two-body Hamiltonian sum_{i<j} J_ij O_i . O_j
    on an open chain, at the minimal count per bond.

    At bond b (between sites b-1 and b) the terms that cross it are
    sum_{i<b<=j} J_ij O_i O_j = sum_k Ohat_k P_k with Ohat = Q_b^T O_{0:b},
    where the orthonormal columns of Q_b span the column space of
    C_b = J[0:b, b:L]; their number r_b = rank(C_b) is the least any MPO can
    carry (the reference reaches it by deleting linearly dependent rows and
    columns, MpoTerms.h:1438-1560).  Since the first b-1 rows of C_{b+1} are
    columns of C_b, Q_{b+1}[0:b] = Q_b T_b with T_b = Q_b^T Q_{b+1}[0:b].

    Returns per site s: (Q_{s+1}[s], T_s, close_s) with
      start -> channel k at bond s+1 : Q_{s+1}[s, k]   (the operator placed at s)
      channel k' -> channel k        : T_s[k', k]       (passed on with the identity)
      channel k' -> done             : close_s[k'] = (Q_s^T J[0:s, s])[k']."""
    J = np.asarray(J, dtype=np.float64)
    L = J.shape[0]
    assert J.shape == (L, L) and np.allclose(J, J.T), "J must be a symmetric L x L matrix"
    Q = [np.zeros((0, 0))]
    for b in range(1, L):
        C = J[:b, b:]
        U, s, _ = np.linalg.svd(C, full_matrices=False)
        r = int(np.sum(s > tol * max(1.0, s[0] if len(s) else 0.0)))
        Q.append(U[:, :r])
    Q.append(np.zeros((L, 0)))
    out = []
    for s in range(L):
        Qs, Qn = Q[s], Q[s + 1]
        start = Qn[s, :] if Qn.shape[1] else np.zeros(0)
        T = Qs.T @ Qn[:s, :] if s > 0 else np.zeros((0, Qn.shape[1]))
        close = Qs.T @ J[:s, s] if s > 0 else np.zeros(0)
        out.append((start, T, close))
    return out


def heisenberg_su2_full_mpo(J, S=0.5, tol=1e-14):
    """H = sum_{i<j} J_ij S_i . S_j for an arbitrary coupling matrix (the
    HeisenbergSU2 'Jfull' Hamiltonian, models/HeisenbergSU2.h:223-229) in
    SU(2)-reduced form, one MpoSite per site.  Auxiliary order at bond b:
    0 done | 1..r_b spin-1 channels | r_b + 1 start, with r_b from
    coupling_channels (D_W = r_b + 2 multiplets).  The first site keeps the
    start row only, the last the done column only.  The channel couplings
    mirror heisenberg_su2_lr_mpo: S opens a channel with sqrt(3) (-red),
    closes it with J red."""
    L = J.shape[0]
    q2 = int(round(2 * S)) + 1
    red = np.array([[np.sqrt(S * (S + 1.0))]])
    Id = np.eye(1)
    qloc = [(q2,)]
    ch = coupling_channels(J)
    r = [0] + [len(ch[s][0]) for s in range(L - 1)] + [0]  # channels at bond b

    def aux(b):
        if b == 0:
            return [(1,)]                       # start only
        if b == L:
            return [(1,)]                       # done only
        return [(1,)] + [(3,)] * r[b] + [(1,)]

    sites = []
    for s in range(L):
        start, T, close = ch[s]
        qin, qout = aux(s), aux(s + 1)
        i_start = 0 if s == 0 else r[s] + 1
        o_start = None if s == L - 1 else r[s + 1] + 1
        ent = []
        if s > 0:
            ent.append((0, 0, Id, (1,)))                                   # done -> done
            for k in range(r[s]):
                if abs(close[k]) > tol:
                    ent.append((1 + k, 0, close[k] * red, (3,)))           # close a channel
        if o_start is not None:
            ent.append((i_start, o_start, Id, (1,)))                        # start -> start
            for k in range(r[s + 1]):
                if abs(start[k]) > tol:
                    ent.append((i_start, 1 + k, start[k] * np.sqrt(3.0) * (-red), (3,)))  # open
            for kp in range(r[s]):
                for k in range(r[s + 1]):
                    if abs(T[kp, k]) > tol:
                        ent.append((1 + kp, 1 + k, T[kp, k] * Id, (1,)))   # pass on
        sites.append(build_mpo_site(SU2, qloc, qin, qout, ent))
    return sites


def hubbard_su2xu1_ladder_mpo(t=1.0, U=6.0, rung=True, position="bulk"):
    """Two-leg Hubbard ladder in snake order (site 2x: leg 0, site 2x+1: leg 1
    of column x), spin SU(2) x charge U(1), reduced form (HubbardSU2xU1 with
    tFull, models/HubbardSU2xU1.h:224-233).  Bonds: rungs (2x, 2x+1), i.e.
    distance 1 from even sites (`rung` = this site is even), legs (i, i+2).
    Auxiliary order: 0 done | 1 c pending, distance 1 (2,+1) | 2 c^dagger
    pending, distance 1 (2,-1) | 3, 4 the same at distance 2 | 5 start;
    D_W = 6.  A distance-2 channel crosses the middle site with the string F."""
    qloc, Id, F, c, cdag, d = hubbard_su2xu1_ops()
    q = [(1, 0), (2, 1), (2, -1), (2, 1), (2, -1), (1, 0)]
    s2 = np.sqrt(2.0)
    ent = [
        (0, 0, Id, (1, 0)),
        (1, 0, c, (2, -1)),
        (2, 0, cdag, (2, 1)),
        (3, 1, F, (1, 0)),
        (4, 2, F, (1, 0)),
        (5, 0, U * d, (1, 0)),
        (5, 3, -s2 * t * (cdag @ F), (2, 1)),
        (5, 4, -s2 * t * (c @ F), (2, -1)),
        (5, 5, Id, (1, 0)),
    ]
    if rung:
        ent += [(5, 1, -s2 * t * (cdag @ F), (2, 1)), (5, 2, -s2 * t * (c @ F), (2, -1))]
    return _place(SU2xU1, qloc, q, ent, position)


def identity_mpo(sym, qloc):
    """W = identity with one vacuum auxiliary state (the overlap operator)."""
    vac = tuple(1 if k == SU2_KIND else 0 for k in sym.kinds)
    return build_mpo_site(sym, qloc, [vac], [vac], [(0, 0, np.eye(len(qloc)), vac)])


def _sector_order(qs):
    """Global index -> position in the sector-contiguous numbering of _aux_sectors."""
    sectors, local = _aux_sectors(qs)
    off, o = {}, 0
    for q, d in sectors:
        off[q] = o
        o += d
    return [off[tuple(q)] + l for q, l in zip(qs, local)]


def _place(sym, qloc, q, ent, position, conj=None):
    n = len(q)
    if position == "bulk":
        site = build_mpo_site(sym, qloc, q, q, ent)
        if conj is not None:
            pos = _sector_order(q)
            m = [0] * n
            for g in range(n):
                m[pos[g]] = pos[conj[g]]
            site.conj_in = m
            site.conj_out = m
        return site
    if position == "first":   # row vector: start row only
        e = _select(ent, [n - 1], None)
        return build_mpo_site(sym, qloc, [q[n - 1]], q, e)
    if position == "last":    # column vector: done column only
        e = _select(ent, None, [0])
        return build_mpo_site(sym, qloc, q, [q[0]], e)
    if position == "single":
        e = _select(ent, [n - 1], [0])
        return build_mpo_site(sym, qloc, [q[n - 1]], [q[0]], e)
    raise ValueError(position)


def chain_mpo(model, L, **params):
    """Full open-boundary MPO as a list of MpoSite."""
    if model == "HeisenbergSU2Full":
        J = np.asarray(params["J"], dtype=np.float64)
        assert J.shape == (L, L)
        return heisenberg_su2_full_mpo(J, **{k: v for k, v in params.items() if k != "J"})
    if model in ("HubbardSU2xU1Ladder", "HeisenbergSU2LR"):
        def f(i, position):
            if model == "HubbardSU2xU1Ladder":
                return hubbard_su2xu1_ladder_mpo(rung=(i % 2 == 0), position=position, **params)
            return heisenberg_su2_lr_mpo(position=position, **params)
        if L == 1:
            return [f(0, "single")]
        return [f(0, "first")] + [f(i, "bulk") for i in range(1, L - 1)] + [f(L - 1, "last")]
    f = {"HeisenbergU1": heisenberg_u1_mpo, "HubbardU1xU1": hubbard_u1xu1_mpo,
         "HeisenbergSU2": heisenberg_su2_mpo, "HubbardSU2xU1": hubbard_su2xu1_mpo}[model]
    if L == 1:
        return [f(position="single", **params)]
    return [f(position="first", **params)] + [f(position="bulk", **params) for _ in range(L - 2)] + \
           [f(position="last", **params)]
