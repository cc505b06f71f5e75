"""Synthetic pivots (two-site for the benchmark workloads; one- and zero-site,
and sums of terms, for the tests), and the routine that hands a pivot to a
library through the builder calls of include/vmps_hxv.h.

A pivot is described by PivotData: the two MPO site tensors, the sectors of
the left and right bonds, dense environments and the block layout of the
wavefunction.  Environments are random N(0,1) blocks for every allowed
(in, out, mid) sector; the wavefunction layout has every block allowed by the
symmetry (extend_A materialises all of them, DmrgContractions.h:1832-1867).
"""
import ctypes
from dataclasses import dataclass, field
import math

import numpy as np

from . import models, pack


# ---------------------------------------------------------------------------
def two_site_basis(sym, qloc12, qloc34):
    """Combined basis loc12.combine(loc34) of the reference (Qbasis.h:686-798):
    states sorted by (qmerge, q1, q3), multiplicities of repeated local quantum
    numbers kept adjacent.  Returns list of (qmerge, s1, s3) in s order."""
    states = []
    for s1, q1 in enumerate(qloc12):
        for s3, q3 in enumerate(qloc34):
            for qm in sym.reduce(q1, q3):
                states.append((tuple(qm), tuple(q1), tuple(q3), s1, s3))
    # within equal (qm, q1, q3): plain index p1 + dim1*p3 (first basis fastest)
    states.sort(key=lambda x: (x[0], x[1], x[2], x[4], x[3]))
    return [(st[0], st[3], st[4]) for st in states]


@dataclass
class PivotTerm:
    """An additional summand of a pivot (PivotMatrix2Terms / PivotMatrix1Terms):
    its own MPO tensors and environments on the same bonds."""
    W12: models.MpoSite
    W34: models.MpoSite
    L: list
    R: list


@dataclass
class PivotData:
    sym: models.Sym
    W12: models.MpoSite   # one-site pivots: the site's W; zero-site: None
    W34: models.MpoSite   # two-site pivots only
    left: list          # [(q, dim)] bond sectors left of site 1
    right: list         # [(q, dim)] bond sectors right of site 2
    L: list = field(default_factory=list)   # [(qin, qout, qmid, [F-ordered blocks])]
    R: list = field(default_factory=list)
    V: list = field(default_factory=list)   # per s: [(qin, qout, rows, cols)]
    basis: list = field(default_factory=list)
    nsite: int = 2
    extra: list = field(default_factory=list)  # [PivotTerm] beyond the first

    @property
    def vsize(self):
        return sum(r * c for blocks in self.V for (_, _, r, c) in blocks)

    def block_offsets(self):
        """[(s, qin, qout, offset, rows, cols)] in packed order."""
        out, off = [], 0
        for s, blocks in enumerate(self.V):
            for (qi, qo, r, c) in blocks:
                out.append((s, qi, qo, off, r, c))
                off += r * c
        return out


def _flow_ok(sym, a, b, c):
    return sym.validate3(a, b, c)


def _symmetrize_env(env, aux, conj, flip):
    """Impose E[(qin, qout, qmid)][a] = E[(qout, qin, flip(qmid))][a']^T with a'
    the conjugate auxiliary index, so the effective Hamiltonian is symmetric."""
    off, o = {}, 0
    for q, d in aux:
        off[tuple(q)] = o
        o += d
    glob2sec = {}
    for q, d in aux:
        for a in range(d):
            glob2sec[off[tuple(q)] + a] = (tuple(q), a)
    idx = {(qin, qout, qmid): i for i, (qin, qout, qmid, _) in enumerate(env)}
    done = set()
    for i, (qin, qout, qmid, blocks) in enumerate(env):
        for a in range(len(blocks)):
            if (i, a) in done:
                continue
            qc, ac = glob2sec[conj[off[qmid] + a]]
            j = idx.get((qout, qin, qc))
            if j is None:
                raise ValueError("environment has no conjugate sector")
            X = blocks[a]
            Y = env[j][3][ac]
            S = 0.5 * (X + Y.T)
            blocks[a] = np.asfortranarray(S)
            env[j][3][ac] = np.asfortranarray(S.T.copy())
            done.add((i, a))
            done.add((j, ac))


def random_envs(sym, Win, Wout, left, right, rng, env_scale=1.0, hermitian=False):
    """Random L over the left bond with Win's incoming auxiliary sectors and R
    over the right bond with Wout's outgoing ones, every allowed sector."""
    ldim = dict(left)
    rdim = dict(right)
    L, R = [], []
    # L: {in (bra), out (ket), mid} with out (x) mid -> in  (DmrgContractions.h:30-33)
    for (qmid, adim) in Win.aux_in:
        for (qout, dout) in left:
            for qin in sym.reduce(qout, qmid):
                if qin in ldim:
                    blocks = [np.asfortranarray(rng.standard_normal((ldim[qin], dout)) * env_scale) for _ in range(adim)]
                    L.append((qin, qout, tuple(qmid), blocks))
    # R: {in (ket), out (bra), mid} with in (x) mid -> out
    for (qmid, adim) in Wout.aux_out:
        for (qin, din) in right:
            for qout in sym.reduce(qin, qmid):
                if qout in rdim:
                    blocks = [np.asfortranarray(rng.standard_normal((din, rdim[qout])) * env_scale) for _ in range(adim)]
                    R.append((qin, qout, tuple(qmid), blocks))
    if hermitian:
        flip = lambda q: tuple(-x for x in q)
        _symmetrize_env(L, Win.aux_in, Win.conj_in, flip)
        _symmetrize_env(R, Wout.aux_out, Wout.conj_out, flip)
    return L, R


def make_pivot1(W, left, right, rng, env_scale=1.0, hermitian=False):
    """One-site pivot (PivotMatrix1) on a site with MPO tensor W: PivotVector
    components are the local states s, with every block (qin, qin + qloc[s])."""
    sym = W.sym
    pd = PivotData(sym, W, None, [(tuple(q), d) for q, d in left], [(tuple(q), d) for q, d in right], nsite=1)
    pd.L, pd.R = random_envs(sym, W, W, pd.left, pd.right, rng, env_scale, hermitian)
    rdim = dict(pd.right)
    pd.basis = [(tuple(q), s) for s, q in enumerate(W.qloc)]
    for q in W.qloc:
        blocks = []
        for (qin, din) in pd.left:
            for qout in sym.reduce(qin, q):
                if qout in rdim:
                    blocks.append((qin, tuple(qout), din, rdim[qout]))
        pd.V.append(blocks)
    return pd


def make_pivot0(W, bond, rng, env_scale=1.0):
    """Zero-site pivot (PivotMatrix0) on a bond whose MPO auxiliary sectors are
    W.aux_out: L and R share the mid sectors, the vector is one component of
    blocks (q, q) for every bond sector q."""
    sym = W.sym
    bond = [(tuple(q), d) for q, d in bond]
    pd = PivotData(sym, None, None, bond, bond, nsite=0)
    bdim = dict(bond)
    for (qmid, adim) in W.aux_out:
        for (qout, dout) in bond:
            for qin in sym.reduce(qout, qmid):
                if qin in bdim:
                    pd.L.append((qin, qout, tuple(qmid),
                                 [np.asfortranarray(rng.standard_normal((bdim[qin], dout)) * env_scale)
                                  for _ in range(adim)]))
                    # R sector {in = L.out, out = L.in, mid}: ket x bra
                    pd.R.append((qout, qin, tuple(qmid),
                                 [np.asfortranarray(rng.standard_normal((dout, bdim[qin])) * env_scale)
                                  for _ in range(adim)]))
    pd.basis = [(None, 0)]
    pd.V = [[(q, q, d, d) for (q, d) in bond]]
    return pd


def add_random_term(pd, rng, W12=None, W34=None, env_scale=1.0):
    """Append a summand with fresh random environments (and optionally other
    MPO tensors) on the same bonds."""
    W12 = W12 or pd.W12
    W34 = W34 or pd.W34
    if pd.nsite == 2:
        L, R = random_envs(pd.sym, W12, W34, pd.left, pd.right, rng, env_scale)
    elif pd.nsite == 1:
        L, R = random_envs(pd.sym, W12, W12, pd.left, pd.right, rng, env_scale)
    else:
        raise ValueError("zero-site pivots have a single term")
    pd.extra.append(PivotTerm(W12, W34, L, R))
    return pd


def make_pivot(W12, W34, left, right, rng, env_scale=1.0, hermitian=False):
    """Random-environment pivot for sites with MPO tensors W12, W34.  With
    hermitian=True the environments are symmetrised through the MPO's
    auxiliary conjugation map, so H_eff is a symmetric matrix (needed by the
    eigensolver); this requires left == right-compatible sector sets."""
    sym = W12.sym
    pd = PivotData(sym, W12, W34, [(tuple(q), d) for q, d in left], [(tuple(q), d) for q, d in right])
    rdim = dict(pd.right)
    pd.L, pd.R = random_envs(sym, W12, W34, pd.left, pd.right, rng, env_scale, hermitian)
    pd.basis = two_site_basis(sym, W12.qloc, W34.qloc)
    pd.V = []
    for (qm, s1, s3) in pd.basis:
        blocks = []
        for (qin, din) in pd.left:
            for qout in sym.reduce(qin, qm):
                if qout in rdim:
                    blocks.append((qin, tuple(qout), din, rdim[qout]))
        pd.V.append(blocks)
    return pd


# ---------------------------------------------------------------------------
# bond sectors
# ---------------------------------------------------------------------------
def gaussian_sectors(centers, sigmas, M, parity_fn=None, span=None, max_dim=None, min_dim=1):
    """Sectors q on a grid around `centers` with dims ~ M * gaussian weight.

    centers/sigmas: per component; parity_fn(q) -> bool keeps allowed q;
    span: per component half-width of the grid."""
    nq = len(centers)
    span = span or [int(4 * s) + 2 for s in sigmas]
    grids = [range(int(c) - w, int(c) + w + 1) for c, w in zip(centers, span)]
    qs = [()]
    for g in grids:
        qs = [q + (x,) for q in qs for x in g]
    qs = [q for q in qs if (parity_fn is None or parity_fn(q))]
    w = np.array([math.exp(-sum((q[i] - centers[i]) ** 2 / (2 * sigmas[i] ** 2) for i in range(nq))) for q in qs])
    dims = np.floor(M * w / w.sum()).astype(int)
    keep = dims >= min_dim
    qs = [q for q, k in zip(qs, keep) if k]
    dims = dims[keep]
    w = w[keep]
    if max_dim:
        dims = np.minimum(dims, max_dim)
    # distribute the remainder to the largest-weight sectors
    order = np.argsort(-w)
    i = 0
    while dims.sum() < M:
        j = order[i % len(order)]
        if not max_dim or dims[j] < max_dim:
            dims[j] += 1
        i += 1
        if i > 100 * len(order):
            break
    while dims.sum() > M:
        j = order[-1 - (i % len(order))]
        if dims[j] > 1:
            dims[j] -= 1
        i += 1
    return [(q, int(d)) for q, d in zip(qs, dims) if d > 0]


def heisenberg_bond(nsites_left, M):
    """2Sz sectors of the bond after `nsites_left` spins 1/2, centred on 0."""
    par = nsites_left % 2
    sigma = max(1.0, 0.55 * math.sqrt(nsites_left) * 1.6)
    return gaussian_sectors([0], [sigma], M, parity_fn=lambda q: (q[0] - par) % 2 == 0 and abs(q[0]) <= nsites_left,
                            span=[min(nsites_left, int(5 * sigma) + 2)])


def stretched_sectors(center, M, alpha, lam, allowed):
    """Sectors with dims ~ exp(-(r/lam)^alpha), r = distance from `center`,
    rounded to integers and scaled so the dims sum to M (heavy tail: many
    sectors keep one or two states, a few central ones keep tens)."""
    qs = [q for q in allowed]
    r = np.array([math.sqrt(sum((qi - ci) ** 2 for qi, ci in zip(q, center))) for q in qs])
    w = np.exp(-(r / lam) ** alpha)
    lo, hi = 1e-6, 1e12
    for _ in range(200):
        sc = math.sqrt(lo * hi)
        if np.floor(w * sc + 0.5).sum() > M:
            hi = sc
        else:
            lo = sc
    dims = np.floor(w * lo + 0.5).astype(int)
    order = np.argsort(-w)
    i = 0
    while dims.sum() < M:  # top up the central sectors
        dims[order[i % 4]] += 1
        i += 1
    return [(tuple(q), int(d)) for q, d in zip(qs, dims) if d > 0]


def heisenberg_su2_bond(nsites_left, M, sigma=None):
    """SU(2) multiplets q = 2S+1 of the bond after `nsites_left` spins 1/2:
    S runs over the values the spins allow, the multiplet counts follow a
    Gaussian in S (sum = M multiplets)."""
    par = nsites_left % 2
    qs = [(q,) for q in range(1 + par, nsites_left + 2, 2)]
    sigma = sigma or max(1.0, 0.5 * math.sqrt(nsites_left))
    w = np.array([math.exp(-((q[0] - 1) / 2.0) ** 2 / (2 * sigma ** 2)) for q in qs])
    dims = np.maximum(1, np.floor(M * w / w.sum() + 0.5)).astype(int)
    keep = dims > 0
    return [(q, int(d)) for q, d, k in zip(qs, dims, keep) if k and (M * w[qs.index(q)] / w.sum()) >= 0.5]


def hubbard_su2xu1_bond(nsites_left, M, filling=1.0, sigma_n=None, sigma_s=None):
    """(q = 2S+1, N) multiplet sectors of the bond after `nsites_left` Hubbard
    sites (spin SU(2) x charge U(1)): N around filling * nsites_left, S small,
    2S = N mod 2; multiplet counts Gaussian in (N, S), summing to M."""
    n0 = filling * nsites_left
    sigma_n = sigma_n or max(1.0, 0.45 * math.sqrt(nsites_left))
    sigma_s = sigma_s or max(1.0, 0.35 * math.sqrt(nsites_left))
    qs, w = [], []
    for N in range(max(0, int(n0 - 5 * sigma_n)), min(2 * nsites_left, int(n0 + 5 * sigma_n)) + 1):
        smax = min(N, 2 * nsites_left - N)  # 2 S_max
        for two_s in range(N % 2, smax + 1, 2):
            qs.append((two_s + 1, N))
            w.append(math.exp(-((N - n0) ** 2) / (2 * sigma_n ** 2) - (two_s / 2.0) ** 2 / (2 * sigma_s ** 2)))
    w = np.array(w)
    dims = np.floor(M * w / w.sum() + 0.5).astype(int)
    return [(q, int(d)) for q, d in zip(qs, dims) if d > 0]


def hubbard_bond(nsites_left, M, filling=1.0, alpha=0.36, lam=0.02):
    """(N, 2Sz) sectors of the bond after `nsites_left` Hubbard sites.  With the
    defaults and M = 512: ~110 sectors of dims 1..~55."""
    n0 = round(filling * nsites_left)
    allowed = [(n, s) for n in range(0, 2 * nsites_left + 1) for s in range(-nsites_left, nsites_left + 1)
               if (n - s) % 2 == 0 and abs(s) <= min(n, 2 * nsites_left - n)]
    return stretched_sectors((n0, 0), M, alpha, lam, allowed)


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
WORKLOADS = {
    # two-site pivot in the middle of a HeisenbergU1 chain, L=32, M=64
    "heisenberg_u1_L32_M64": dict(model="HeisenbergU1", L=32, M=64),
    # HubbardU1xU1 chain L=48, half filling, M=512, D_W=6
    "hubbard_u1xu1_L48_M512": dict(model="HubbardU1xU1", L=48, M=512),
    # large-block pivots (MFMA-bound regime)
    "heisenberg_u1_L64_M2048": dict(model="HeisenbergU1", L=64, M=2048),
    "hubbard_u1xu1_L48_M2048": dict(model="HubbardU1xU1", L=48, M=2048, alpha=1.0, lam=2.2),
    # two-leg Hubbard ladder 2 x 32 in snake order, spin SU(2) x charge U(1),
    # t=1, U=6, filling n = 0.875, M = 2048 multiplets, D_W = 6 (rung + leg hopping)
    "hubbard_su2xu1_ladder_L64_M2048": dict(model="HubbardSU2xU1Ladder", L=64, M=2048, filling=0.875),
    # long-range SU(2) Heisenberg, 60 sites, random couplings up to distance
    # 38 (D_W = 40), M = 4096 multiplets: a synthetic large-block pivot with a
    # wider MPO than the C60 Jfull model's (at most 12 multiplet channels,
    # lattice.cut_ranks); smaller variants for parity tests
    "heisenberg_su2_lr40_L60_M4096": dict(model="HeisenbergSU2LR", L=60, M=4096, R=38),
    "heisenberg_su2_lr40_L60_M1024": dict(model="HeisenbergSU2LR", L=60, M=1024, R=38),
}


def lr_couplings(R, rng):
    """Distance couplings J_1 .. J_R of the long-range workloads: J_1 = 1, the
    rest random in [-0.5, 0.5] with a 1/d envelope."""
    return [1.0] + [float(rng.uniform(-0.5, 0.5)) / d for d in range(2, R + 1)]


BATCHES = {
    # 64 independent pivots of HeisenbergU1 (D_W = 5) and HubbardU1xU1 (D_W = 6)
    # chains at random bonds, bond dimension log-uniform in [128, 2048]
    "batch64_u1_M128_2048": dict(n=64, Mmin=128, Mmax=2048),
}


def make_batch(name, seed=20261202):
    """List of PivotData for a batched workload (deterministic in the seed)."""
    cfg = BATCHES[name]
    rng = np.random.default_rng(seed)
    out = []
    heis = models.heisenberg_u1_mpo()
    hub = models.hubbard_u1xu1_mpo()
    for i in range(cfg["n"]):
        M = int(round(math.exp(rng.uniform(math.log(cfg["Mmin"]), math.log(cfg["Mmax"])))))
        if i % 2 == 0:
            nleft = int(rng.integers(12, 52))
            left, right = heisenberg_bond(nleft, M), heisenberg_bond(nleft + 2, M)
            W = heis
        else:
            nleft = int(rng.integers(10, 38))
            left = hubbard_bond(nleft, M, alpha=1.0, lam=2.2)
            right = hubbard_bond(nleft + 2, M, alpha=1.0, lam=2.2)
            W = hub
        out.append(make_pivot(W, W, left, right, rng))
    return out


def make_workload(name, seed=20261202, hermitian=False, **override):
    cfg = dict(WORKLOADS[name])
    cfg.update(override)
    rng = np.random.default_rng(seed)
    L = cfg["L"]
    loc = L // 2 - 1  # pivot sites (loc, loc+1)
    if cfg["model"] == "HeisenbergU1":
        W = models.heisenberg_u1_mpo(J=cfg.get("J", 1.0))
        left = heisenberg_bond(loc, cfg["M"])
        right = heisenberg_bond(loc + 2, cfg["M"])
    elif cfg["model"] == "HubbardU1xU1":
        W = models.hubbard_u1xu1_mpo(t=cfg.get("t", 1.0), U=cfg.get("U", 8.0))
        kw = {k: cfg[k] for k in ("alpha", "lam") if k in cfg}
        left = hubbard_bond(loc, cfg["M"], **kw)
        right = hubbard_bond(loc + 2, cfg["M"], **kw)
    elif cfg["model"] == "HubbardSU2xU1Ladder":
        # sites loc (odd: no rung to its right) and loc + 1 (even: rung)
        t, U = cfg.get("t", 1.0), cfg.get("U", 6.0)
        W12 = models.hubbard_su2xu1_ladder_mpo(t=t, U=U, rung=(loc % 2 == 0))
        W34 = models.hubbard_su2xu1_ladder_mpo(t=t, U=U, rung=((loc + 1) % 2 == 0))
        left = hubbard_su2xu1_bond(loc, cfg["M"], filling=cfg["filling"])
        right = hubbard_su2xu1_bond(loc + 2, cfg["M"], filling=cfg["filling"])
        return make_pivot(W12, W34, left, right, rng, hermitian=hermitian)
    elif cfg["model"] == "HeisenbergSU2LR":
        W = models.heisenberg_su2_lr_mpo(lr_couplings(cfg["R"], rng))
        left = heisenberg_su2_bond(loc, cfg["M"])
        right = heisenberg_su2_bond(loc + 2, cfg["M"])
    else:
        raise ValueError(cfg["model"])
    return make_pivot(W, W, left, right, rng, hermitian=hermitian)


# ---------------------------------------------------------------------------
# handing a pivot to a library
# ---------------------------------------------------------------------------
def _iarr(seq):
    a = np.ascontiguousarray(np.array(seq, dtype=np.int32).ravel())
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _feed_envs(fn, h, L, R, check, keep, nq=None, bulk=False):
    if bulk:
        for which, env in ((0, L), (1, R)):
            if hasattr(env, "handle"):  # env.DeviceEnv: blocks already on the device
                check(fn("pivot2_add_env_device")(h, which, env.handle))
                continue
            pk = pack.tripod(env, nq)
            keep.append(pk)
            check(fn("pivot2_add_env_sectors")(h, which, pk.count, *pk.args()))
        return
    for which, env in ((0, L), (1, R)):
        for (qin, qout, qmid, blocks) in env:
            a, pa = _iarr(qin)
            b, pb = _iarr(qout)
            c, pc = _iarr(qmid)
            rows = np.array([blk.shape[0] if blk is not None else 0 for blk in blocks], dtype=np.int64)
            cols = np.array([blk.shape[1] if blk is not None else 0 for blk in blocks], dtype=np.int64)
            data = np.concatenate([blk.ravel(order="F") for blk in blocks if blk is not None] or [np.zeros(0)])
            data = np.ascontiguousarray(data, dtype=np.float64)
            keep.append((a, b, c, rows, cols, data))
            check(fn("pivot2_add_env")(h, which, pa, pb, pc, len(blocks),
                                       rows.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       cols.ctypes.data_as(ctypes.POINTER(ctypes.c_int64)),
                                       data.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def _feed_W(fn, h, slots, check, nq=None, bulk=False):
    if bulk:
        for which, W in slots:
            pk = pack.w_blocks(W.blocks, nq)
            check(fn("pivot2_add_W_blocks")(h, which, pk.count, *pk.args()))
        return
    for which, W in slots:
        for (s1, s2, k, qi, qo, rows, cols, r, c, v) in W.blocks:
            a, pa = _iarr(qi)
            b, pb = _iarr(qo)
            r = np.ascontiguousarray(r, dtype=np.int32)
            c = np.ascontiguousarray(c, dtype=np.int32)
            v = np.ascontiguousarray(v, dtype=np.float64)
            check(fn("pivot2_add_W")(h, which, s1, s2, k, pa, pb, rows, cols, len(v),
                                     r.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                     v.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))


def _w_slots(pd, W12, W34):
    if pd.nsite == 2:
        return ((12, W12), (34, W34))
    if pd.nsite == 1:
        return ((12, W12),)
    return ()


def feed(lib, prefix, pd, check, projectors=None, Epenalty=0.0, terms=None):
    """Build the pivot in `lib` through the `<prefix>pivot*` builder calls
    (identical signatures for the engine and the oracle).  `terms` selects
    which summands to feed (0 = the first, i >= 1 = pd.extra[i-1]); the
    default is all of them.  Returns the handle."""
    fn = lambda name: getattr(lib, prefix + name)
    bulk = prefix == "vmps_"  # the engine takes whole Tripods / W slots / layouts per call
    sym = pd.sym
    kinds, pk = _iarr(sym.kinds)
    h = ctypes.c_void_p()
    if pd.nsite == 2:
        q12, p12 = _iarr(pd.W12.qloc)
        q34, p34 = _iarr(pd.W34.qloc)
        o12, po12 = _iarr(pd.W12.qop)
        o34, po34 = _iarr(pd.W34.qop)
        check(fn("pivot2_create")(sym.nq, pk, len(pd.W12.qloc), p12, len(pd.W34.qloc), p34, len(pd.W12.qop), po12,
                                  len(pd.W34.qop), po34, ctypes.byref(h)))
    elif pd.nsite == 1:
        q12, p12 = _iarr(pd.W12.qloc)
        o12, po12 = _iarr(pd.W12.qop)
        check(fn("pivot1_create")(sym.nq, pk, len(pd.W12.qloc), p12, len(pd.W12.qop), po12, ctypes.byref(h)))
    else:
        check(fn("pivot0_create")(sym.nq, pk, ctypes.byref(h)))
    keep = []
    all_terms = [PivotTerm(pd.W12, pd.W34, pd.L, pd.R)] + list(pd.extra)
    sel = range(len(all_terms)) if terms is None else list(terms)
    for n, i in enumerate(sel):
        t = all_terms[i]
        if n > 0:
            if prefix != "vmps_":
                raise ValueError("the oracle takes one term per pivot")
            o12, po12 = _iarr(t.W12.qop if t.W12 is not None else [0] * sym.nq)
            o34, po34 = _iarr(t.W34.qop if t.W34 is not None else [0] * sym.nq)
            n12 = len(t.W12.qop) if t.W12 is not None else 0
            n34 = len(t.W34.qop) if t.W34 is not None else 0
            check(fn("pivot_add_term")(h, n12, po12, n34, po34, None))
        _feed_envs(fn, h, t.L, t.R, check, keep, sym.nq, bulk)
        _feed_W(fn, h, _w_slots(pd, t.W12, t.W34), check, sym.nq, bulk)
    if bulk:
        pk = pack.v_blocks(pd.V, sym.nq)
        check(fn("pivot2_add_V_blocks")(h, pk.count, *pk.args(), None))
    off = ctypes.c_int64()
    for s, blocks in (() if bulk else enumerate(pd.V)):
        for (qi, qo, rr, cc) in blocks:
            a, pa = _iarr(qi)
            b, pb = _iarr(qo)
            check(fn("pivot2_add_V_block")(h, s, pa, pb, rr, cc, ctypes.byref(off)))
    if projectors is not None and len(projectors):
        A0 = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float64) for x in projectors]))
        check(fn("pivot2_set_projectors")(h, float(Epenalty), len(projectors),
                                          A0.ctypes.data_as(ctypes.POINTER(ctypes.c_double))))
    return h


def random_vector(pd, seed):
    return np.random.default_rng(seed).standard_normal(pd.vsize)
