"""Exact effective Hamiltonians of short chains (test helper).

With MPS site tensors that keep the full bond bases (selection isometries,
no truncation), the environments contracted from the chain ends turn the
pivot's H_eff into the chain Hamiltonian restricted to the target sector,
written in a non-orthonormal basis.  Applying the same pivot machinery to
the identity MPO gives that basis's overlap operator N, so the generalized
eigenvalues of (H_eff, N_eff) are the exact spectrum of H in the sector.
For SU(2) this pins every coupling coefficient on the path (coeff_HPsi,
coeff_Apair, coeff_tensorProd, coeff_buildL, coeff_buildR) and the reduced
MPO conventions against exact diagonalisation, independently of any
orthonormalisation convention of the reference.

The CPU path uses a restatement of contract_L / contract_R
(tensors/DmrgContractions.h:34-311 with LAWA / AWAR,
tensors/DmrgIndexGymnastics.h:32-124) whose SU(2) factors come from the
oracle's coefficient functions; the device path uses vmps_amd.env.
"""
import ctypes

import numpy as np

from vmps_amd import models, synth


def vacuum(sym):
    return tuple(1 if k == models.SU2_KIND else 0 for k in sym.kinds)


def preimages(sym, qout, ql):
    """q with qout in q (x) ql."""
    comps = []
    for i, k in enumerate(sym.kinds):
        if k == models.U1_KIND:
            comps.append([qout[i] - ql[i]])
        else:
            comps.append([x for x in range(abs(qout[i] - ql[i]) + 1, qout[i] + ql[i], 2)])
    out = [()]
    for c in comps:
        out = [o + (x,) for o in out for x in c]
    return [q for q in out if all(x >= 1 for x, k in zip(q, sym.kinds) if k == models.SU2_KIND)]


# ---------------------------------------------------------------------------
# full bond bases and selection isometries
# ---------------------------------------------------------------------------
def left_bases(sym, sites):
    """B[l] = multiplets (q, mult) of the bond left of site l; B[0] = vacuum."""
    B = [[(vacuum(sym), 1)]]
    for W in sites:
        mult = {}
        for (qin, m) in B[-1]:
            for ql in W.qloc:
                for qo in sym.reduce(qin, ql):
                    mult[qo] = mult.get(qo, 0) + m
        B.append(sorted(mult.items()))
    return B


def right_bases(sym, sites, Qt):
    """C[l] = multiplets of the bond left of site l, for sites l.. coupled to Qt; C[L] = Qt."""
    C = [[(tuple(Qt), 1)]]
    for W in reversed(sites):
        mult = {}
        for (qo, m) in C[0]:
            for ql in W.qloc:
                for qi in preimages(sym, qo, ql):
                    mult[qi] = mult.get(qi, 0) + m
        C.insert(0, sorted(mult.items()))
    return C


def left_isometry(sym, W, bin_, bout):
    """A[s] blocks (qin, qout, M): each (qin multiplet, s) gets its own columns of qout."""
    dout = dict(bout)
    used = {q: 0 for q in dout}
    A = [[] for _ in W.qloc]
    for (qin, m) in bin_:
        for s, ql in enumerate(W.qloc):
            for qo in sym.reduce(qin, ql):
                if qo not in dout:
                    continue
                M = np.zeros((m, dout[qo]))
                M[np.arange(m), used[qo] + np.arange(m)] = 1.0
                used[qo] += m
                A[s].append((qin, qo, np.asfortranarray(M)))
    return A


def right_isometry(sym, W, bin_, bout):
    """A[s] blocks (qin, qout, M): each (qout multiplet, s) gets its own rows of qin."""
    din = dict(bin_)
    used = {q: 0 for q in din}
    A = [[] for _ in W.qloc]
    for (qo, m) in bout:
        for s, ql in enumerate(W.qloc):
            for qi in preimages(sym, qo, ql):
                if qi not in din:
                    continue
                M = np.zeros((din[qi], m))
                M[used[qi] + np.arange(m), np.arange(m)] = 1.0
                used[qi] += m
                A[s].append((qi, qo, np.asfortranarray(M)))
    return A


# ---------------------------------------------------------------------------
# contract_L / contract_R restated for any of the supported symmetries
# ---------------------------------------------------------------------------
class OracleCoeffs:
    def __init__(self, sym):
        import oracle
        self.lib = oracle.load()
        self.sym = sym
        self.kinds = (ctypes.c_int * sym.nq)(*sym.kinds)

    def _call(self, fn, qs):
        flat = [x for q in qs for x in q]
        arr = (ctypes.c_int * len(flat))(*flat)
        return fn(self.sym.nq, self.kinds, arr)

    def buildL(self, *qs):
        return self._call(self.lib.oracle_coeff_buildL, qs)

    def buildR(self, *qs):
        return self._call(self.lib.oracle_coeff_buildR, qs)


def _wblocks(W):
    out = {}
    for (s1, s2, k, qi, qo, rows, cols, r, c, v) in W.blocks:
        out.setdefault((s1, s2, k), {})[(tuple(qi), tuple(qo))] = list(zip(r, c, v))
    return out


def _adict(A):
    return [{(tuple(qi), tuple(qo)): M for (qi, qo, M) in blocks} for blocks in A]


def contract_L_ref(sym, W, Lold, Abra, Aket, coeffs):
    """Lnew[{Rout,Rin,Rmid}][b] += f W(a,b) Abra^T Lold[a] Aket (DmrgContractions.h:34-158)."""
    NA = sym.non_abelian
    wb = _wblocks(W)
    Ab, Ak = _adict(Abra), _adict(Aket)
    dim_out = dict(W.aux_out)
    new = {}
    order = []
    for s1, q1 in enumerate(W.qloc):
        for s2, q2 in enumerate(W.qloc):
            for k, qop in enumerate(W.qop):
                if not sym.validate3(q2, qop, q1):
                    continue
                blocksW = wb.get((s1, s2, k), {})
                for (Lin, Lout, Lmid, blocks) in Lold:
                    for Rout in sym.reduce(Lin, q1):
                        Mb = Ab[s1].get((tuple(Lin), tuple(Rout)))
                        if Mb is None:
                            continue
                        for Rin in sym.reduce(Lout, q2):
                            Mk = Ak[s2].get((tuple(Lout), tuple(Rin)))
                            if Mk is None:
                                continue
                            for Rmid in sym.reduce(Lmid, qop):
                                nz = blocksW.get((tuple(Lmid), tuple(Rmid)))
                                if nz is None or not sym.validate3(Rin, Rmid, Rout):
                                    continue
                                f = coeffs.buildL(Lout, q2, Rin, Lmid, qop, Rmid, Lin, q1, Rout) if NA else 1.0
                                if abs(f) < 1e-14:
                                    continue
                                key = (tuple(Rout), tuple(Rin), tuple(Rmid))
                                for (a, b, v) in nz:
                                    if blocks[a] is None:
                                        continue
                                    M = f * v * (Mb.T @ blocks[a] @ Mk)
                                    if key not in new:
                                        new[key] = [None] * dim_out[tuple(Rmid)]
                                        order.append(key)
                                    new[key][b] = M if new[key][b] is None else new[key][b] + M
    return [(k[0], k[1], k[2], [None if x is None else np.asfortranarray(x) for x in new[k]]) for k in order]


def contract_R_ref(sym, W, Rold, Abra, Aket, coeffs):
    """Rnew[{Lout,Lin,Lmid}][a] += f W(a,b) Aket Rold[b] Abra^T (DmrgContractions.h:179-311)."""
    NA = sym.non_abelian
    wb = _wblocks(W)
    Ab, Ak = _adict(Abra), _adict(Aket)
    dim_in = dict(W.aux_in)
    new = {}
    order = []
    for s1, q1 in enumerate(W.qloc):
        for s2, q2 in enumerate(W.qloc):
            for k, qop in enumerate(W.qop):
                if not sym.validate3(q2, qop, q1):
                    continue
                blocksW = wb.get((s1, s2, k), {})
                for (Rin, Rout, Rmid, blocks) in Rold:
                    for Lin in preimages(sym, Rout, q1):
                        Mb = Ab[s1].get((tuple(Lin), tuple(Rout)))
                        if Mb is None:
                            continue
                        for Lout in preimages(sym, Rin, q2):
                            Mk = Ak[s2].get((tuple(Lout), tuple(Rin)))
                            if Mk is None:
                                continue
                            for Lmid in preimages(sym, Rmid, qop):
                                nz = blocksW.get((tuple(Lmid), tuple(Rmid)))
                                if nz is None or not sym.validate3(Lout, Lmid, Lin):
                                    continue
                                f = coeffs.buildR(Lout, q2, Rin, Lmid, qop, Rmid, Lin, q1, Rout) if NA else 1.0
                                if abs(f) < 1e-14:
                                    continue
                                key = (tuple(Lout), tuple(Lin), tuple(Lmid))
                                for (a, b, v) in nz:
                                    if blocks[b] is None:
                                        continue
                                    M = f * v * (Mk @ blocks[b] @ Mb.T)
                                    if key not in new:
                                        new[key] = [None] * dim_in[tuple(Lmid)]
                                        order.append(key)
                                    new[key][a] = M if new[key][a] is None else new[key][a] + M
    return [(k[0], k[1], k[2], [None if x is None else np.asfortranarray(x) for x in new[k]]) for k in order]


# ---------------------------------------------------------------------------
# exact pivots of a chain
# ---------------------------------------------------------------------------
def chain_pivot(sym, mpo, Qt, p, nsite, contract_L, contract_R):
    """Pivot on sites p (.. p+nsite-1) of the chain `mpo` (list of MpoSite) in
    the sector Qt, with environments contracted from full-basis isometries."""
    L = len(mpo)
    B = left_bases(sym, mpo)
    C = right_bases(sym, mpo, Qt)
    vac = vacuum(sym)
    Lenv = [(vac, vac, tuple(mpo[0].aux_in[0][0]), [np.ones((1, 1))])]
    for l in range(p):
        A = left_isometry(sym, mpo[l], B[l], B[l + 1])
        Lenv = contract_L(mpo[l], Lenv, A, A)
    Renv = [(tuple(Qt), tuple(Qt), tuple(mpo[-1].aux_out[0][0]), [np.ones((1, 1))])]
    for l in range(L - 1, p + nsite - 1, -1):
        A = right_isometry(sym, mpo[l], C[l], C[l + 1])
        Renv = contract_R(mpo[l], Renv, A, A)
    left = [(q, m) for (q, m) in B[p]]
    right = [(q, m) for (q, m) in C[p + nsite]]
    rdim = dict(right)
    if nsite == 0:
        # zero-site pivot on the bond left of site p (PivotMatrix0): the bond
        # matrix C has one block (q, q) per multiplet q both sides carry
        ld = dict(left)
        common = [q for q, _ in left if q in rdim]
        pd = synth.PivotData(sym, None, None, [(q, ld[q]) for q in common], [(q, rdim[q]) for q in common], nsite=0)
        cd = set(common)
        pd.L = [e for e in Lenv if tuple(e[0]) in cd and tuple(e[1]) in cd]
        pd.R = [e for e in Renv if tuple(e[0]) in cd and tuple(e[1]) in cd]
        pd.basis = [(None, 0)]
        pd.V = [[(q, q, ld[q], rdim[q]) for q in common]]
        return pd
    if nsite == 2:
        pd = synth.PivotData(sym, mpo[p], mpo[p + 1], left, right)
        pd.basis = synth.two_site_basis(sym, mpo[p].qloc, mpo[p + 1].qloc)
        comps = [qm for (qm, _, _) in pd.basis]
    else:
        pd = synth.PivotData(sym, mpo[p], None, left, right, nsite=1)
        pd.basis = [(tuple(q), s) for s, q in enumerate(mpo[p].qloc)]
        comps = [tuple(q) for q in mpo[p].qloc]
    # keep only environment sectors on the pivot's bonds
    ld = dict(left)
    pd.L = [e for e in Lenv if tuple(e[0]) in ld and tuple(e[1]) in ld]
    pd.R = [e for e in Renv if tuple(e[0]) in rdim and tuple(e[1]) in rdim]
    for qm in comps:
        blocks = []
        for (qin, din) in left:
            for qout in sym.reduce(qin, qm):
                if qout in rdim:
                    blocks.append((tuple(qin), tuple(qout), din, rdim[qout]))
        pd.V.append(blocks)
    return pd


def operator_matrix(pd, factory):
    """Matrix of the pivot's HxV; factory(pd) returns v -> HxV(v)."""
    apply = factory(pd)
    n = pd.vsize
    return np.stack([apply(np.eye(n)[i]) for i in range(n)], axis=1)


def oracle_factory(pd):
    import oracle
    from pivots import oracle_hxv
    lib = oracle.load()
    h = synth.feed(lib, "oracle_", pd, oracle.check)
    oracle.check(lib.oracle_pivot2_precalc(h))

    def apply(v):
        return oracle_hxv(lib, h, v)
    return apply


def device_factory(pd):
    from vmps_amd import pivot
    H = {2: pivot.PivotMatrix2, 1: pivot.PivotMatrix1, 0: pivot.PivotMatrix0}[pd.nsite](pd)
    return lambda v: pivot.HxV(H, v)


def oracle_contracts(sym):
    co = OracleCoeffs(sym)
    return (lambda W, E, Ab, Ak: contract_L_ref(sym, W, E, Ab, Ak, co),
            lambda W, E, Ab, Ak: contract_R_ref(sym, W, E, Ab, Ak, co))


def device_contracts(sym):
    from vmps_amd import env
    return (lambda W, E, Ab, Ak: env.contract_L(W, E, Ab, Ak),
            lambda W, E, Ab, Ak: env.contract_R(W, E, Ab, Ak))


def generalized_spectrum(Hm, Nm):
    ev = np.linalg.eigvals(np.linalg.solve(Nm, Hm))
    assert np.max(np.abs(ev.imag)) < 1e-8 * max(1.0, np.max(np.abs(ev.real)))
    return np.sort(ev.real)


def exact_spectrum(sym, mpo, Qt, p, nsite, contracts, factory):
    """Generalized eigenvalues of (H_eff, N_eff) of the pivot at p of the chain `mpo`."""
    cL, cR = contracts
    pd_h = chain_pivot(sym, mpo, Qt, p, nsite, cL, cR)
    ids = [models.identity_mpo(sym, W.qloc) for W in mpo]
    pd_n = chain_pivot(sym, ids, Qt, p, nsite, cL, cR)
    assert pd_h.vsize == pd_n.vsize and pd_h.vsize > 0
    return generalized_spectrum(operator_matrix(pd_h, factory), operator_matrix(pd_n, factory))
