"""Two-site DMRG driver for abelian (U(1), U(1)xU(1)) chains: the caller of the
effective-Hamiltonian engine, following DmrgSolver::edgeState / halfsweep /
iteration_two (solvers/DmrgSolver.h:749-874, :1398-1523, :1669-1792):

  per two-site pivot (l, l+1):
    Psi   = contract_AA2(A[l], A[l+1]), all allowed blocks   (DmrgContractions.h:1832-1927)
    E, Psi <- lowest eigenpair of H_eff (L[l], W[l], W[l+1], R[l+2])
    A[l], A[l+1] = split_AA2(Psi): per-sector SVD, keep the largest M singular
                   values, truncation weight = sum of discarded sigma^2
                   (DmrgContractions.h:2460-2630, tensors/Biped.h:741-897)
    L[l+1] = contract_L(L[l], A[l], W[l], A[l])   (left-to-right half sweep)
    R[l+1] = contract_R(R[l+2], A[l+1], W[l+1], A[l+1]) (right-to-left)

The eigensolver and the environment updates are pluggable (a backend); the
default backend runs both on the GPU through the C ABI.  Bond sectors carry
the total charge of the sites to their left; the right boundary carries the
target charge Q.
"""
import math
import time

import numpy as np

from . import models, synth


# ---------------------------------------------------------------------------
# MPS helpers (abelian: every combination of charges is a single irrep)
# ---------------------------------------------------------------------------
def _add(a, b):
    return tuple(x + y for x, y in zip(a, b))


def _sub(a, b):
    return tuple(x - y for x, y in zip(a, b))


def reachable_counts(qloc, nsites, vac):
    """Number of product states with each total charge after nsites sites."""
    counts = {tuple(vac): 1}
    for _ in range(nsites):
        new = {}
        for q, c in counts.items():
            for s in qloc:
                k = _add(q, s)
                new[k] = new.get(k, 0) + c
        counts = new
    return counts


def random_mps(sites, Q, M, rng, spread=2):
    """Random N(0,1) MPS with charge Q, bond dims at most M: each bond keeps the
    sectors reachable from both ends whose charge lies within `spread` of the
    uniform profile l*Q/L, dims split evenly and capped by the number of
    product states on either side."""
    L = len(sites)
    qloc = sites[0].qloc
    vac = tuple(0 for _ in Q)
    bonds = []
    for l in range(L + 1):
        cl = reachable_counts(qloc, l, vac)
        cr = reachable_counts(qloc, L - l, vac)
        centre = tuple(l * x / L for x in Q)
        allowed = {q: min(c, cr[_sub(Q, q)]) for q, c in cl.items()
                   if _sub(Q, q) in cr and all(abs(a - b) <= spread for a, b in zip(q, centre))}
        share = max(1, M // max(1, len(allowed)))
        bonds.append({q: max(1, min(c, share)) for q, c in sorted(allowed.items())})
    A = []
    for l in range(L):
        site = []
        for s in qloc:
            blocks = {}
            for qin, din in bonds[l].items():
                qout = _add(qin, s)
                if qout in bonds[l + 1]:
                    blocks[(qin, qout)] = rng.standard_normal((din, bonds[l + 1][qout]))
            site.append(blocks)
        A.append(site)
    return A, bonds


def right_normalize(A, bonds, qloc):
    """Right-to-left sweep of QR decompositions (Mps::rightSweep): every site
    but the first becomes right-normalized."""
    L = len(A)
    for l in range(L - 1, 0, -1):
        # group by the left charge: matrix [qin] x [(s, qout)]
        newA = [dict() for _ in qloc]
        carry = {}
        for qin, din in bonds[l].items():
            cols, parts = [], []
            for s, qs in enumerate(qloc):
                qout = _add(qin, qs)
                if (qin, qout) in A[l][s]:
                    parts.append((s, qout, A[l][s][(qin, qout)]))
            if not parts:
                continue
            Mm = np.concatenate([p[2] for p in parts], axis=1)
            q, r = np.linalg.qr(Mm.T)  # Mm = r^T q^T
            Bt = q.T  # rows: min(din, cols) orthonormal
            k = Bt.shape[0]
            c0 = 0
            for (s, qout, blk) in parts:
                newA[s][(qin, qout)] = Bt[:, c0:c0 + blk.shape[1]].copy()
                c0 += blk.shape[1]
            carry[qin] = r.T  # din x k
        A[l] = newA
        bonds[l] = {q: carry[q].shape[1] for q in carry}
        # absorb into the left neighbour
        for s in range(len(qloc)):
            for key in list(A[l - 1][s].keys()):
                qin, qout = key
                if qout in carry:
                    A[l - 1][s][key] = A[l - 1][s][key] @ carry[qout]
                else:
                    del A[l - 1][s][key]
    return A, bonds


def two_site_psi(pd, Al, Ar, qloc):
    """contract_AA2: Psi[s][(qin, qout)] = sum_qm A_l[s1][(qin, qm)] A_r[s3][(qm, qout)]
    laid out in the pivot's vector (all allowed blocks, zero where nothing contributes)."""
    v = np.zeros(pd.vsize)
    off = 0
    for s, blocks in enumerate(pd.V):
        _, s1, s3 = pd.basis[s]
        for (qin, qout, r, c) in blocks:
            qm = _add(qin, qloc[s1])
            a = Al[s1].get((qin, qm))
            b = Ar[s3].get((qm, qout))
            if a is not None and b is not None:
                v[off:off + r * c] = (a @ b).ravel(order="F")
            off += r * c
    return v


def split_psi(pd, v, qloc, M, eps_svd=1e-14, left_to_right=True):
    """split_AA2 + truncateSVD for abelian symmetries: SVD per middle charge,
    global selection of the M largest singular values above eps_svd.
    Returns (A_l, A_r, truncation weight, entropy, bond)."""
    # gather the two-site blocks by middle charge qm = qin + q(s1)
    blocks = {}
    off = 0
    for s, bl in enumerate(pd.V):
        _, s1, s3 = pd.basis[s]
        for (qin, qout, r, c) in bl:
            qm = _add(qin, qloc[s1])
            blocks.setdefault(qm, []).append((s1, s3, qin, qout, v[off:off + r * c].reshape((r, c), order="F")))
            off += r * c
    svd = {}
    allsv = []
    for qm, parts in blocks.items():
        rows = sorted({(p[0], p[2]) for p in parts})
        cols = sorted({(p[1], p[3]) for p in parts})
        rdim = {k: next(p[4].shape[0] for p in parts if (p[0], p[2]) == k) for k in rows}
        cdim = {k: next(p[4].shape[1] for p in parts if (p[1], p[3]) == k) for k in cols}
        ro, o = {}, 0
        for k in rows:
            ro[k] = o
            o += rdim[k]
        nr = o
        co, o = {}, 0
        for k in cols:
            co[k] = o
            o += cdim[k]
        nc = o
        Mm = np.zeros((nr, nc))
        for (s1, s3, qin, qout, blk) in parts:
            Mm[ro[(s1, qin)]:ro[(s1, qin)] + blk.shape[0], co[(s3, qout)]:co[(s3, qout)] + blk.shape[1]] = blk
        U, S, Vt = np.linalg.svd(Mm, full_matrices=False)
        svd[qm] = (U, S, Vt, ro, co, rdim, cdim)
        allsv.extend((sv, qm, i) for i, sv in enumerate(S))
    allsv.sort(key=lambda x: -x[0])
    keep = [x for x in allsv[:M] if x[0] > eps_svd]
    kept = {}
    for (sv, qm, i) in keep:
        kept[qm] = max(kept.get(qm, 0), i + 1)
    trunc = float(sum(x[0] ** 2 for x in allsv[len(keep):]))
    p = np.array([x[0] ** 2 for x in keep])
    p = p / p.sum() if p.sum() > 0 else p
    entropy = float(-(p[p > 0] * np.log(p[p > 0])).sum())
    Al = [dict() for _ in qloc]
    Ar = [dict() for _ in qloc]
    for qm, k in kept.items():
        U, S, Vt, ro, co, rdim, cdim = svd[qm]
        U, S, Vt = U[:, :k], S[:k], Vt[:k, :]
        if left_to_right:
            X, Y = U, S[:, None] * Vt
        else:
            X, Y = U * S[None, :], Vt
        for (s1, qin), o in ro.items():
            Al[s1][(qin, qm)] = X[o:o + rdim[(s1, qin)], :].copy()
        for (s3, qout), o in co.items():
            Ar[s3][(qm, qout)] = Y[:, o:o + cdim[(s3, qout)]].copy()
    return Al, Ar, trunc, entropy, kept


def _mps_to_blocks(Asite):
    return [[(qin, qout, np.asfortranarray(b)) for (qin, qout), b in sorted(blocks.items())] for blocks in Asite]


# ---------------------------------------------------------------------------
class GpuBackend:
    """Pivot eigensolver and environment updates on the device."""

    def __init__(self, precision="fp64", refine_fp64=True, tol_eigval=1e-10, tol_state=1e-6, max_iter=100):
        self.precision = precision
        self.refine = refine_fp64
        self.tol_eigval, self.tol_state, self.max_iter = tol_eigval, tol_state, max_iter
        self.hxv_calls = 0

    def solve(self, pd, v0):
        from .pivot import PivotMatrix2, lanczos
        H = PivotMatrix2(pd, precision=self.precision)
        v, r = lanczos(H, v0=v0, precision=self.precision, refine_fp64=self.refine, max_iter=self.max_iter,
                       tol_eigval=self.tol_eigval, tol_state=self.tol_state)
        H.close()
        self.hxv_calls += r["hxv_calls"]
        return r["energy"], v, r

    def contract_L(self, site, Lold, A):
        from . import env
        blocks = _mps_to_blocks(A)
        return env.contract_L(site, Lold, blocks, blocks)

    def contract_R(self, site, Rold, A):
        from . import env
        blocks = _mps_to_blocks(A)
        return env.contract_R(site, Rold, blocks, blocks)


def run_dmrg(model, L, M, sweeps, Q=None, seed=20261202, backend=None, params=None, verbose=False):
    """Ground state by two-site DMRG.  Returns a dict with energies per half
    sweep, the final energy, the truncation weights and timings."""
    params = params or {}
    backend = backend or GpuBackend()
    sites = models.chain_mpo(model, L, **params)
    qloc = sites[0].qloc
    nq = len(qloc[0])
    if Q is None:
        Q = (0,) if model == "HeisenbergU1" else (L, 0)
    Q = tuple(Q)
    rng = np.random.default_rng(seed)
    A, bonds = random_mps(sites, Q, M, rng)
    A, bonds = right_normalize(A, bonds, qloc)
    vac = tuple(0 for _ in range(nq))
    # environments: L[l] sits on bond l, R[l] on bond l
    Lenv = [None] * (L + 1)
    Renv = [None] * (L + 1)
    Lenv[0] = [(vac, vac, tuple(sites[0].aux_in[0][0]), [np.ones((1, 1))])]
    Renv[L] = [(Q, Q, tuple(sites[-1].aux_out[0][0]), [np.ones((1, 1))])]
    for l in range(L - 1, 1, -1):
        Renv[l] = backend.contract_R(sites[l], Renv[l + 1], A[l])
    energies, truncs = [], []
    t_sweeps = []
    E = None
    for sweep in range(sweeps):
        t0 = time.perf_counter()
        for direction in ("right", "left"):
            ls = range(0, L - 1) if direction == "right" else range(L - 2, -1, -1)
            for l in ls:
                pd = synth.PivotData(sites[0].sym, sites[l], sites[l + 1],
                                     sorted(bonds[l].items()), sorted(bonds[l + 2].items()))
                pd.L = Lenv[l]
                pd.R = Renv[l + 2]
                pd.basis = synth.two_site_basis(pd.sym, sites[l].qloc, sites[l + 1].qloc)
                rdim = bonds[l + 2]
                pd.V = []
                for (qm, s1, s3) in pd.basis:
                    bl = []
                    for qin, din in sorted(bonds[l].items()):
                        qout = _add(qin, qm)
                        if qout in rdim:
                            bl.append((qin, qout, din, rdim[qout]))
                    pd.V.append(bl)
                v0 = two_site_psi(pd, A[l], A[l + 1], qloc)
                if not np.any(v0):
                    v0 = rng.uniform(-1, 1, pd.vsize)
                E, v, _ = backend.solve(pd, v0)
                Al, Ar, tw, ent, kept = split_psi(pd, v, qloc, M, left_to_right=(direction == "right"))
                A[l], A[l + 1] = Al, Ar
                bonds[l + 1] = {}
                for qm, k in kept.items():
                    bonds[l + 1][qm] = k
                truncs.append(tw)
                if direction == "right":
                    Lenv[l + 1] = backend.contract_L(sites[l], Lenv[l], A[l])
                else:
                    Renv[l + 1] = backend.contract_R(sites[l + 1], Renv[l + 2], A[l + 1])
            energies.append(E)
            if verbose:
                print("sweep %d %s  E = %.12f  max trunc = %.2e" % (sweep, direction, E, max(truncs[-(L - 1):])))
        t_sweeps.append(time.perf_counter() - t0)
    return {"energy": E, "energies": energies, "truncation": truncs, "sweep_seconds": t_sweeps,
            "bond_dims": [sum(b.values()) for b in bonds]}
