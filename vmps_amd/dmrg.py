"""Two-site DMRG driver for U(1), U(1)xU(1), SU(2) and SU(2)xU(1) chains: the
caller of the effective-Hamiltonian engine, following DmrgSolver::edgeState /
halfsweep / iteration_two (solvers/DmrgSolver.h:749-874, :1398-1523,
:1669-1792):

  per two-site pivot (l, l+1):
    Psi   = contract_AA2(A[l], A[l+1]), all allowed blocks, SU(2) recoupled with
            coeff_twoSiteGate                    (DmrgContractions.h:1832-1927)
    E, Psi <- lowest eigenpair of H_eff (L[l], W[l], W[l+1], R[l+2])
    A[l], A[l+1] = split_AA2(Psi): recoupling with coeff_splitAA, SVD per
                   middle sector, keep the largest M singular values
                   (multiplets), truncation weight = sum of discarded
                   deg(q) sigma^2   (DmrgContractions.h:2460-2630, tensors/Biped.h:741-897)
    L[l+1] = contract_L(L[l], A[l], W[l], A[l])   (left-to-right half sweep)
    R[l+1] = contract_R(R[l+2], A[l+1], W[l+1], A[l+1]) (right-to-left)

The eigensolver and the environment updates are pluggable (a backend); the
default backend runs both on the GPU through the C ABI.  Bond sectors carry
the total quantum numbers of the sites to their left; the right boundary
carries the target Q.  SU(2) bond sectors are multiplets q = 2S+1 and M counts
multiplets.
"""
import ctypes
import math
import os
import time

import numpy as np

from . import _lib, models, synth


# ---------------------------------------------------------------------------
# SU(2) recoupling factors of contract_AA2 / split_AA2 (symmetry/SU2.h; the
# U(1) components contribute 1).  phase() follows symmetry/functions.h:50-55
# with C++ integer division.
# ---------------------------------------------------------------------------
def _cdiv2(x):
    return int(x / 2)


def _phase(x):
    return -1.0 if x % 2 else 1.0


def _6j(q1, q2, q3, q4, q5, q6):
    return _lib.load().vmps_coupling_6j(q1, q2, q3, q4, q5, q6)


def coeff_two_site_gate(sym, q1, q2, q3, q4, q5, q6):
    """coeff_twoSiteGate: 6j{q1 q2 q3; q4 q5 q6} sqrt(q3 q6) (-1)^((q1+q2+q4+q5-4)/2)."""
    out = 1.0
    for i, k in enumerate(sym.kinds):
        if k == models.SU2_KIND:
            out *= _6j(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i]) * math.sqrt(q3[i] * q6[i]) * \
                _phase(_cdiv2(q1[i] + q2[i] + q4[i] + q5[i] - 4))
    return out


def coeff_split6(sym, q1, q2, q3, q4, q5, q6):
    """coeff_splitAA (six arguments): 6j{q1 q2 q3; q4 q5 q6} sqrt(q2 q3) (-1)^((q1+q5+q6-3)/2)."""
    out = 1.0
    for i, k in enumerate(sym.kinds):
        if k == models.SU2_KIND:
            out *= _6j(q1[i], q2[i], q3[i], q4[i], q5[i], q6[i]) * math.sqrt(q2[i] * q3[i]) * \
                _phase(_cdiv2(q1[i] + q5[i] + q6[i] - 3))
    return out


def coeff_split3(sym, q1, q2, q3):
    """coeff_splitAA (three arguments): (-1)^((q1-q2-q3-3)/2) sqrt(q1/q2)."""
    out = 1.0
    for i, k in enumerate(sym.kinds):
        if k == models.SU2_KIND:
            out *= _phase(_cdiv2(q1[i] - q2[i] - q3[i] - 3)) * math.sqrt(q1[i] / q2[i])
    return out


# ---------------------------------------------------------------------------
# MPS helpers
# ---------------------------------------------------------------------------
def left_counts(sym, qloc, nsites):
    """Number of multiplets (SU(2)) or states (U(1)) of each quantum number
    after nsites sites, coupled from the vacuum."""
    counts = {sym.vacuum(): 1}
    for _ in range(nsites):
        new = {}
        for q, c in counts.items():
            for s in qloc:
                for k in sym.reduce(q, s):
                    new[k] = new.get(k, 0) + c
        counts = new
    return counts


def right_counts(sym, qloc, nsites, Q):
    """Same for the bond left of the last nsites sites, coupled to Q at the end."""
    counts = {tuple(Q): 1}
    for _ in range(nsites):
        new = {}
        for q, c in counts.items():
            for s in qloc:
                for k in sym.preimages(q, s):
                    new[k] = new.get(k, 0) + c
        counts = new
    return counts


def random_mps(sites, Q, M, rng, spread=2, max_spin2=6):
    """Random N(0,1) MPS with total quantum number Q, bond dims at most M:
    each bond keeps the sectors reachable from both ends whose U(1) charges
    lie within `spread` of the uniform profile l*Q/L (SU(2) components up to
    2S = max_spin2), dims split evenly and capped by the counts on either side."""
    L = len(sites)
    sym = sites[0].sym
    qloc = sites[0].qloc
    bonds = []
    for l in range(L + 1):
        cl = left_counts(sym, qloc, l)
        cr = right_counts(sym, qloc, L - l, Q)
        allowed = {}
        for q, c in cl.items():
            if q not in cr:
                continue
            ok = True
            for i, k in enumerate(sym.kinds):
                if k == models.U1_KIND and abs(q[i] - l * Q[i] / L) > spread:
                    ok = False
                if k == models.SU2_KIND and q[i] - 1 > max_spin2:
                    ok = False
            if ok:
                allowed[q] = min(c, cr[q])
        share = max(1, M // max(1, len(allowed)))
        bonds.append({q: max(1, min(c, share)) for q, c in sorted(allowed.items())})
    A = []
    for l in range(L):
        site = []
        for s in qloc:
            blocks = {}
            for qin, din in bonds[l].items():
                for qout in sym.reduce(qin, s):
                    if qout in bonds[l + 1]:
                        blocks[(qin, qout)] = rng.standard_normal((din, bonds[l + 1][qout]))
            site.append(blocks)
        A.append(site)
    return A, bonds


def su2_gaussian_bonds(sym, qloc, L, Q, M, width=0.5):
    """Bond sectors of a chain with at most M states (multiplets for SU(2))
    per bond: every sector reachable from both ends, with counts ~
    exp(-sum_i (q_i - c_i)^2 / (2 sigma^2)), sigma = width * sqrt(min(b, L - b)),
    around c_i = 0 in the spin S of SU(2) components and b Q_i / L in U(1)
    charges, each capped by the exact count on either side of the bond (so
    the bonds near the ends keep the full space)."""
    bonds = []
    lefts, rights = [{sym.vacuum(): 1}], [{tuple(Q): 1}]
    for _ in range(L):  # exact counts after 1..L sites from either end
        new = {}
        for q, c in lefts[-1].items():
            for sq in qloc:
                for k in sym.reduce(q, sq):
                    new[k] = new.get(k, 0) + c
        lefts.append(new)
        new = {}
        for q, c in rights[-1].items():
            for sq in qloc:
                for k in sym.preimages(q, sq):
                    new[k] = new.get(k, 0) + c
        rights.append(new)
    for b in range(L + 1):
        cl, cr = lefts[b], rights[L - b]
        caps = {q: min(c, cr[q]) for q, c in cl.items() if q in cr}
        if sum(caps.values()) <= M:
            bonds.append(dict(sorted(caps.items())))
            continue
        sigma = max(1.0, width * math.sqrt(min(b, L - b)))
        qs = sorted(caps)

        def dist2(q):
            d = 0.0
            for i, k in enumerate(sym.kinds):
                x = (q[i] - 1) / 2.0 if k == models.SU2_KIND else q[i] - b * Q[i] / L
                d += x * x
            return d
        w = np.array([math.exp(-dist2(q) / (2 * sigma ** 2)) for q in qs])
        cap = np.array([caps[q] for q in qs], dtype=np.float64)
        lo, hi = 1e-300, 1e300
        for _ in range(200):  # largest scale whose rounded, capped counts fit in M (geometric bisection)
            mid = math.sqrt(lo * hi)
            if np.minimum(cap, np.floor(mid * w + 0.5)).sum() <= M:
                lo = mid
            else:
                hi = mid
        dims = np.minimum(cap, np.floor(lo * w + 0.5)).astype(int)
        bonds.append({q: int(d) for q, d in zip(qs, dims) if d > 0})
    return bonds


def random_blocks(sym, qloc, bonds, rng):
    """N(0,1) site tensors for every block the bond sectors allow."""
    A = []
    for l in range(len(bonds) - 1):
        site = []
        for s in qloc:
            blocks = {}
            for qin, din in bonds[l].items():
                for qout in sym.reduce(qin, s):
                    if qout in bonds[l + 1]:
                        blocks[(qin, qout)] = rng.standard_normal((din, bonds[l + 1][qout]))
            site.append(blocks)
        A.append(site)
    return A


def left_normalize(A, bonds, qloc, sym, stop):
    """Left-to-right sweep of QR decompositions (Mps::leftSweep) over the
    sites l < stop: each becomes left-normalized,
    sum_{s, qin} A[s](qin,qout)^T A[s](qin,qout) = 1 (the normalisation the
    left factor U of split_psi has), and its R factor moves into site l+1."""
    for l in range(stop):
        newA = [dict() for _ in qloc]
        carry = {}
        for qout, dout in bonds[l + 1].items():
            parts = [(s, qin, A[l][s][(qin, qout)]) for s in range(len(qloc)) for qin in bonds[l]
                     if (qin, qout) in A[l][s]]
            if not parts:
                continue
            Mm = np.concatenate([p[2] for p in parts], axis=0)
            q, r = np.linalg.qr(Mm)
            r0 = 0
            for (s, qin, blk) in parts:
                newA[s][(qin, qout)] = q[r0:r0 + blk.shape[0], :]
                r0 += blk.shape[0]
            carry[qout] = r
        A[l] = newA
        bonds[l + 1] = {q: carry[q].shape[0] for q in carry}
        for s in range(len(qloc)):
            for key in list(A[l + 1][s].keys()):
                qin, qout = key
                if qin in carry:
                    A[l + 1][s][key] = carry[qin] @ A[l + 1][s][key]
                else:
                    del A[l + 1][s][key]
    return A, bonds


def right_normalize(A, bonds, qloc, sym=None, stop=1):
    """Right-to-left sweep of QR decompositions (Mps::rightSweep): every site
    l >= stop becomes right-normalized,
    sum_{s, qout} (deg qout / deg qin) A[s](qin,qout) A[s](qin,qout)^T = 1
    (deg = coeff_dot; the weight is 1 for U(1)), and its R factor moves into
    site l-1."""
    L = len(A)
    sym = sym or models.Sym([models.U1_KIND] * len(qloc[0]))
    for l in range(L - 1, stop - 1, -1):
        newA = [dict() for _ in qloc]
        carry = {}
        for qin, din in bonds[l].items():
            parts = []
            for s, qs in enumerate(qloc):
                for qout in sym.reduce(qin, qs):
                    if (qin, qout) in A[l][s]:
                        w = math.sqrt(sym.coeff_dot(qout) / sym.coeff_dot(qin))
                        parts.append((s, qout, w, A[l][s][(qin, qout)]))
            if not parts:
                continue
            Mm = np.concatenate([p[2] * p[3] for p in parts], axis=1)
            q, r = np.linalg.qr(Mm.T)  # Mm = r^T q^T
            Bt = q.T  # orthonormal rows
            c0 = 0
            for (s, qout, w, blk) in parts:
                newA[s][(qin, qout)] = Bt[:, c0:c0 + blk.shape[1]] / w
                c0 += blk.shape[1]
            carry[qin] = r.T  # din x k
        A[l] = newA
        bonds[l] = {q: carry[q].shape[1] for q in carry}
        for s in range(len(qloc)):
            for key in list(A[l - 1][s].keys()):
                qin, qout = key
                if qout in carry:
                    A[l - 1][s][key] = A[l - 1][s][key] @ carry[qout]
                else:
                    del A[l - 1][s][key]
    return A, bonds


def two_site_psi(pd, Al, Ar, qloc):
    """contract_AA2: Psi[s1s3](qin, qout) = sum_qm coeff_twoSiteGate(qin, q1, qm,
    q3, qout, qmerge) A_l[s1](qin, qm) A_r[s3](qm, qout), laid out in the
    pivot's vector (all allowed blocks, zero where nothing contributes)."""
    sym = pd.sym
    v = np.zeros(pd.vsize)
    off = 0
    for s, blocks in enumerate(pd.V):
        qmerge, s1, s3 = pd.basis[s]
        for (qin, qout, r, c) in blocks:
            # the block's column-major slot, written through its transpose
            # (row-major c x r): (A_l A_r)^T = A_r^T A_l^T lands in place
            dst_t = v[off:off + r * c].reshape((c, r))
            first = True
            for qm in sym.reduce(qin, qloc[s1]):
                a = Al[s1].get((qin, qm))
                b = Ar[s3].get((qm, qout))
                if a is None or b is None:
                    continue
                f = coeff_two_site_gate(sym, qin, qloc[s1], qm, qloc[s3], qout, qmerge)
                if abs(f) < 1e-14:
                    continue
                if first:
                    np.matmul(b.T, a.T, out=dst_t)
                    dst_t *= f
                    first = False
                else:
                    t = np.matmul(b.T, a.T)
                    t *= f
                    dst_t += t
            off += r * c
    return v


def two_site_psi_terms(pd, Al, Ar, qloc):
    """The products of contract_AA2 (two_site_psi) in the pivot's layout,
    for the device build (pivot.contract_aa2): the A_l and A_r blocks they
    use, and per product the offset of its output block, the indices of its
    two blocks and its coeff_twoSiteGate factor."""
    sym = pd.sym
    idx1, idx2, blocks1, blocks2 = {}, {}, [], []
    out_off, i1, i2, factor = [], [], [], []
    off = 0
    for s, blocks in enumerate(pd.V):
        qmerge, s1, s3 = pd.basis[s]
        for (qin, qout, r, c) in blocks:
            for qm in sym.reduce(qin, qloc[s1]):
                a = Al[s1].get((qin, qm))
                b = Ar[s3].get((qm, qout))
                if a is None or b is None:
                    continue
                f = coeff_two_site_gate(sym, qin, qloc[s1], qm, qloc[s3], qout, qmerge)
                if abs(f) < 1e-14:
                    continue
                k1, k2 = (s1, qin, qm), (s3, qm, qout)
                if k1 not in idx1:
                    idx1[k1] = len(blocks1)
                    blocks1.append(a)
                if k2 not in idx2:
                    idx2[k2] = len(blocks2)
                    blocks2.append(b)
                out_off.append(off)
                i1.append(idx1[k1])
                i2.append(idx2[k2])
                factor.append(f)
            off += r * c
    return blocks1, blocks2, out_off, i1, i2, factor


def block_svd_device(mats):
    """Thin SVD of every matrix of `mats` in one device call (vmps_block_svd,
    one-sided Jacobi in FP64): list of (U, S, Vt) with S descending."""
    lib = _lib.load()
    if not mats:
        return []
    rows = np.array([m.shape[0] for m in mats], dtype=np.int64)
    cols = np.array([m.shape[1] for m in mats], dtype=np.int64)
    r = np.minimum(rows, cols)
    A = np.concatenate([np.asarray(m, dtype=np.float64).ravel(order="F") for m in mats] + [np.zeros(1)])
    U = np.empty(int((rows * r).sum()) + 1)
    S = np.empty(int(r.sum()) + 1)
    Vt = np.empty(int((r * cols).sum()) + 1)
    P = lambda a, t=ctypes.c_double: a.ctypes.data_as(ctypes.POINTER(t))
    _lib.check(lib.vmps_block_svd(len(mats), P(rows, ctypes.c_int64), P(cols, ctypes.c_int64), P(A), P(U), P(S),
                                  P(Vt), None))
    out, uo, so, vo = [], 0, 0, 0
    for m_, n_, k in zip(rows, cols, r):
        m_, n_, k = int(m_), int(n_), int(k)
        out.append((U[uo:uo + m_ * k].reshape((m_, k), order="F"), S[so:so + k].copy(),
                    Vt[vo:vo + k * n_].reshape((k, n_), order="F")))
        uo, so, vo = uo + m_ * k, so + k, vo + k * n_
    return out


_HOST_SVD = {}


def _host_workers():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def svd_robust(m, full_matrices=False):
    """Thin SVD by LAPACK's divide and conquer (gesdd, numpy), falling back to
    the QR-iteration driver (gesvd, scipy) when gesdd reports that it did not
    converge, which it can for some matrices that gesvd handles (the
    reference picks Eigen's BDCSVD, or JacobiSVD with DONT_USE_BDCSVD, at
    compile time, tensors/Biped.h:754-759).  Non-finite input is refused."""
    if not np.all(np.isfinite(m)):
        raise FloatingPointError("split: non-finite entries in the two-site wavefunction")
    try:
        return np.linalg.svd(m, full_matrices=full_matrices)
    except np.linalg.LinAlgError:
        import scipy.linalg
        return scipy.linalg.svd(m, full_matrices=full_matrices, lapack_driver="gesvd")


def block_svd_host(mats):
    """Thin SVD of every matrix with host LAPACK (svd_robust).  When the
    work is large enough, the blocks go to a thread pool, largest first,
    with one BLAS thread per call (LAPACK releases the GIL; nested BLAS
    threads would oversubscribe the cores).  The result of each block is the
    same as a single-threaded call's."""
    cost = [m.shape[0] * m.shape[1] * min(m.shape) for m in mats]
    if len(mats) < 2 or sum(cost) < 4_000_000:
        return [svd_robust(m) for m in mats]
    if not _HOST_SVD:
        from concurrent.futures import ThreadPoolExecutor
        from threadpoolctl import ThreadpoolController
        _HOST_SVD["pool"] = ThreadPoolExecutor(_host_workers())
        _HOST_SVD["blas"] = ThreadpoolController()
    pool = _HOST_SVD["pool"]
    with _HOST_SVD["blas"].limit(limits=1, user_api="blas"):
        futs = {i: pool.submit(svd_robust, mats[i], False) for i in sorted(range(len(mats)), key=lambda i: -cost[i])}
        return [futs[i].result() for i in range(len(mats))]


def _packed_base(mats):
    """The buffer that holds `mats` back to back in column-major order (as
    split_psi lays them out, with one spare element at the end), or None."""
    base = mats[0].base
    if base is None or base.ndim != 1 or base.dtype != np.float64 or not base.flags.c_contiguous:
        return None
    ptr = base.ctypes.data
    off = 0
    for m in mats:
        if m.base is not base or not m.flags.f_contiguous or m.ctypes.data != ptr + 8 * off:
            return None
        off += m.size
    return base if base.size == off + 1 else None


def block_eigh_device(mats, side):
    """Density-matrix factorisation of every matrix of `mats` on the device
    (vmps_block_eigh_split): side 0 gives (lambda, W, C) with
    Psi Psi^T = W diag(lambda) W^T and C = W^T Psi, side 1 the same with
    Psi^T Psi and C = Psi W; lambda descending (= sigma^2)."""
    lib = _lib.load()
    if not mats:
        return []
    rows = np.array([m.shape[0] for m in mats], dtype=np.int64)
    cols = np.array([m.shape[1] for m in mats], dtype=np.int64)
    k = rows if side == 0 else cols
    A = _packed_base(mats)
    if A is None:
        A = np.concatenate([np.asarray(m, dtype=np.float64).ravel(order="F") for m in mats] + [np.zeros(1)])
    lam = np.empty(int(k.sum()) + 1)
    W = np.empty(int((k * k).sum()) + 1)
    C = np.empty(A.size)
    P = lambda a, t=ctypes.c_double: a.ctypes.data_as(ctypes.POINTER(t))
    _lib.check(lib.vmps_block_eigh_split(len(mats), P(rows, ctypes.c_int64), P(cols, ctypes.c_int64), P(A), side,
                                         P(lam), P(W), P(C)))
    out, lo, wo, co = [], 0, 0, 0
    for r, c, kk in zip(rows, cols, k):
        r, c, kk = int(r), int(c), int(kk)
        l_ = lam[lo:lo + kk][::-1].copy()
        w_ = W[wo:wo + kk * kk].reshape((kk, kk), order="F")[:, ::-1]
        c_ = C[co:co + r * c].reshape((r, c), order="F")
        c_ = c_[::-1, :] if side == 0 else c_[:, ::-1]
        out.append((l_, w_, c_))  # views in descending order of lambda
        lo, wo, co = lo + kk, wo + kk * kk, co + r * c
    return out


class DevicePsi:
    """A two-site wavefunction left on the device by the eigensolver (a plan
    vector): the split reads it there (vmps_block_eigh_split_gather) and
    copies it to the host only when the factorisation runs on the host."""

    def __init__(self, H, dv):
        self.H, self.dv = H, dv
        self.size = H.vsize

    def download(self):
        return self.dv.download()

    def close(self):
        if self.dv is not None:
            self.dv.free()
            self.dv = None
        if self.H is not None:
            self.H.close()
            self.H = None


def block_eigh_gather_device(dev, shapes, dst, side):
    """block_eigh_device on sector matrices built on the device from the
    blocks of a DevicePsi: shapes = [(rows, cols)] per sector, dst = list of
    (sector, row0, col0, r, c, [(src offset, factor), ...]); same return
    format as block_eigh_device."""
    lib = _lib.load()
    rows = np.array([r for r, _ in shapes] or [0], dtype=np.int64)
    cols = np.array([c for _, c in shapes] or [0], dtype=np.int64)
    db = np.array([d[0] for d in dst] or [0], dtype=np.int32)
    dr0 = np.array([d[1] for d in dst] or [0], dtype=np.int64)
    dc0 = np.array([d[2] for d in dst] or [0], dtype=np.int64)
    dr = np.array([d[3] for d in dst] or [0], dtype=np.int64)
    dc = np.array([d[4] for d in dst] or [0], dtype=np.int64)
    pb = np.zeros(len(dst) + 1, dtype=np.int64)
    pb[1:] = np.cumsum([len(d[5]) for d in dst])
    src = np.array([o for d in dst for (o, _) in d[5]] or [0], dtype=np.int64)
    fac = np.array([f for d in dst for (_, f) in d[5]] or [0.0], dtype=np.float64)
    k = rows if side == 0 else cols
    k = k[:len(shapes)]
    lam = np.empty(int(k.sum()) + 1)
    W = np.empty(int((k * k).sum()) + 1)
    C = np.empty(int((rows[:len(shapes)] * cols[:len(shapes)]).sum()) + 1)
    P = lambda a, t=ctypes.c_double: a.ctypes.data_as(ctypes.POINTER(t))
    Pi = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    _lib.check(lib.vmps_block_eigh_split_gather(
        dev.dv.ptr, dev.size, len(shapes), P(rows, ctypes.c_int64), P(cols, ctypes.c_int64), len(dst), Pi(db),
        P(dr0, ctypes.c_int64), P(dc0, ctypes.c_int64), P(dr, ctypes.c_int64), P(dc, ctypes.c_int64),
        P(pb, ctypes.c_int64), P(src, ctypes.c_int64), P(fac), side, P(lam), P(W), P(C)))
    out, lo, wo, co = [], 0, 0, 0
    for (r, c), kk in zip(shapes, k):
        kk = int(kk)
        l_ = lam[lo:lo + kk][::-1].copy()
        w_ = W[wo:wo + kk * kk].reshape((kk, kk), order="F")[:, ::-1]
        c_ = C[co:co + r * c].reshape((r, c), order="F")
        c_ = c_[::-1, :] if side == 0 else c_[:, ::-1]
        out.append((l_, w_, c_))
        lo, wo, co = lo + kk, wo + kk * kk, co + r * c
    return out


def split_psi(pd, v, qloc, M, eps_svd=1e-14, left_to_right=True, svd=None, eigh=None, eps_rho=1e-15,
              timings=None, eigh_gather=None):
    """split_AA2 + truncateSVD: recouple the two-site blocks to the middle
    sector Qm (coeff_splitAA), factor each Qm, keep the M largest singular
    values (multiplets) above eps_svd.  Returns (A_l, A_r, truncation weight =
    sum of discarded deg(Qm) sigma^2, entropy, kept dims per Qm).

    `svd` maps a list of matrices to their thin SVDs (default: host LAPACK).
    With `eigh` (block_eigh_device, or a function that returns None to leave
    the matrices to `svd`) the factorisation is that of the reduced
    density matrix of the side that becomes orthonormal: sigma^2 are its
    eigenvalues, the new site tensor its eigenvectors and the centre W^T Psi
    (or Psi W), so no singular value is divided by; sigma^2 below eps_rho
    (the eigenvalues' resolution relative to ||Psi||^2) count as discarded."""
    sym = pd.sym
    t0 = time.perf_counter()
    dev = v if isinstance(v, DevicePsi) else None
    if dev is None and not np.all(np.isfinite(v)):
        raise FloatingPointError("split: the two-site wavefunction has non-finite entries")
    # first pass: which blocks go to which middle sector Qm, with their factors
    parts = {}  # Qm -> [(row key, col key, offset of the block in v, r, c, factor)]
    rdims, cdims = {}, {}
    off = 0
    for s, bl in enumerate(pd.V):
        q12, sl, sr = pd.basis[s]
        for (qin, qout, r, c) in bl:
            for Qm in sym.reduce(qin, qloc[sl]):
                if qout not in sym.reduce(Qm, qloc[sr]):
                    continue
                f = coeff_split6(sym, qin, qout, q12, qloc[sr], qloc[sl], Qm)
                if abs(f) < 1e-14:
                    continue
                parts.setdefault(Qm, []).append(((sl, qin), (sr, qout), off, r, c, f))
                rdims.setdefault(Qm, {})[(sl, qin)] = r
                cdims.setdefault(Qm, {})[(sr, qout)] = c
            off += r * c
    # second pass: every sector matrix written in place, column-major, into one
    # packed buffer (the layout the device factorisation reads), no temporaries
    meta, shapes, total = [], [], 0
    for Qm in parts:
        ro, o = {}, 0
        for k in sorted(rdims[Qm]):
            ro[k] = o
            o += rdims[Qm][k]
        nr = o
        co, o = {}, 0
        for k in sorted(cdims[Qm]):
            co[k] = o
            o += cdims[Qm][k]
        meta.append((Qm, ro, co))
        shapes.append((nr, o, total))
        total += nr * o
    efac = None
    if dev is not None:
        # with the wavefunction on the device, the sector matrices are built
        # there (or, when the factorisation declines the device, on the host
        # from a copy)
        if eigh_gather is not None:
            dst = []
            for b, (Qm, ro, co) in enumerate(meta):
                groups = {}
                for (rk, ck, o0, r, c, f) in parts[Qm]:
                    groups.setdefault((rk, ck), (r, c, []))[2].append((o0, f))
                for (rk, ck), (r, c, lst) in groups.items():
                    dst.append((b, ro[rk], co[ck], r, c, lst))
            t1 = time.perf_counter()
            efac = eigh_gather(dev, [(nr, nc) for (nr, nc, _) in shapes], dst, 0 if left_to_right else 1)
        if efac is None:
            v = dev.download()
            if not np.all(np.isfinite(v)):
                raise FloatingPointError("split: the two-site wavefunction has non-finite entries")
    if efac is not None:
        norm2 = float(sum(np.sum(lam) for (lam, _, _) in efac))  # trace of rho = ||Psi||^2
        if not np.isfinite(norm2):
            raise FloatingPointError("split: the two-site wavefunction has non-finite entries")
        packed = None
    else:
        packed = np.zeros(total + 1)
    if packed is not None:
        mats = [packed[o0:o0 + nr * nc].reshape((nr, nc), order="F") for (nr, nc, o0) in shapes]
        for (Qm, ro, co), Mm in zip(meta, mats):
            seen = set()
            for (rk, ck, o0, r, c, f) in parts[Qm]:
                blk = v[o0:o0 + r * c].reshape((r, c), order="F")
                dst = Mm[ro[rk]:ro[rk] + r, co[ck]:co[ck] + c]
                if (rk, ck) in seen:
                    dst += f * blk
                else:
                    np.multiply(blk, f, out=dst)
                    seen.add((rk, ck))
        t1 = time.perf_counter()
        efac = eigh(mats, 0 if left_to_right else 1) if eigh is not None else None  # None: declined
        if efac is not None:
            norm2 = float(np.dot(packed, packed))
    fac = {}
    allsv = []
    if efac is not None:
        floor = max(eps_svd, math.sqrt(eps_rho * norm2))
        for (Qm, ro, co), (lam, W, C) in zip(meta, efac):
            S = np.sqrt(np.maximum(lam, 0.0))
            X, Y = (W, C) if left_to_right else (C, W.T)  # X Y = Psi, X or Y orthonormal
            fac[Qm] = (X, S, Y, ro, co)
            allsv.extend((sv, Qm, i) for i, sv in enumerate(S))
    else:
        floor = eps_svd
        for (Qm, ro, co), (U, S, Vt) in zip(meta, (svd or block_svd_host)(mats)):  # host packing above
            X, Y = (U, S[:, None] * Vt) if left_to_right else (U * S[None, :], Vt)
            fac[Qm] = (X, S, Y, ro, co)
            allsv.extend((sv, Qm, i) for i, sv in enumerate(S))
    t2 = time.perf_counter()
    allsv.sort(key=lambda x: -x[0])
    keep = [x for x in allsv[:M] if x[0] > floor]
    kept = {}
    for (sv, Qm, i) in keep:
        kept[Qm] = max(kept.get(Qm, 0), i + 1)
    deg = lambda q: sym.coeff_dot(q)
    trunc = float(sum(deg(x[1]) * x[0] ** 2 for x in allsv[len(keep):]))
    p = np.array([x[0] ** 2 for x in keep])
    w = np.array([deg(x[1]) for x in keep])
    tot = float((w * p).sum())
    entropy = 0.0
    if tot > 0:
        p = p / tot
        entropy = float(-(w[p > 0] * p[p > 0] * np.log(p[p > 0])).sum())
    Al = [dict() for _ in qloc]
    Ar = [dict() for _ in qloc]
    for Qm, k in kept.items():
        X, S, Y, ro, co = fac[Qm]
        X, Y = X[:, :k], Y[:k, :]
        # site tensors in column-major order, the layout the environment
        # update and the next pivot read without another copy
        for (sl, qin), o in ro.items():
            Al[sl][(qin, Qm)] = X[o:o + rdims[Qm][(sl, qin)], :].copy(order="F")
        for (sr, qout), o in co.items():
            src = Y[:, o:o + cdims[Qm][(sr, qout)]]
            dst = np.empty(src.shape, order="F")
            np.multiply(src, coeff_split3(sym, Qm, qout, qloc[sr]), out=dst)
            Ar[sr][(Qm, qout)] = dst
    if timings is not None:
        for key, dt in (("split_assemble", t1 - t0), ("split_factor", t2 - t1), ("split_finish", time.perf_counter() - t2)):
            timings[key] = timings.get(key, 0.0) + dt
    return Al, Ar, trunc, entropy, kept


def _mps_to_blocks(Asite):
    return [[(qin, qout, np.asfortranarray(b)) for (qin, qout), b in sorted(blocks.items())] for blocks in Asite]


# ---------------------------------------------------------------------------
_HOST_MEMORY_RETAINED = False


def retain_freed_host_memory():
    """Serve large host arrays from the heap and keep freed memory mapped
    (glibc mallopt: no mmap'd chunks, trim threshold at its maximum).  A DMRG
    step at M = 4096 allocates and frees a few hundred MB of host arrays (the
    two-site vector, the packed sectors of the split, the site tensors, the
    operand packing of the environment update); served from fresh mappings,
    every 4 KB page of them faults in again at each step (30 ms of the
    environment hand-over alone on the C60 steps).  Process-wide, applied
    once; returns whether the allocator took the settings."""
    global _HOST_MEMORY_RETAINED
    if _HOST_MEMORY_RETAINED:
        return True
    try:
        libc = ctypes.CDLL("libc.so.6")
        M_TRIM_THRESHOLD, M_MMAP_MAX = -1, -4
        ok = libc.mallopt(M_MMAP_MAX, 0) == 1 and libc.mallopt(M_TRIM_THRESHOLD, 2**31 - 1) == 1
    except (OSError, AttributeError):
        ok = False
    _HOST_MEMORY_RETAINED = ok
    return ok


class GpuBackend:
    """Pivot eigensolver and environment updates on the device.

    With a mixed precision the Lanczos solve runs in that precision to its
    rounding level and is refined in FP64 to the same tolerances as an FP64
    solve.  Pivots whose algorithmic work is below `mixed_min_flops` per HxV
    are latency-bound (the mixed products are no faster there) and are solved
    in FP64 directly.  The default floor, 4 GFLOP per HxV, is where the
    split-BF16 HxV runs about twice as fast as FP64 (M = 2048 U(1) chains);
    below it the FP64 refinement (about a third more HxV per solve) costs more
    than the mixed iterations save (HeisenbergU1 L=32, M=1024: 3.00 vs 2.60 s
    per sweep with a 0.5 GFLOP floor).

    With `device_envs` (default) every environment the sweep contracts stays
    on the device (env.DeviceEnv) and feeds the next contraction and pivot
    without a host copy; the results are bit-identical to host environments.

    The split factors on the device through the reduced density matrix
    (block_eigh_device) when a middle sector has both dimensions >=
    device_split_min (default 256), on host LAPACK otherwise.  Pivots with at
    least tune_min_flops of work per HxV get autotuned tiles (cached per
    block-size histogram, so a sweep tunes once per shape).

    Pivots of at least `device_psi_min` elements build their Lanczos start
    vector on the device (contract_AA2) and leave the eigenvector there for
    the split (DevicePsi).

    With `retain_host_memory` (default) the backend asks the C library's
    allocator to keep freed host memory mapped
    (retain_freed_host_memory)."""

    def __init__(self, precision="fp64", refine_fp64=True, tol_eigval=1e-10, tol_state=1e-6, max_iter=100,
                 mixed_min_flops=4e9, device_svd_min=0, device_envs=True, refine_method="correction",
                 device_split_min=256, tune_min_flops=1e9, retain_host_memory=True, device_psi_min=65536):
        if retain_host_memory:
            retain_freed_host_memory()
        self.precision = precision
        self.device_envs = device_envs
        self.refine = refine_fp64
        self.refine_method = refine_method
        self.tol_eigval, self.tol_state, self.max_iter = tol_eigval, tol_state, max_iter
        self.mixed_min_flops = mixed_min_flops
        self.device_svd_min = device_svd_min
        self.device_split_min = device_split_min
        self.tune_min_flops = tune_min_flops
        self.device_psi_min = device_psi_min
        self.hxv_calls = 0
        self.mixed_solves = 0
        self.device_svds = 0
        self.device_splits = 0

    def eigh(self, mats, side):
        """Density-matrix factorisation of the split on the device, or None
        (host SVD) when every middle sector is small."""
        if not self.device_split_min or not mats or max(min(m.shape) for m in mats) < self.device_split_min:
            return None
        self.device_splits += 1
        return block_eigh_device(mats, side)

    def eigh_gather(self, dev, shapes, dst, side):
        """eigh() with the sector matrices built on the device from a
        DevicePsi (block_eigh_gather_device), or None as eigh() declines."""
        if not self.device_split_min or not shapes or max(min(sh) for sh in shapes) < self.device_split_min:
            return None
        self.device_splits += 1
        return block_eigh_gather_device(dev, shapes, dst, side)

    def svd(self, mats):
        """Two-site split SVDs: on the device (vmps_block_svd) when
        device_svd_min > 0 and the largest block has both dimensions >=
        device_svd_min, else host LAPACK.  The device Jacobi SVD (one
        workgroup per block) is slower than LAPACK on the box's 16 host cores
        at every size measured (64 to 1024), so the default keeps the host."""
        if self.device_svd_min and mats and max(min(m.shape) for m in mats) >= self.device_svd_min:
            self.device_svds += 1
            return block_svd_device(mats)
        return block_svd_host(mats)

    def prepare(self, pd):
        """The pivot's plan and tile choice (they do not need the start
        vector).  Returns (plan, precision, seconds)."""
        from .pivot import PivotMatrix2
        t0 = time.perf_counter()
        H = PivotMatrix2(pd, precision="fp64")
        st = H.stats()
        flops = st["flops_alg"]
        tune = self.tune_min_flops is not None and flops >= self.tune_min_flops
        prec = self.precision
        if prec != "fp64" and flops < self.mixed_min_flops:
            prec = "fp64"
        # A mixed solve applies the FP64 operator twice (the refinement's
        # residuals): on pivots with blocks of 256 and more, whose FP64 HxV
        # takes the big tiles in any case, its tuning (six FP64 HxV) would
        # cost more than it could save, and the split-BF16 tuning below
        # chooses the shared explicit-U setting again anyway.
        if tune and (prec == "fp64" or st["max_block_dim"] < 256):
            H.autotune(iters=1)
        if prec != "fp64":
            H.set_precision(prec)
            if tune:
                H.autotune(iters=1)
            self.mixed_solves += 1
        return H, prec, time.perf_counter() - t0

    def device_psi(self, pd):
        """Whether the step builds its start vector on the device and leaves
        the eigenvector there for the split: for pivots of at least
        device_psi_min elements (on smaller ones the block transfers and
        GEMM launches cost more than the host products)."""
        return pd.vsize >= self.device_psi_min

    def solve_two_site(self, pd, Al, Ar, qloc, seed=0, keep_device=False):
        """contract_AA2 on the device into the start vector (a zero vector is
        replaced by a random one), then solve() on it without a host copy of
        the start vector.  Returns (energy, eigenvector, result dict with
        "psi_seconds", the time of the start-vector build); the eigenvector is
        a host array, or with `keep_device` a DevicePsi (the caller closes it,
        which also releases the plan)."""
        from .pivot import DeviceVector, contract_aa2, lanczos_device, vec_fill_random, vec_norm
        H, prec, t_plan = self.prepare(pd)
        dv = None
        kept = False
        try:
            t0 = time.perf_counter()
            dv = DeviceVector(H)
            contract_aa2(H, dv, *two_site_psi_terms(pd, Al, Ar, qloc))
            if vec_norm(H, dv) == 0.0:
                vec_fill_random(H, dv, seed)
            t_psi = time.perf_counter() - t0
            r = lanczos_device(H, dv, precision=prec, refine_fp64=self.refine, max_iter=self.max_iter,
                               tol_eigval=self.tol_eigval, tol_state=self.tol_state,
                               refine_tol_eigval=self.tol_eigval, refine_tol_state=self.tol_state,
                               refine_method=self.refine_method)
            r["flops_alg"] = H.stats()["flops_alg"]
            if keep_device:
                v = DevicePsi(H, dv)
                kept = True
            else:
                v = dv.download()
        finally:
            if not kept:
                if dv is not None:
                    dv.free()
                H.close()
        r["precision"] = prec
        r["plan_seconds"] = t_plan
        r["psi_seconds"] = t_psi
        self.hxv_calls += r["hxv_calls"]
        return r["energy"], v, r

    def solve(self, pd, v0, prepared=None):
        from .pivot import lanczos
        H, prec, t_plan = prepared if prepared is not None else self.prepare(pd)
        try:
            v, r = lanczos(H, v0=v0, precision=prec, refine_fp64=self.refine, max_iter=self.max_iter,
                           tol_eigval=self.tol_eigval, tol_state=self.tol_state,
                           refine_tol_eigval=self.tol_eigval, refine_tol_state=self.tol_state,
                           refine_method=self.refine_method, overwrite_v0=True)
            r["flops_alg"] = H.stats()["flops_alg"]
        finally:
            H.close()
        r["precision"] = prec
        r["plan_seconds"] = t_plan  # plan set-up and tile choice, part of the solve
        self.hxv_calls += r["hxv_calls"]
        return r["energy"], v, r

    def contract_L(self, site, Lold, A, timings=None):
        from . import env
        blocks = _mps_to_blocks(A)
        return env.contract_L(site, Lold, blocks, blocks, device=self.device_envs, timings=timings)

    def contract_R(self, site, Rold, A, timings=None):
        from . import env
        blocks = _mps_to_blocks(A)
        return env.contract_R(site, Rold, blocks, blocks, device=self.device_envs, timings=timings)


class Chain:
    """An MPS with its MPO and environments: A[l] per site (list over the
    local index of {(qin, qout): block}), bonds[b] = {q: dim} for b = 0..L,
    Lenv[b] / Renv[b] the environments on bond b (None where not built)."""

    def __init__(self, sites, Q, M, A, bonds):
        self.sites = sites
        self.sym = sites[0].sym
        self.qloc = sites[0].qloc
        self.L = len(sites)
        self.Q = tuple(Q)
        self.M = M
        self.A = A
        self.bonds = bonds
        vac = self.sym.vacuum()
        self.Lenv = [None] * (self.L + 1)
        self.Renv = [None] * (self.L + 1)
        self.Lenv[0] = [(vac, vac, tuple(sites[0].aux_in[0][0]), [np.ones((1, 1))])]
        self.Renv[self.L] = [(self.Q, self.Q, tuple(sites[-1].aux_out[0][0]), [np.ones((1, 1))])]

    def copy(self):
        """A second handle on the same state: steps replace A[l], bonds[b] and
        environments rather than change them in place, so the lists are copied
        and their entries shared."""
        c = Chain.__new__(Chain)
        c.__dict__.update(self.__dict__)
        c.A = list(self.A)
        c.bonds = [dict(b) for b in self.bonds]
        c.Lenv = list(self.Lenv)
        c.Renv = list(self.Renv)
        return c

    def pivot(self, l):
        """Two-site pivot (l, l+1): PivotData with the environments Lenv[l],
        Renv[l+2] and every wavefunction block the bonds allow."""
        sym, sites, bonds = self.sym, self.sites, self.bonds
        pd = synth.PivotData(sym, sites[l], sites[l + 1], sorted(bonds[l].items()), sorted(bonds[l + 2].items()))
        pd.L = self.Lenv[l]
        pd.R = self.Renv[l + 2]
        pd.basis = synth.two_site_basis(sym, sites[l].qloc, sites[l + 1].qloc)
        rdim = bonds[l + 2]
        pd.V = []
        for (qm, s1, s3) in pd.basis:
            bl = []
            for qin, din in sorted(bonds[l].items()):
                for qout in sym.reduce(qin, qm):
                    if qout in rdim:
                        bl.append((qin, qout, din, rdim[qout]))
            pd.V.append(bl)
        return pd


def build_envs(chain, backend, p):
    """Environments Lenv[1..p] and Renv[p+2..L-1] of the pivot (p, p+1),
    contracted from the chain's ends through the backend."""
    for b in range(1, p + 1):
        chain.Lenv[b] = backend.contract_L(chain.sites[b - 1], chain.Lenv[b - 1], chain.A[b - 1])
    for b in range(chain.L - 1, p + 1, -1):
        chain.Renv[b] = backend.contract_R(chain.sites[b], chain.Renv[b + 1], chain.A[b])
    return chain


def center_chain(sites, Q, M, p, backend, rng, bonds=None):
    """A random state prepared for two-site steps at the pivot (p, p+1): sites
    left of p left-normalized, sites right of p+1 right-normalized, and the
    environments Lenv[0..p] and Renv[p+2..L] contracted through the backend.
    Bond sectors default to su2_gaussian_bonds (SU(2) chains)."""
    sym, qloc, L = sites[0].sym, sites[0].qloc, len(sites)
    bonds = bonds or su2_gaussian_bonds(sym, qloc, L, Q, M)
    A = random_blocks(sym, qloc, bonds, rng)
    A, bonds = left_normalize(A, bonds, qloc, sym, p)
    A, bonds = right_normalize(A, bonds, qloc, sym, stop=p + 2)
    return build_envs(Chain(sites, Q, M, A, bonds), backend, p)


def warm_chain(sites, Q, schedule, p, backend, rng, log=None):
    """A DMRG state converged at small bond dimensions, ready for steps at the
    pivot (p, p+1): from a random state, one full sweep (right, then left) at
    each M of `schedule`, then a right half sweep up to the pivot (p-1, p), so
    that sites < p are left-normalized, site p carries the state and sites >
    p are right-normalized."""
    L = len(sites)
    sym, qloc = sites[0].sym, sites[0].qloc
    A, bonds = random_mps(sites, Q, schedule[0], rng)
    A, bonds = right_normalize(A, bonds, qloc, sym)
    chain = Chain(sites, Q, schedule[0], A, bonds)
    for b in range(L - 1, 1, -1):
        chain.Renv[b] = backend.contract_R(sites[b], chain.Renv[b + 1], chain.A[b])
    for M in schedule:
        chain.M = M
        t0 = time.perf_counter()
        for l in range(0, L - 1):
            E, _, _ = dmrg_step(chain, backend, l, "right", rng=rng)
        for l in range(L - 2, -1, -1):
            E, _, _ = dmrg_step(chain, backend, l, "left", rng=rng)
        if log:
            log("warm-up sweep M=%d  E = %.10f  (%.1f s)" % (M, E, time.perf_counter() - t0))
    for l in range(0, p):
        E, _, _ = dmrg_step(chain, backend, l, "right", rng=rng)
    chain.energy = E
    return chain


def _orth_complement(Old, ncols, rng):
    """ncols orthonormal columns orthogonal to the orthonormal columns of Old
    (random, two rounds of projection, then QR)."""
    N = rng.standard_normal((Old.shape[0], ncols))
    for _ in range(2):
        if Old.shape[1]:
            N -= Old @ (Old.T @ N)
    q, _ = np.linalg.qr(N)
    return q


def expand_chain(chain, M, p, backend, rng):
    """The same state with larger bonds: every bond grows towards the sectors
    of su2_gaussian_bonds(M) (never below its current dimensions), the new
    basis states orthonormal completions of the old ones, so sites < p stay
    left-normalized, sites > p right-normalized, and site p holds the old
    state in its old-old block (zeros elsewhere).  The environments of the
    pivot (p, p+1) are rebuilt.  This is a DMRG state right after its bond
    dimension was raised: the next two-site steps start from an excellent
    guess in a much larger space."""
    sym, qloc, L = chain.sym, chain.qloc, chain.L
    old, A = chain.bonds, chain.A
    tgt = su2_gaussian_bonds(sym, qloc, L, chain.Q, M)
    new = [dict(b) for b in old]
    newA = [None] * L
    # left part: site l defines bond l + 1
    for l in range(p):
        site = [dict() for _ in qloc]
        nb = {}
        for qout in sorted(set(old[l + 1]) | set(tgt[l + 1])):
            parts = [(s, qin) for s in range(len(qloc)) for qin in sorted(new[l]) if qout in sym.reduce(qin, qloc[s])]
            rows = sum(new[l][qin] for _, qin in parts)
            d_old = old[l + 1].get(qout, 0)
            D = min(max(d_old, tgt[l + 1].get(qout, 0)), rows)
            if D == 0:
                continue
            X = np.zeros((rows, D))
            r0, offs = 0, []
            for (s, qin) in parts:
                blk = A[l][s].get((qin, qout))
                if blk is not None and d_old:
                    X[r0:r0 + blk.shape[0], :d_old] = blk
                offs.append(r0)
                r0 += new[l][qin]
            if D > d_old:
                X[:, d_old:] = _orth_complement(X[:, :d_old], D - d_old, rng)
            for (s, qin), o in zip(parts, offs):
                site[s][(qin, qout)] = np.ascontiguousarray(X[o:o + new[l][qin], :])
            nb[qout] = D
        new[l + 1] = nb
        newA[l] = site
    # right part: site l defines bond l
    for l in range(L - 1, p, -1):
        site = [dict() for _ in qloc]
        nb = {}
        for qin in sorted(set(old[l]) | set(tgt[l])):
            parts = [(s, qout) for s in range(len(qloc)) for qout in sorted(new[l + 1])
                     if qout in sym.reduce(qin, qloc[s])]
            cols = sum(new[l + 1][qo] for _, qo in parts)
            d_old = old[l].get(qin, 0)
            D = min(max(d_old, tgt[l].get(qin, 0)), cols)
            if D == 0:
                continue
            w = [math.sqrt(sym.coeff_dot(qo) / sym.coeff_dot(qin)) for _, qo in parts]
            Y = np.zeros((D, cols))
            c0, offs = 0, []
            for (s, qo), wi in zip(parts, w):
                blk = A[l][s].get((qin, qo))
                if blk is not None and d_old:
                    Y[:d_old, c0:c0 + blk.shape[1]] = wi * blk
                offs.append(c0)
                c0 += new[l + 1][qo]
            if D > d_old:
                Y[d_old:, :] = _orth_complement(Y[:d_old, :].T, D - d_old, rng).T
            for (s, qo), wi, o in zip(parts, w, offs):
                site[s][(qin, qo)] = np.ascontiguousarray(Y[:, o:o + new[l + 1][qo]] / wi)
            nb[qin] = D
        new[l] = nb
        newA[l] = site
    # centre: the old state in the old-old block
    site = [dict() for _ in qloc]
    for s in range(len(qloc)):
        for qin, din in new[p].items():
            for qout in sym.reduce(qin, qloc[s]):
                if qout not in new[p + 1]:
                    continue
                blk = np.zeros((din, new[p + 1][qout]))
                b = A[p][s].get((qin, qout))
                if b is not None:
                    blk[:b.shape[0], :b.shape[1]] = b
                site[s][(qin, qout)] = blk
    newA[p] = site
    out = Chain(chain.sites, chain.Q, M, newA, new)
    return build_envs(out, backend, p)


def dmrg_step(chain, backend, l, direction, rng=None, timings=None):
    """One two-site DMRG step on the pivot (l, l+1) (DmrgSolver::iteration_two,
    solvers/DmrgSolver.h:1398-1523): contract_AA2, lowest eigenpair of H_eff,
    split_AA2 with truncation to chain.M, and the environment update in the
    sweep direction ("right": Lenv[l+1], "left": Renv[l+1]).  Returns
    (energy, truncation weight, solver info); `timings`, a dict, collects
    the seconds of each phase."""
    t = time.perf_counter
    t0 = t()
    pd = chain.pivot(l)
    if hasattr(backend, "solve_two_site") and backend.device_psi(pd):  # contract_AA2 on the device, inside the solve
        t1 = t()
        E, v, info = backend.solve_two_site(pd, chain.A[l], chain.A[l + 1], chain.qloc, seed=l + 1,
                                            keep_device=hasattr(backend, "eigh_gather"))
        t2 = t()
        t1 += info["psi_seconds"]
    else:
        v0 = two_site_psi(pd, chain.A[l], chain.A[l + 1], chain.qloc)
        if not np.any(v0):
            v0 = (rng or np.random.default_rng(l)).uniform(-1, 1, pd.vsize)
        t1 = t()
        E, v, info = backend.solve(pd, v0)
        t2 = t()
    try:
        Al, Ar, tw, ent, kept = split_psi(pd, v, chain.qloc, chain.M, left_to_right=(direction == "right"),
                                          svd=getattr(backend, "svd", None), eigh=getattr(backend, "eigh", None),
                                          timings=timings, eigh_gather=getattr(backend, "eigh_gather", None))
    finally:
        if isinstance(v, DevicePsi):
            v.close()
    chain.A[l], chain.A[l + 1] = Al, Ar
    chain.bonds[l + 1] = {qm: k for qm, k in kept.items()}
    t3 = t()
    kw = {"timings": timings} if timings is not None else {}  # backends without sub-phase timings take none
    if direction == "right":
        chain.Lenv[l + 1] = backend.contract_L(chain.sites[l], chain.Lenv[l], chain.A[l], **kw)
    else:
        chain.Renv[l + 1] = backend.contract_R(chain.sites[l + 1], chain.Renv[l + 2], chain.A[l + 1], **kw)
    t4 = t()
    if timings is not None:
        for key, dt in (("psi", t1 - t0), ("solve", t2 - t1), ("split", t3 - t2), ("env", t4 - t3)):
            timings[key] = timings.get(key, 0.0) + dt
        if "plan_seconds" in info:  # the part of "solve" spent building the plan
            timings["plan"] = timings.get("plan", 0.0) + info["plan_seconds"]
    return E, tw, info


def run_dmrg(model, L, M, sweeps, Q=None, seed=20261202, backend=None, params=None, verbose=False):
    """Ground state by two-site DMRG.  Returns a dict with energies per half
    sweep, the final energy, the truncation weights and timings."""
    params = params or {}
    backend = backend or GpuBackend()
    sites = models.chain_mpo(model, L, **params)
    qloc = sites[0].qloc
    sym = sites[0].sym
    if Q is None:
        Q = {"HeisenbergU1": (0,), "HubbardU1xU1": (L, 0), "HeisenbergSU2": (1,), "HubbardSU2xU1": (1, L),
             "HeisenbergSU2LR": (1,), "HeisenbergSU2Full": (1,), "HubbardSU2xU1Ladder": (1, L)}[model]
    Q = tuple(Q)
    rng = np.random.default_rng(seed)
    A, bonds = random_mps(sites, Q, M, rng)
    A, bonds = right_normalize(A, bonds, qloc, sym)
    chain = Chain(sites, Q, M, A, bonds)
    for l in range(L - 1, 1, -1):
        chain.Renv[l] = backend.contract_R(sites[l], chain.Renv[l + 1], chain.A[l])
    energies, truncs = [], []
    t_sweeps = []
    E = None
    for sweep in range(sweeps):
        t0 = time.perf_counter()
        for direction in ("right", "left"):
            ls = range(0, L - 1) if direction == "right" else range(L - 2, -1, -1)
            for l in ls:
                E, tw, _ = dmrg_step(chain, backend, l, direction, rng=rng)
                truncs.append(tw)
            energies.append(E)
            if verbose:
                print("sweep %d %s  E = %.12f  max trunc = %.2e" % (sweep, direction, E, max(truncs[-(L - 1):])))
        t_sweeps.append(time.perf_counter() - t0)
    return {"energy": E, "energies": energies, "truncation": truncs, "sweep_seconds": t_sweeps,
            "bond_dims": [sum(b.values()) for b in chain.bonds]}
