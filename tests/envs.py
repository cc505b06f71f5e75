"""Random environments and site tensors for the contract_L / contract_R tests,
and their dense forms (abelian symmetries)."""
import numpy as np


def _off(sectors):
    off, o = {}, 0
    for q, d in sectors:
        off[tuple(q)] = o
        o += d
    return off, o


def random_site_tensor(site, left, right, rng):
    """A[s] blocks (qin, qin + qloc[s]) for every pair of bond sectors."""
    rdim = dict(right)
    A = []
    for qs in site.qloc:
        blocks = []
        for (qin, din) in left:
            for qout in site.sym.reduce(qin, qs):
                if qout in rdim:
                    blocks.append((qin, qout, np.asfortranarray(rng.standard_normal((din, rdim[qout])))))
        A.append(blocks)
    return A


def random_env(site, bond, rng, side):
    """L: {in, out, mid} with out + mid = in; R: {in, out, mid} with in + mid = out."""
    d = dict(bond)
    out = []
    aux = site.aux_in if side == "L" else site.aux_out
    for (qmid, adim) in aux:
        for (qa, da) in bond:
            for qb in site.sym.reduce(qa, qmid):
                if qb in d:
                    if side == "L":  # qa = out (ket), qb = in (bra)
                        blocks = [np.asfortranarray(rng.standard_normal((d[qb], da))) for _ in range(adim)]
                        out.append((qb, qa, tuple(qmid), blocks))
                    else:  # qa = in (ket), qb = out (bra)
                        blocks = [np.asfortranarray(rng.standard_normal((da, d[qb]))) for _ in range(adim)]
                        out.append((qa, qb, tuple(qmid), blocks))
    return out


def dense_A(A, left, right, d):
    lo, lt = _off(left)
    ro, rt = _off(right)
    D = np.zeros((d, lt, rt))
    for s, blocks in enumerate(A):
        for (qin, qout, b) in blocks:
            D[s, lo[qin]:lo[qin] + b.shape[0], ro[qout]:ro[qout] + b.shape[1]] = b
    return D


def dense_env(E, bond, aux):
    bo, bt = _off(bond)
    ao, at = _off(aux)
    D = np.zeros((at, bt, bt))
    for (qin, qout, qmid, blocks) in E:
        for a, b in enumerate(blocks):
            if b is not None:
                D[ao[tuple(qmid)] + a, bo[qin]:bo[qin] + b.shape[0], bo[qout]:bo[qout] + b.shape[1]] = b
    return D
