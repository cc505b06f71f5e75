"""Brute-force dense effective Hamiltonian for abelian pivots (test helper).

H_eff acts as  out[i', s1, s3, j'] = sum L_a[i', i] W12[a, w, s1, s2]
W34[w, b, s3, s4] V[i, s2, s4, j] R_b[j, j']  on dense bond/local indices,
which needs none of the block-structure bookkeeping under test.
"""
import numpy as np


def _off(sectors):
    off, o = {}, 0
    for q, d in sectors:
        off[tuple(q)] = o
        o += d
    return off, o


def dense_operands(pd):
    assert not pd.sym.non_abelian
    loff, lt = _off(pd.left)
    roff, rt = _off(pd.right)
    aoff12, na = _off(pd.W12.aux_in)
    aoff34, nb = _off(pd.W34.aux_out)
    W12 = pd.W12.dense()
    W34 = pd.W34.dense()
    L = np.zeros((na, lt, lt))
    for (qin, qout, qmid, blocks) in pd.L:
        for a, blk in enumerate(blocks):
            if blk is None:
                continue
            L[aoff12[qmid] + a, loff[qin]:loff[qin] + blk.shape[0], loff[qout]:loff[qout] + blk.shape[1]] = blk
    R = np.zeros((nb, rt, rt))
    for (qin, qout, qmid, blocks) in pd.R:
        for b, blk in enumerate(blocks):
            if blk is None:
                continue
            R[aoff34[qmid] + b, roff[qin]:roff[qin] + blk.shape[0], roff[qout]:roff[qout] + blk.shape[1]] = blk
    return L, W12, W34, R, loff, roff, lt, rt


def to_dense(pd, v, loff, roff, lt, rt):
    d12, d34 = len(pd.W12.qloc), len(pd.W34.qloc)
    D = np.zeros((lt, d12, d34, rt))
    for (s, qin, qout, off, r, c) in pd.block_offsets():
        _, s1, s3 = pd.basis[s]
        D[loff[qin]:loff[qin] + r, s1, s3, roff[qout]:roff[qout] + c] = v[off:off + r * c].reshape((r, c), order="F")
    return D


def from_dense(pd, D, loff, roff):
    out = np.zeros(pd.vsize)
    for (s, qin, qout, off, r, c) in pd.block_offsets():
        _, s1, s3 = pd.basis[s]
        out[off:off + r * c] = D[loff[qin]:loff[qin] + r, s1, s3, roff[qout]:roff[qout] + c].ravel(order="F")
    return out


def dense_hxv(pd, vin):
    L, W12, W34, R, loff, roff, lt, rt = dense_operands(pd)
    V = to_dense(pd, vin, loff, roff, lt, rt)
    T = np.einsum("axy,yqsz->axqsz", L, V)            # L_a V
    T = np.einsum("awpq,axqsz->wxpsz", W12, T)        # W12
    T = np.einsum("wbrs,wxpsz->bxprz", W34, T)        # W34
    out = np.einsum("bxprz,bzv->xprv", T, R)          # R_b
    return from_dense(pd, out, loff, roff)
