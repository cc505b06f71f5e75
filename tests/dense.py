"""Brute-force dense effective Hamiltonians for abelian pivots (test helper).

Two sites:  out[i', s1, s3, j'] = sum L_a[i', i] W12[a, w, s1, s2] W34[w, b, s3, s4] V[i, s2, s4, j] R_b[j, j']
One site:   out[i', s1, j']     = sum L_a[i', i] W[a, b, s1, s2] V[i, s2, j] R_b[j, j']
Zero sites: out[i', j']         = sum_a L_a[i', i] C[i, j] R_a[j, j']
summed over the pivot's terms, on dense bond / local / auxiliary indices,
which needs none of the block-structure bookkeeping under test.
"""
import numpy as np


def _off(sectors):
    off, o = {}, 0
    for q, d in sectors:
        off[tuple(q)] = o
        o += d
    return off, o


def _env_dense(env, bond_in, bond_out, aux):
    ioff, it = _off(bond_in)
    ooff, ot = _off(bond_out)
    aoff, na = _off(aux)
    E = np.zeros((na, it, ot))
    for (qin, qout, qmid, blocks) in env:
        for a, blk in enumerate(blocks):
            if blk is None:
                continue
            E[aoff[qmid] + a, ioff[qin]:ioff[qin] + blk.shape[0], ooff[qout]:ooff[qout] + blk.shape[1]] = blk
    return E


def _aux_of(env):
    seen = {}
    for (_, _, qmid, blocks) in env:
        seen.setdefault(tuple(qmid), len(blocks))
    return list(seen.items())


def dense_operands(pd, W12=None, W34=None, L=None, R=None):
    """Dense L, W12, W34, R of one term (default: the first) of a two-site pivot."""
    assert not pd.sym.non_abelian
    W12 = W12 or pd.W12
    W34 = W34 or pd.W34
    L = pd.L if L is None else L
    R = pd.R if R is None else R
    loff, lt = _off(pd.left)
    roff, rt = _off(pd.right)
    return (_env_dense(L, pd.left, pd.left, W12.aux_in), W12.dense(), W34.dense(),
            _env_dense(R, pd.right, pd.right, W34.aux_out), loff, roff, lt, rt)


def _local_index(pd, s):
    if pd.nsite == 2:
        _, s1, s3 = pd.basis[s]
        return (s1, s3)
    if pd.nsite == 1:
        return (pd.basis[s][1],)
    return ()


def _local_shape(pd):
    if pd.nsite == 2:
        return (len(pd.W12.qloc), len(pd.W34.qloc))
    if pd.nsite == 1:
        return (len(pd.W12.qloc),)
    return ()


def to_dense(pd, v, loff, roff, lt, rt):
    D = np.zeros((lt,) + _local_shape(pd) + (rt,))
    for (s, qin, qout, off, r, c) in pd.block_offsets():
        idx = (slice(loff[qin], loff[qin] + r),) + _local_index(pd, s) + (slice(roff[qout], roff[qout] + c),)
        D[idx] = v[off:off + r * c].reshape((r, c), order="F")
    return D


def from_dense(pd, D, loff, roff):
    out = np.zeros(pd.vsize)
    for (s, qin, qout, off, r, c) in pd.block_offsets():
        idx = (slice(loff[qin], loff[qin] + r),) + _local_index(pd, s) + (slice(roff[qout], roff[qout] + c),)
        out[off:off + r * c] = D[idx].ravel(order="F")
    return out


def _terms(pd):
    return [(pd.W12, pd.W34, pd.L, pd.R)] + [(t.W12, t.W34, t.L, t.R) for t in pd.extra]


def dense_hxv(pd, vin):
    assert not pd.sym.non_abelian
    loff, lt = _off(pd.left)
    roff, rt = _off(pd.right)
    V = to_dense(pd, vin, loff, roff, lt, rt)
    out = np.zeros_like(V)
    for (W12, W34, L, R) in _terms(pd):
        if pd.nsite == 2:
            Ld = _env_dense(L, pd.left, pd.left, W12.aux_in)
            Rd = _env_dense(R, pd.right, pd.right, W34.aux_out)
            T = np.einsum("axy,yqsz->axqsz", Ld, V)             # L_a V
            T = np.einsum("awpq,axqsz->wxpsz", W12.dense(), T)  # W12
            T = np.einsum("wbrs,wxpsz->bxprz", W34.dense(), T)  # W34
            out += np.einsum("bxprz,bzv->xprv", T, Rd)          # R_b
        elif pd.nsite == 1:
            Ld = _env_dense(L, pd.left, pd.left, W12.aux_in)
            Rd = _env_dense(R, pd.right, pd.right, W12.aux_out)
            T = np.einsum("axy,yqz->axqz", Ld, V)
            T = np.einsum("abpq,axqz->bxpz", W12.dense(), T)
            out += np.einsum("bxpz,bzv->xpv", T, Rd)
        else:
            aux = _aux_of(L)
            Ld = _env_dense(L, pd.left, pd.left, aux)
            Rd = _env_dense(R, pd.right, pd.right, aux)
            out += np.einsum("axy,yz,azv->xv", Ld, V, Rd)
    return from_dense(pd, out, loff, roff)
