"""Environment updates contract_L / contract_R on the device against a dense
brute-force contraction (abelian symmetries: all CGC = 1)."""
import numpy as np
import pytest

from vmps_amd import env, models, synth

pytestmark = pytest.mark.gpu


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


@pytest.mark.parametrize("model", ["heis", "hub"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("bf16x3", 5e-5)])
def test_contract_L_and_R_vs_dense(model, precision, tol):
    rng = np.random.default_rng(4)
    if model == "heis":
        site = models.heisenberg_u1_mpo()
        left, right = synth.heisenberg_bond(5, 14), synth.heisenberg_bond(6, 14)
    else:
        site = models.hubbard_u1xu1_mpo()
        left = synth.hubbard_bond(3, 20, alpha=1.0, lam=1.2)
        right = synth.hubbard_bond(4, 20, alpha=1.0, lam=1.2)
    d = len(site.qloc)
    Abra = random_site_tensor(site, left, right, rng)
    Aket = random_site_tensor(site, left, right, rng)
    W = site.dense()  # [a, b, s1, s2]
    DAb, DAk = dense_A(Abra, left, right, d), dense_A(Aket, left, right, d)

    Lold = random_env(site, left, rng, "L")
    Lnew = env.contract_L(site, Lold, Abra, Aket, precision=precision)
    DL = dense_env(Lold, left, site.aux_in)
    ref = np.einsum("sxp,axy,absq,qyr->bpr", DAb, DL, W, DAk)  # Abra^T L_a Aket, weighted by W
    got = dense_env(Lnew, right, site.aux_out)
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)

    Rold = random_env(site, right, rng, "R")
    Rnew = env.contract_R(site, Rold, Abra, Aket, precision=precision)
    DR = dense_env(Rold, right, site.aux_out)
    ref = np.einsum("qxy,byw,abpq,pzw->axz", DAk, DR, W, DAb)  # Aket R_b Abra^T
    got = dense_env(Rnew, left, site.aux_in)
    assert np.linalg.norm(got - ref) <= tol * np.linalg.norm(ref)
