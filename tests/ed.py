"""Exact diagonalisation of small open chains, built from Kronecker products
of local operators (independent of the MPO construction)."""
import numpy as np
from scipy.sparse import csr_matrix, identity, kron
from scipy.sparse.linalg import eigsh


def _op(local, i, L, d, string=None):
    """local operator on site i; `string` is applied on every site < i."""
    mats = []
    for j in range(L):
        if j == i:
            mats.append(csr_matrix(local))
        elif string is not None and j < i:
            mats.append(csr_matrix(string))
        else:
            mats.append(identity(d, format="csr"))
    out = mats[0]
    for m in mats[1:]:
        out = kron(out, m, format="csr")
    return out


def heisenberg_ground(L, J=1.0, two_sz=0):
    Sp = np.array([[0.0, 0.0], [1.0, 0.0]])
    Sm = Sp.T
    Sz = np.diag([-0.5, 0.5])
    H = 0
    for i in range(L - 1):
        H = H + J * (_op(Sz, i, L, 2) @ _op(Sz, i + 1, L, 2)
                     + 0.5 * (_op(Sp, i, L, 2) @ _op(Sm, i + 1, L, 2) + _op(Sm, i, L, 2) @ _op(Sp, i + 1, L, 2)))
    # sector 2Sz: basis index bits, state 1 = up (+1), 0 = down (-1); site 0 is the most significant
    idx = [n for n in range(2 ** L) if sum(1 if (n >> (L - 1 - i)) & 1 else -1 for i in range(L)) == two_sz]
    Hs = H[idx][:, idx]
    return float(eigsh(Hs, k=1, which="SA", tol=1e-13)[0][0]) if Hs.shape[0] > 50 else \
        float(np.linalg.eigvalsh(Hs.toarray())[0])


def hubbard_ground(L, t=1.0, U=8.0, N=None, two_sz=0):
    """Fermions with an explicit Jordan-Wigner string; local basis |0>, |up>, |dn>, |updn>."""
    N = L if N is None else N
    c_up = np.zeros((4, 4))
    c_dn = np.zeros((4, 4))
    c_up[0, 1] = 1.0
    c_up[2, 3] = 1.0
    c_dn[0, 2] = 1.0
    c_dn[1, 3] = -1.0
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    n_up = c_up.T @ c_up
    n_dn = c_dn.T @ c_dn
    H = 0
    for i in range(L):
        H = H + U * _op(n_up @ n_dn, i, L, 4)
    for i in range(L - 1):
        for c in (c_up, c_dn):
            ci = _op(c, i, L, 4, string=F)
            cj = _op(c, i + 1, L, 4, string=F)
            H = H - t * (ci.T @ cj + cj.T @ ci)
    nloc = [0, 1, 1, 2]
    szloc = [0, 1, -1, 0]
    idx = []
    for n in range(4 ** L):
        digits = [(n // 4 ** (L - 1 - i)) % 4 for i in range(L)]
        if sum(nloc[x] for x in digits) == N and sum(szloc[x] for x in digits) == two_sz:
            idx.append(n)
    Hs = H[idx][:, idx]
    return float(eigsh(Hs, k=1, which="SA", tol=1e-13)[0][0])


def heisenberg_dense(L, J=1.0, S=0.5, Js=None, Jmat=None):
    """Dense spin-S Heisenberg chain and total Sz, S^2 (basis: site 0 most
    significant).  Js = [J_1, J_2, ...] couples sites at distance d with J_d
    (default: J at distance 1 only); Jmat, an L x L matrix, couples every
    pair (i, j) with Jmat[i, j] instead."""
    d = int(round(2 * S)) + 1
    m = np.arange(S, -S - 1, -1)[::-1]  # local Sz values, index 0 = lowest
    Sz = np.diag(m)
    Sp = np.zeros((d, d))
    for i in range(d - 1):
        Sp[i + 1, i] = np.sqrt(S * (S + 1) - m[i] * (m[i] + 1))
    Sm = Sp.T
    ops = lambda A, i: _op(A, i, L, d).toarray()
    H = np.zeros((d ** L, d ** L))
    if Jmat is not None:
        pairs = [(i, j, Jmat[i][j]) for i in range(L) for j in range(i + 1, L) if Jmat[i][j] != 0.0]
    else:
        Js = [J] if Js is None else list(Js)
        pairs = [(i, i + dist, Jd) for dist, Jd in enumerate(Js, start=1) for i in range(L - dist)]
    for i, j, Jd in pairs:
        H += Jd * (ops(Sz, i) @ ops(Sz, j) + 0.5 * (ops(Sp, i) @ ops(Sm, j) + ops(Sm, i) @ ops(Sp, j)))
    Szt = sum(ops(Sz, i) for i in range(L))
    Spt = sum(ops(Sp, i) for i in range(L))
    S2 = Spt.T @ Spt + Szt @ Szt + Szt  # S^2 = S-S+ + Sz^2 + Sz
    return H, np.diag(Szt), S2


def heisenberg_sector_spectrum(L, two_sz, J=1.0):
    """All eigenvalues of the S=1/2 chain with 2 Sz_total = two_sz."""
    H, sz, _ = heisenberg_dense(L, J)
    idx = np.where(np.isclose(2 * sz, two_sz))[0]
    return np.sort(np.linalg.eigvalsh(H[np.ix_(idx, idx)]))


def heisenberg_multiplet_spectrum(L, S_tot, J=1.0, Js=None, Jmat=None):
    """One eigenvalue per spin-S_tot multiplet: H restricted to the highest-weight
    states (Sz = S_tot and S^2 = S_tot(S_tot+1))."""
    H, sz, S2 = heisenberg_dense(L, J, Js=Js, Jmat=Jmat)
    idx = np.where(np.isclose(sz, S_tot))[0]
    w, U = np.linalg.eigh(S2[np.ix_(idx, idx)])
    P = U[:, np.isclose(w, S_tot * (S_tot + 1))]
    if P.shape[1] == 0:
        return np.zeros(0)
    return np.sort(np.linalg.eigvalsh(P.T @ H[np.ix_(idx, idx)] @ P))


def ladder_bonds(L):
    """Bonds of the two-leg ladder in snake order: rungs (2x, 2x+1), legs (i, i+2)."""
    return [(i, i + 1) for i in range(0, L - 1, 2)] + [(i, i + 2) for i in range(L - 2)]


def hubbard_dense(L, t=1.0, U=8.0, bonds=None):
    """Dense Hubbard model with an explicit Jordan-Wigner string; returns H,
    the diagonals of N and 2 Sz, and S^2.  bonds: hopping pairs (i, j), i < j
    (default: the open chain)."""
    c_up = np.zeros((4, 4))
    c_dn = np.zeros((4, 4))
    c_up[0, 1] = 1.0
    c_up[2, 3] = 1.0
    c_dn[0, 2] = 1.0
    c_dn[1, 3] = -1.0
    F = np.diag([1.0, -1.0, -1.0, 1.0])
    ops = lambda A, i, s=None: _op(A, i, L, 4, string=s).toarray()
    H = np.zeros((4 ** L, 4 ** L))
    for i in range(L):
        H += U * ops(c_up.T @ c_up @ c_dn.T @ c_dn, i)
    for (i, j) in (bonds if bonds is not None else [(i, i + 1) for i in range(L - 1)]):
        for c in (c_up, c_dn):
            ci, cj = ops(c, i, F), ops(c, j, F)
            H -= t * (ci.T @ cj + cj.T @ ci)
    n_up, n_dn = c_up.T @ c_up, c_dn.T @ c_dn
    N = sum(np.diag(ops(n_up + n_dn, i)) for i in range(L))
    two_sz = sum(np.diag(ops(n_up - n_dn, i)) for i in range(L))
    Sp = c_up.T @ c_dn  # on-site S+ (bosonic: no string)
    Spt = sum(ops(Sp, i) for i in range(L))
    Szt = np.diag(0.5 * two_sz)
    S2 = Spt.T @ Spt + Szt @ Szt + Szt
    return H, N, two_sz, S2


def hubbard_multiplet_spectrum(L, N_tot, S_tot, t=1.0, U=8.0, bonds=None):
    """One eigenvalue per spin-S_tot multiplet at particle number N_tot."""
    H, N, two_sz, S2 = hubbard_dense(L, t, U, bonds)
    idx = np.where(np.isclose(N, N_tot) & np.isclose(two_sz, 2 * S_tot))[0]
    w, Uv = np.linalg.eigh(S2[np.ix_(idx, idx)])
    P = Uv[:, np.isclose(w, S_tot * (S_tot + 1))]
    if P.shape[1] == 0:
        return np.zeros(0)
    return np.sort(np.linalg.eigvalsh(P.T @ H[np.ix_(idx, idx)] @ P))
