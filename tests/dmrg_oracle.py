"""CPU backend of the DMRG driver, built from the oracle (test infrastructure):
the oracle's HxV with scipy's ARPACK as eigensolver, and a Python
restatement of contract_L / contract_R for abelian symmetries
(tensors/DmrgContractions.h:34-311 with all CGC = 1)."""
import ctypes

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

import oracle
from vmps_amd import synth


class OracleBackend:
    def __init__(self, tol=1e-13):
        self.lib = oracle.load()
        self.tol = tol

    def solve(self, pd, v0):
        lib = self.lib
        h = synth.feed(lib, "oracle_", pd, oracle.check)
        oracle.check(lib.oracle_pivot2_precalc(h))
        P = ctypes.POINTER(ctypes.c_double)

        def mv(x):
            x = np.ascontiguousarray(np.asarray(x, dtype=np.float64).ravel())
            out = np.zeros_like(x)
            oracle.check(lib.oracle_hxv_omp(h, x.ctypes.data_as(P), out.ctypes.data_as(P), 8))
            return out

        n = pd.vsize
        try:
            if n <= 3:
                Hm = np.stack([mv(np.eye(n)[i]) for i in range(n)], axis=1)
                w, V = np.linalg.eigh(0.5 * (Hm + Hm.T))
                E, v = w[0], V[:, 0]
            else:
                op = LinearOperator((n, n), matvec=mv, dtype=np.float64)
                w, V = eigsh(op, k=1, which="SA", tol=self.tol, v0=v0 if np.any(v0) else None)
                E, v = w[0], V[:, 0]
        finally:
            lib.oracle_pivot2_destroy(h)
        return float(E), v / np.linalg.norm(v), {}

    @staticmethod
    def _wnz(site):
        """(s1, s2, qin, qout, a, b, value) of every MPO nonzero."""
        out = []
        for (s1, s2, k, qi, qo, rows, cols, r, c, v) in site.blocks:
            for a, b, x in zip(r, c, v):
                out.append((s1, s2, tuple(qi), tuple(qo), int(a), int(b), float(x)))
        return out

    def contract_L(self, site, Lold, A):
        new = {}
        idx = {(qi, qo, qm): blocks for (qi, qo, qm, blocks) in Lold}
        for (s1, s2, qwi, qwo, a, b, w) in self._wnz(site):
            for (qin, qout, qmid), blocks in idx.items():
                if qmid != qwi or a >= len(blocks) or blocks[a] is None:
                    continue
                for (bi, bo), Ab in A[s1].items():
                    if bi != qin:
                        continue
                    for (ki, ko), Ak in A[s2].items():
                        if ki != qout:
                            continue
                        key = (bo, ko, qwo)
                        dim = dict(site.aux_out)[qwo]
                        slot = new.setdefault(key, [None] * dim)
                        m = w * (Ab.T @ blocks[a] @ Ak)
                        slot[b] = m if slot[b] is None else slot[b] + m
        return [(k[0], k[1], k[2], [None if x is None else np.asfortranarray(x) for x in v])
                for k, v in new.items()]

    def contract_R(self, site, Rold, A):
        new = {}
        idx = {(qi, qo, qm): blocks for (qi, qo, qm, blocks) in Rold}
        for (s1, s2, qwi, qwo, a, b, w) in self._wnz(site):
            for (qin, qout, qmid), blocks in idx.items():
                if qmid != qwo or b >= len(blocks) or blocks[b] is None:
                    continue
                for (ki, ko), Ak in A[s2].items():
                    if ko != qin:
                        continue
                    for (bi, bo), Ab in A[s1].items():
                        if bo != qout:
                            continue
                        key = (ki, bi, qwi)
                        dim = dict(site.aux_in)[qwi]
                        slot = new.setdefault(key, [None] * dim)
                        m = w * (Ak @ blocks[b] @ Ab.T)
                        slot[a] = m if slot[a] is None else slot[a] + m
        return [(k[0], k[1], k[2], [None if x is None else np.asfortranarray(x) for x in v])
                for k, v in new.items()]
