"""CPU backend of the DMRG driver, built from the oracle (test infrastructure):
the oracle's HxV with scipy's ARPACK as eigensolver, and the restatement of
contract_L / contract_R in fullbasis.py (tensors/DmrgContractions.h:34-311;
SU(2) factors from the oracle's coefficient functions).

For SU(2) the effective Hamiltonian is self-adjoint under the coeff_dot
weighted inner product (pivot/DmrgPivotVector.h:296-339), so ARPACK works on
D^1/2 H D^-1/2 with D the per-element weights."""
import ctypes

import numpy as np
from scipy.sparse.linalg import LinearOperator, eigsh

import fullbasis
import oracle
from vmps_amd import dmrg, synth


def element_weights(pd):
    return np.concatenate([np.full(r * c, pd.sym.coeff_dot(qo)) for (_, _, qo, _, r, c) in pd.block_offsets()]
                          or [np.zeros(0)])


class OracleBackend:
    def __init__(self, tol=1e-13):
        self.lib = oracle.load()
        self.tol = tol
        self._coeffs = {}

    def solve(self, pd, v0):
        lib = self.lib
        h = synth.feed(lib, "oracle_", pd, oracle.check)
        oracle.check(lib.oracle_pivot2_precalc(h))
        P = ctypes.POINTER(ctypes.c_double)
        w = element_weights(pd)
        sw = np.sqrt(w)

        def mv(y):
            x = np.ascontiguousarray(np.asarray(y, dtype=np.float64).ravel() / sw)
            out = np.zeros_like(x)
            oracle.check(lib.oracle_hxv_omp(h, x.ctypes.data_as(P), out.ctypes.data_as(P), 8))
            return out * sw

        n = pd.vsize
        try:
            if n <= 3:
                Hm = np.stack([mv(np.eye(n)[i]) for i in range(n)], axis=1)
                ev, V = np.linalg.eigh(0.5 * (Hm + Hm.T))
                E, y = ev[0], V[:, 0]
            else:
                op = LinearOperator((n, n), matvec=mv, dtype=np.float64)
                y0 = v0 * sw if np.any(v0) else None
                ev, V = eigsh(op, k=1, which="SA", tol=self.tol, v0=y0)
                E, y = ev[0], V[:, 0]
        finally:
            lib.oracle_pivot2_destroy(h)
        v = y / sw
        return float(E), v / np.sqrt(np.sum(w * v * v)), {}

    def _co(self, sym):
        if sym.kinds not in self._coeffs:
            self._coeffs[sym.kinds] = fullbasis.OracleCoeffs(sym)
        return self._coeffs[sym.kinds]

    def contract_L(self, site, Lold, A):
        blocks = dmrg._mps_to_blocks(A)
        return fullbasis.contract_L_ref(site.sym, site, Lold, blocks, blocks, self._co(site.sym))

    def contract_R(self, site, Rold, A):
        blocks = dmrg._mps_to_blocks(A)
        return fullbasis.contract_R_ref(site.sym, site, Rold, blocks, blocks, self._co(site.sym))
