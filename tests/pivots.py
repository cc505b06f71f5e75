"""Small pivots shared by the tests."""
import numpy as np

from vmps_amd import models, synth


def small_heisenberg(seed=1, M=12, nleft=5):
    W = models.heisenberg_u1_mpo(J=1.0)
    rng = np.random.default_rng(seed)
    return synth.make_pivot(W, W, synth.heisenberg_bond(nleft, M), synth.heisenberg_bond(nleft + 2, M), rng)


def small_hubbard(seed=2, M=16, nleft=4, U=8.0):
    W = models.hubbard_u1xu1_mpo(t=1.0, U=U)
    rng = np.random.default_rng(seed)
    left = synth.hubbard_bond(nleft, M, alpha=1.0, lam=1.2)
    right = synth.hubbard_bond(nleft + 2, M, alpha=1.0, lam=1.2)
    return synth.make_pivot(W, W, left, right, rng)


def build_oracle(lib, pd, precalc=True, projectors=None, Epenalty=0.0):
    import oracle
    h = synth.feed(lib, "oracle_", pd, oracle.check, projectors=projectors, Epenalty=Epenalty)
    if precalc:
        oracle.check(lib.oracle_pivot2_precalc(h))
    return h


def oracle_hxv(lib, h, vin, nthreads=0):
    import ctypes
    import oracle
    vin = np.ascontiguousarray(vin, dtype=np.float64)
    out = np.zeros_like(vin)
    P = ctypes.POINTER(ctypes.c_double)
    if nthreads:
        oracle.check(lib.oracle_hxv_omp(h, vin.ctypes.data_as(P), out.ctypes.data_as(P), nthreads))
    else:
        oracle.check(lib.oracle_hxv(h, vin.ctypes.data_as(P), out.ctypes.data_as(P)))
    return out
