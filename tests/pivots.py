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


def small_heisenberg1(seed=5, M=12, nleft=5):
    W = models.heisenberg_u1_mpo(J=1.0)
    rng = np.random.default_rng(seed)
    return synth.make_pivot1(W, synth.heisenberg_bond(nleft, M), synth.heisenberg_bond(nleft + 1, M), rng)


def small_hubbard1(seed=6, M=16, nleft=4):
    W = models.hubbard_u1xu1_mpo(t=1.0, U=8.0)
    rng = np.random.default_rng(seed)
    return synth.make_pivot1(W, synth.hubbard_bond(nleft, M, alpha=1.0, lam=1.2),
                             synth.hubbard_bond(nleft + 1, M, alpha=1.0, lam=1.2), rng)


def small_heisenberg0(seed=7, M=14, nleft=6):
    W = models.heisenberg_u1_mpo(J=1.0)
    return synth.make_pivot0(W, synth.heisenberg_bond(nleft, M), np.random.default_rng(seed))


def small_hubbard0(seed=8, M=20, nleft=4):
    W = models.hubbard_u1xu1_mpo(t=1.0, U=8.0)
    return synth.make_pivot0(W, synth.hubbard_bond(nleft, M, alpha=1.0, lam=1.2), np.random.default_rng(seed))


def with_terms(pd, n=2, seed=9):
    """pd plus n more random summands (two-site terms alternate Heisenberg J)."""
    rng = np.random.default_rng(seed)
    for _ in range(n):
        synth.add_random_term(pd, rng)
    return pd


ALL_KINDS = {
    "heis2": small_heisenberg, "hub2": small_hubbard,
    "heis1": small_heisenberg1, "hub1": small_hubbard1,
    "heis0": small_heisenberg0, "hub0": small_hubbard0,
}


def oracle_hxv_terms(lib, pd, vin):
    """Oracle result of a pivot with several terms: the reference sums the
    per-term OxV results (DmrgPivotMatrix2.h:233-245); the oracle takes one
    term per pivot, so each term is built and applied on its own."""
    import oracle
    out = np.zeros(pd.vsize)
    for t in range(1 + len(pd.extra)):
        h = synth.feed(lib, "oracle_", pd, oracle.check, terms=[t])
        oracle.check(lib.oracle_pivot2_precalc(h))
        out += oracle_hxv(lib, h, vin)
        lib.oracle_pivot2_destroy(h)
    return out
