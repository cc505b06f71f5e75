"""CPU checks that pin the oracle and the host-side block structure."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import _lib, synth
from vmps_amd.pivot import get_blockstructure

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import dense  # noqa: E402
import pivots  # noqa: E402


@pytest.mark.parametrize("make", [pivots.small_heisenberg, pivots.small_hubbard])
@pytest.mark.parametrize("threads", [0, 4])
def test_oracle_matches_dense_heff(make, threads):
    """Oracle OxV == brute-force dense L.W12.W34.R contraction (abelian: all CGC = 1)."""
    pd = make()
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    for seed in (1, 2):
        v = synth.random_vector(pd, seed)
        out = pivots.oracle_hxv(lib, h, v, nthreads=threads)
        ref = dense.dense_hxv(pd, v)
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    lib.oracle_pivot2_destroy(h)


def test_oracle_plants_are_caught():
    """A mistake planted in the environment (one L block transposed within a
    square sector) changes the oracle's result: the dense check is sensitive."""
    pd = pivots.small_heisenberg()
    v = synth.random_vector(pd, 3)
    ref = dense.dense_hxv(pd, v)
    for i, (qin, qout, qmid, blocks) in enumerate(pd.L):
        if blocks[0].shape[0] == blocks[0].shape[1] and blocks[0].shape[0] > 1:
            blocks = [np.asfortranarray(blocks[0].T)] + blocks[1:]
            pd.L[i] = (qin, qout, qmid, blocks)
            break
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    out = pivots.oracle_hxv(lib, h, v)
    lib.oracle_pivot2_destroy(h)
    assert np.linalg.norm(out - ref) > 1e-6 * np.linalg.norm(ref)


@pytest.mark.parametrize("make", [pivots.small_heisenberg, pivots.small_hubbard,
                                  lambda: synth.make_workload("heisenberg_u1_L32_M64")])
def test_engine_blockstructure_equals_oracle(make):
    """Host-side precalc_blockStructure of the engine == oracle's, entry for entry."""
    pd = make()
    olib = oracle.load()
    vlib = _lib.load()
    ho = pivots.build_oracle(olib, pd)
    hv = synth.feed(vlib, "vmps_", pd, _lib.check)
    _lib.check(vlib.vmps_pivot2_precalc(hv))
    bo = get_blockstructure(olib, ho, prefix="oracle_")
    bv = get_blockstructure(vlib, hv)
    for a, b in zip(bo, bv):
        assert np.array_equal(a, b)
    olib.oracle_pivot2_destroy(ho)
    vlib.vmps_pivot2_destroy(hv)


def test_local_basis_agrees():
    """Combined two-site basis: Python restatement, engine and oracle agree."""
    from vmps_amd import models
    olib = oracle.load()
    vlib = _lib.load()
    for W in (models.heisenberg_u1_mpo(), models.hubbard_u1xu1_mpo()):
        pd = synth.make_pivot(W, W, [((0,) * W.sym.nq, 1)], [((0,) * W.sym.nq, 1)], np.random.default_rng(0))
        ho = synth.feed(olib, "oracle_", pd, oracle.check)
        hv = synth.feed(vlib, "vmps_", pd, _lib.check)
        res = []
        for lib, h, pre in ((olib, ho, "oracle_"), (vlib, hv, "vmps_")):
            ns = ctypes.c_int()
            q = np.zeros(64, dtype=np.int32)
            getattr(lib, pre + "pivot2_local_basis")(h, ctypes.byref(ns), q.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
            res.append(q[:ns.value * W.sym.nq].reshape(ns.value, -1))
        py = np.array([b[0] for b in pd.basis])
        assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], py)
        # U(1): states sorted by total charge
        assert all(tuple(py[i]) <= tuple(py[i + 1]) for i in range(len(py) - 1))
        olib.oracle_pivot2_destroy(ho)
        vlib.vmps_pivot2_destroy(hv)


@pytest.mark.parametrize("kind", ["heis1", "hub1", "heis0", "hub0"])
@pytest.mark.parametrize("threads", [0, 3])
def test_oracle_one_and_zero_site_match_dense(kind, threads):
    """One-site OxV (DmrgPivotMatrix1.h) and zero-site HxV (DmrgPivotMatrix0.h)
    of the oracle == dense contraction."""
    pd = pivots.ALL_KINDS[kind]()
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    for seed in (1, 2):
        v = synth.random_vector(pd, seed)
        out = pivots.oracle_hxv(lib, h, v, nthreads=threads)
        ref = dense.dense_hxv(pd, v)
        assert np.linalg.norm(ref) > 0
        assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)
    lib.oracle_pivot2_destroy(h)


@pytest.mark.parametrize("kind", ["heis2", "hub1"])
def test_oracle_sum_of_terms_matches_dense(kind):
    """A pivot with three terms: the sum of the oracle's per-term results ==
    the dense sum over terms."""
    pd = pivots.with_terms(pivots.ALL_KINDS[kind]())
    v = synth.random_vector(pd, 4)
    out = pivots.oracle_hxv_terms(oracle.load(), pd, v)
    ref = dense.dense_hxv(pd, v)
    assert np.linalg.norm(out - ref) <= 1e-13 * np.linalg.norm(ref)


@pytest.mark.parametrize("kind", ["heis1", "hub1"])
def test_engine_one_site_blockstructure_equals_oracle(kind):
    pd = pivots.ALL_KINDS[kind]()
    olib = oracle.load()
    vlib = _lib.load()
    ho = pivots.build_oracle(olib, pd)
    hv = synth.feed(vlib, "vmps_", pd, _lib.check)
    _lib.check(vlib.vmps_pivot2_precalc(hv))
    bo = get_blockstructure(olib, ho, prefix="oracle_")
    bv = get_blockstructure(vlib, hv)
    assert len(bo[0]) > 0
    for a, b in zip(bo, bv):
        assert np.array_equal(a, b)
    olib.oracle_pivot2_destroy(ho)
    vlib.vmps_pivot2_destroy(hv)


def test_engine_pivot_terms_and_kinds():
    """Builder bookkeeping of the engine: pivot kind and term count."""
    vlib = _lib.load()
    for kind, ns in (("heis2", 2), ("hub1", 1), ("hub0", 0)):
        pd = pivots.ALL_KINDS[kind]()
        if ns:
            pivots.with_terms(pd, 2)
        hv = synth.feed(vlib, "vmps_", pd, _lib.check)
        nsite, nterms = ctypes.c_int(), ctypes.c_int()
        _lib.check(vlib.vmps_pivot_info(hv, ctypes.byref(nsite), ctypes.byref(nterms)))
        assert (nsite.value, nterms.value) == (ns, 3 if ns else 1)
        assert vlib.vmps_pivot_select_term(hv, nterms.value) == 1  # VMPS_ERR_ARG
        _lib.check(vlib.vmps_pivot2_precalc(hv))
        vlib.vmps_pivot2_destroy(hv)
