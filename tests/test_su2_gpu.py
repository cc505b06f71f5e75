"""SU(2) on the device: HxV of SU(2)-reduced pivots against the oracle, the
coeff_dot-weighted inner product, and the exact-pivot spectrum of the
HeisenbergSU2 chain (device contract_L/R + device HxV) against exact
diagonalisation."""
import ctypes
import os
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import models, synth
from vmps_amd.pivot import HxV, PivotMatrix1, PivotMatrix2, dot

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ed  # noqa: E402
import fullbasis as fb  # noqa: E402
import pivots  # noqa: E402
from test_su2_cpu import heis_su2_chain, su2_pivot, su2_pivot1  # noqa: E402

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle(pd, v):
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    try:
        return pivots.oracle_hxv(lib, h, v, nthreads=8)
    finally:
        lib.oracle_pivot2_destroy(h)


@pytest.mark.parametrize("make", [su2_pivot, su2_pivot1, lambda: su2_pivot(nleft=24, M=40)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("bf16x3", 5e-5)])
def test_su2_hxv_matches_oracle(make, precision, tol):
    pd = make()
    print("su2 pivot nsite", pd.nsite, "vsize", pd.vsize, precision, flush=True)
    H = (PivotMatrix2 if pd.nsite == 2 else PivotMatrix1)(pd, precision=precision)
    for seed in (1, 2):
        v = synth.random_vector(pd, seed)
        assert _rel(HxV(H, v), _oracle(pd, v)) <= tol


def test_su2_weighted_dot():
    pd = su2_pivot()
    H = PivotMatrix2(pd)
    x, y = synth.random_vector(pd, 5), synth.random_vector(pd, 6)
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    r = ctypes.c_double()
    P = ctypes.POINTER(ctypes.c_double)
    oracle.check(lib.oracle_dot(h, x.ctypes.data_as(P), y.ctypes.data_as(P), ctypes.byref(r)))
    lib.oracle_pivot2_destroy(h)
    w = np.concatenate([np.full(r_ * c_, float(qo[0])) for (_, _, qo, _, r_, c_) in pd.block_offsets()])
    assert abs(dot(H, x, y) - r.value) <= 1e-12 * abs(r.value)
    assert abs(r.value - float(np.sum(w * x * y))) <= 1e-12 * abs(r.value)


@pytest.mark.parametrize("nsite,p", [(2, 2), (1, 3)])
@pytest.mark.parametrize("S", [0, 1, 2])
def test_su2_exact_pivot_spectrum_device(nsite, p, S):
    """Environments from the device contract_L/R, H_eff and N_eff from the
    device HxV: the generalized spectrum is the spin-S multiplet spectrum."""
    print("su2 exact pivot", nsite, p, S, flush=True)
    L = 6
    ev = fb.exact_spectrum(models.SU2, heis_su2_chain(L), (2 * S + 1,), p, nsite, fb.device_contracts(models.SU2),
                           fb.device_factory)
    ref = ed.heisenberg_multiplet_spectrum(L, S)
    assert ev.shape == ref.shape
    assert np.allclose(ev, ref, atol=1e-10)


@pytest.mark.parametrize("S2,N", [(1, 4), (2, 3)])
def test_hubbard_su2xu1_exact_pivot_spectrum_device(S2, N):
    print("hubbard su2xu1 exact pivot", S2, N, flush=True)
    L = 4
    mpo = models.chain_mpo("HubbardSU2xU1", L, t=1.0, U=8.0)
    ev = fb.exact_spectrum(models.SU2xU1, mpo, (S2, N), 1, 2, fb.device_contracts(models.SU2xU1),
                           fb.device_factory)
    ref = ed.hubbard_multiplet_spectrum(L, N, (S2 - 1) / 2)
    assert ev.shape == ref.shape
    assert np.allclose(ev, ref, atol=1e-10)
