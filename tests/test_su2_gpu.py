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


@pytest.mark.parametrize("nsite,p", [(2, 2), (1, 3), (0, 3)])
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


@pytest.mark.parametrize("nsite,p", [(2, 1), (0, 2)])
@pytest.mark.parametrize("S2,N", [(1, 4), (2, 3)])
def test_hubbard_su2xu1_exact_pivot_spectrum_device(S2, N, nsite, p):
    print("hubbard su2xu1 exact pivot", S2, N, nsite, p, flush=True)
    L = 4
    mpo = models.chain_mpo("HubbardSU2xU1", L, t=1.0, U=8.0)
    ev = fb.exact_spectrum(models.SU2xU1, mpo, (S2, N), p, nsite, fb.device_contracts(models.SU2xU1),
                           fb.device_factory)
    ref = ed.hubbard_multiplet_spectrum(L, N, (S2 - 1) / 2)
    assert ev.shape == ref.shape
    assert np.allclose(ev, ref, atol=1e-10)


@pytest.mark.parametrize("S,nsite,p", [(1, 2, 2), (0, 2, 2), (1, 1, 3)])
def test_su2_penalty_term_device_matches_oracle(S, nsite, p):
    """Excited-state penalty Epenalty <A0|Vin> A0 on exact SU(2) chain pivots
    (the overlap weighted like the dot, pinned in test_projector_cpu): the
    device HxV with projectors equals the oracle's."""
    import oracle
    import pivots
    from test_projector_cpu import exact_pivot
    from vmps_amd import pivot
    print("su2 penalty", S, nsite, p, flush=True)
    pd, _ = exact_pivot(S, nsite, p)
    rng = np.random.default_rng(23)
    A0 = [rng.standard_normal(pd.vsize) for _ in range(2)]
    Ep = 3.5
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd, projectors=A0, Epenalty=Ep)
    cls = pivot.PivotMatrix2 if nsite == 2 else pivot.PivotMatrix1
    H = cls(pd, projectors=A0, Epenalty=Ep)
    for seed in (1, 2):
        v = np.random.default_rng(seed).standard_normal(pd.vsize)
        ref = pivots.oracle_hxv(lib, h, v)
        out = pivot.HxV(H, v)
        assert np.linalg.norm(out - ref) <= 1e-12 * np.linalg.norm(ref)


def _env_dict(E):
    return {(tuple(a), tuple(b), tuple(c)): blocks for (a, b, c, blocks) in E}


def _close(got, ref, tol):
    """Same sectors (absent and zero blocks alike), blocks within tol of the reference norm."""
    g, r = _env_dict(got), _env_dict(ref)
    scale = np.sqrt(sum(float(np.sum(b * b)) for bl in r.values() for b in bl if b is not None))
    err = 0.0
    for key in set(g) | set(r):
        gb, rb = g.get(key), r.get(key)
        n = len(gb) if gb is not None else len(rb)
        for a in range(n):
            x = gb[a] if gb is not None and a < len(gb) else None
            y = rb[a] if rb is not None and a < len(rb) else None
            if x is None and y is None:
                continue
            if x is None or y is None:
                z = x if y is None else y
                err += float(np.sum(z * z))
                continue
            assert x.shape == y.shape, key
            err += float(np.sum((x - y) ** 2))
    return np.sqrt(err) <= tol * scale, np.sqrt(err) / scale


@pytest.mark.parametrize("which", [0, 1])
@pytest.mark.parametrize("model", ["heis_su2", "c60_site", "hub_su2xu1"])
def test_su2_contract_random_matches_restatement(which, model):
    """contract_L / contract_R on random SU(2) tensors at M ~ 150-250
    multiplets (device, with the engine's coeff_buildL/R) against the CPU
    restatement of DmrgContractions.h:34-311 with the oracle's coefficients."""
    import envs
    from vmps_amd import env, lattice, synth
    rng = np.random.default_rng(40 + which)
    if model == "heis_su2":
        site, sym = models.heisenberg_su2_mpo(), models.SU2
        left, right = synth.heisenberg_su2_bond(14, 160), synth.heisenberg_su2_bond(15, 160)
    elif model == "c60_site":
        site = models.chain_mpo("HeisenbergSU2Full", 60, J=lattice.c60_couplings())[30]
        sym = models.SU2
        left, right = synth.heisenberg_su2_bond(30, 220), synth.heisenberg_su2_bond(31, 220)
    else:
        site, sym = models.hubbard_su2xu1_ladder_mpo(t=1.0, U=6.0, rung=True), models.SU2xU1
        left = synth.hubbard_su2xu1_bond(10, 150, filling=0.875)
        right = synth.hubbard_su2xu1_bond(11, 150, filling=0.875)
    print("su2 contract", "LR"[which], model, flush=True)
    Abra = envs.random_site_tensor(site, left, right, rng)
    Aket = envs.random_site_tensor(site, left, right, rng)
    co = fb.OracleCoeffs(sym)
    if which == 0:
        Eold = envs.random_env(site, left, rng, "L")
        got = env.contract_L(site, Eold, Abra, Aket)
        ref = fb.contract_L_ref(sym, site, Eold, Abra, Aket, co)
    else:
        Eold = envs.random_env(site, right, rng, "R")
        got = env.contract_R(site, Eold, Abra, Aket)
        ref = fb.contract_R_ref(sym, site, Eold, Abra, Aket, co)
    ok, rel = _close(got, ref, 1e-12)
    assert ok, rel
