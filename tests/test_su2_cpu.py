"""SU(2) on the CPU: exact 6j/9j fixtures pin the oracle's Racah evaluation;
full-basis chain pivots pin the coupling coefficients and reduced MPO
conventions against exact diagonalisation; the engine's SU(2) block
structure equals the oracle's."""
import json
import os
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import _lib, models, synth
from vmps_amd.pivot import get_blockstructure

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import ed  # noqa: E402
import fullbasis as fb  # noqa: E402
import pivots  # noqa: E402


def heis_su2_chain(L, J=1.0):
    return [models.heisenberg_su2_mpo(J=J, position="first")] + \
           [models.heisenberg_su2_mpo(J=J) for _ in range(L - 2)] + [models.heisenberg_su2_mpo(J=J, position="last")]


def heis_u1_chain(L, J=1.0):
    return models.chain_mpo("HeisenbergU1", L, J=J)


@pytest.mark.parametrize("nsite,p", [(2, 2), (2, 0), (1, 3), (0, 3)])
@pytest.mark.parametrize("two_sz", [0, 2, 4])
def test_fullbasis_u1_spectrum_equals_ed(nsite, p, two_sz):
    """Harness check with U(1): the exact pivot's generalized spectrum is the
    Sz sector's spectrum."""
    L = 6
    ev = fb.exact_spectrum(models.U1, heis_u1_chain(L), (two_sz,), p, nsite, fb.oracle_contracts(models.U1),
                           fb.oracle_factory)
    ref = ed.heisenberg_sector_spectrum(L, two_sz)
    assert ev.shape == ref.shape
    assert np.allclose(ev, ref, atol=1e-10)


@pytest.mark.parametrize("nsite,p", [(2, 2), (2, 0), (2, 4), (1, 2), (0, 3), (0, 2)])
@pytest.mark.parametrize("S", [0, 1, 2])
def test_fullbasis_su2_spectrum_equals_ed(nsite, p, S):
    """SU(2)-reduced HeisenbergSU2 pivot (oracle HxV + restated contract_L/R
    with the oracle's coeff_buildL/R): one eigenvalue per spin-S multiplet,
    equal to exact diagonalisation of the S=1/2 chain."""
    L = 6
    ev = fb.exact_spectrum(models.SU2, heis_su2_chain(L), (2 * S + 1,), p, nsite, fb.oracle_contracts(models.SU2),
                           fb.oracle_factory)
    ref = ed.heisenberg_multiplet_spectrum(L, S)
    assert ev.shape == ref.shape, (ev, ref)
    assert np.allclose(ev, ref, atol=1e-10), (ev, ref)


def test_oracle_wigner_matches_exact():
    """Racah 6j / 9j of the oracle == sympy's exact symbols (2j up to 60 / 24)."""
    with open(os.path.join(HERE, "golden", "wigner.json")) as f:
        fx = json.load(f)
    lib = oracle.load()
    for key, fn in (("6j", lib.oracle_wigner6j), ("9j", lib.oracle_wigner9j)):
        for (j, _, val) in fx[key]:
            v = float(val)
            got = fn(*j)
            assert abs(got - v) <= 1e-13 * max(abs(v), 1e-3), (key, j, got, v)
    # planted mistake: a transposed argument pair changes the value
    j = fx["6j"][0][0]
    assert abs(lib.oracle_wigner6j(j[1], j[0], j[2], j[4], j[3], j[5] + 2) - float(fx["6j"][0][2])) > 1e-6


def su2_pivot(seed=3, M=18, nleft=9, J=1.0):
    W = models.heisenberg_su2_mpo(J=J)
    rng = np.random.default_rng(seed)
    return synth.make_pivot(W, W, synth.heisenberg_su2_bond(nleft, M), synth.heisenberg_su2_bond(nleft + 2, M), rng)


def su2_pivot1(seed=4, M=18, nleft=9):
    W = models.heisenberg_su2_mpo()
    rng = np.random.default_rng(seed)
    return synth.make_pivot1(W, synth.heisenberg_su2_bond(nleft, M), synth.heisenberg_su2_bond(nleft + 1, M), rng)


@pytest.mark.parametrize("make", [su2_pivot, su2_pivot1, lambda: su2_pivot(nleft=24, M=40)])
def test_engine_su2_blockstructure_equals_oracle(make):
    """SU(2) precalc of the engine (6j/9j factors) == the oracle's."""
    pd = make()
    olib, vlib = oracle.load(), _lib.load()
    ho = pivots.build_oracle(olib, pd)
    hv = synth.feed(vlib, "vmps_", pd, _lib.check)
    _lib.check(vlib.vmps_pivot2_precalc(hv))
    bo = get_blockstructure(olib, ho, prefix="oracle_")
    bv = get_blockstructure(vlib, hv)
    olib.oracle_pivot2_destroy(ho)
    vlib.vmps_pivot2_destroy(hv)
    assert len(bo[0]) > 0
    for a, b in zip(bo[:3], bv[:3]):
        assert np.array_equal(a, b)
    assert np.allclose(bo[3], bv[3], rtol=1e-13, atol=1e-15)
    # non-trivial factors occur
    assert np.any(np.abs(np.abs(bo[3]) - 1.0) > 1e-3)


@pytest.mark.parametrize("nsite,p", [(2, 1), (1, 2), (0, 2)])
@pytest.mark.parametrize("S2,N", [(1, 4), (3, 4), (2, 3), (2, 5), (1, 2)])
def test_fullbasis_hubbard_su2xu1_spectrum_equals_ed(nsite, p, S2, N):
    """HubbardSU2xU1 (reduced c, c^dagger with the fermionic sign, U d):
    exact-pivot spectrum at (2S+1, N) == Hubbard ED multiplets, L = 4."""
    L = 4
    mpo = models.chain_mpo("HubbardSU2xU1", L, t=1.0, U=8.0)
    ev = fb.exact_spectrum(models.SU2xU1, mpo, (S2, N), p, nsite, fb.oracle_contracts(models.SU2xU1),
                           fb.oracle_factory)
    ref = ed.hubbard_multiplet_spectrum(L, N, (S2 - 1) / 2)
    assert ev.shape == ref.shape
    assert np.allclose(ev, ref, atol=1e-10)
