"""The C60 Heisenberg model (HeisenbergSU2 with the full coupling matrix of
the truncated icosahedron, Cuthill-McKee chain order): HxV of real two-site
pivots of a DMRG state against the oracle in the three precisions, at a
warm-up bond dimension and after the bonds are expanded, and the mixed solve
against the FP64 solve of the same pivot."""
import numpy as np
import pytest

from vmps_amd import dmrg, lattice, models, synth
from vmps_amd.pivot import HxV, PivotMatrix2, lanczos

pytestmark = pytest.mark.gpu

TOL = {"fp64": 1e-12, "fp32": 1e-5, "bf16x3": 5e-5}


@pytest.fixture(scope="module")
def c60_chains():
    """A state from DMRG sweeps at M = 32 and 64 centred on pivot (30, 31),
    and the same state with its bonds expanded towards M = 512 (host
    environments, so that the oracle can read them)."""
    sites = models.chain_mpo("HeisenbergSU2Full", 60, J=lattice.c60_couplings())
    be = dmrg.GpuBackend(precision="fp64", device_envs=False, tol_eigval=1e-10, tol_state=1e-6)
    warm = dmrg.warm_chain(sites, (1,), [32, 64], 30, be, np.random.default_rng(3))
    big = dmrg.expand_chain(warm, 512, 30, be, np.random.default_rng(4))
    return warm, big


def _check_against_oracle(pd, seed):
    import oracle
    import pivots
    v = synth.random_vector(pd, seed)
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    ref = pivots.oracle_hxv(lib, h, v, 16)
    lib.oracle_pivot2_destroy(h)
    for prec, tol in TOL.items():
        H = PivotMatrix2(pd, precision=prec)
        out = HxV(H, v)
        H.close()
        err = np.linalg.norm(out - ref) / np.linalg.norm(ref)
        print("C60 pivot vsize", pd.vsize, prec, "relative error %.2e" % err, flush=True)
        assert err <= tol


def test_c60_warm_pivot_hxv_matches_oracle(c60_chains):
    warm, _ = c60_chains
    _check_against_oracle(warm.pivot(30), 5)


def test_c60_expanded_pivot_hxv_matches_oracle(c60_chains):
    _, big = c60_chains
    _check_against_oracle(big.pivot(30), 6)


def test_c60_mixed_solve_matches_fp64(c60_chains):
    """Split-BF16x3 Lanczos refined in FP64 against FP64 Lanczos on the
    expanded pivot, from the same start vector (the state's two-site vector):
    the same eigenvalue to 1e-9 and the same eigenvector."""
    _, big = c60_chains
    pd = big.pivot(30)
    v0 = dmrg.two_site_psi(pd, big.A[30], big.A[31], big.qloc)
    opts = dict(max_iter=200, tol_eigval=1e-11, tol_state=1e-7)
    H = PivotMatrix2(pd, precision="fp64")
    v64, r64 = lanczos(H, v0=v0, **opts)
    H.set_precision("bf16x3")
    vm, rm = lanczos(H, v0=v0, precision="bf16x3", refine_tol_eigval=1e-11, refine_tol_state=1e-7, **opts)
    H.close()
    print("C60 M=512 pivot: E fp64 %.14f (%d HxV), mixed %.14f (%d HxV, %d FP64)" % (
        r64["energy"], r64["hxv_calls"], rm["energy"], rm["hxv_calls"], rm["hxv_fp64"]), flush=True)
    assert abs(rm["energy"] - r64["energy"]) <= 1e-9
    assert rm["residual"] <= 1e-7
    w = np.concatenate([np.full(r * c, pd.sym.coeff_dot(qo)) for (_, _, qo, _, r, c) in pd.block_offsets()])
    assert abs(abs(np.sum(w * vm * v64)) - 1.0) <= 1e-8
