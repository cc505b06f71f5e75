"""Device Lanczos (stands in for the external LanczosSolver) and the device
PivotVector algebra, on symmetric pivots."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import models, synth
from vmps_amd.pivot import PivotMatrix2, lanczos

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import dense  # noqa: E402
import pivots  # noqa: E402

pytestmark = pytest.mark.gpu


def _herm_small(kind, seed=3):
    if kind == "heis":
        W = models.heisenberg_u1_mpo()
        lb, rb = synth.heisenberg_bond(5, 10), synth.heisenberg_bond(7, 10)
    else:
        W = models.hubbard_u1xu1_mpo()
        lb = synth.hubbard_bond(3, 12, alpha=1.0, lam=1.2)
        rb = synth.hubbard_bond(5, 12, alpha=1.0, lam=1.2)
    return synth.make_pivot(W, W, lb, rb, np.random.default_rng(seed), hermitian=True)


@pytest.mark.parametrize("kind", ["heis", "hub"])
def test_lanczos_fp64_vs_dense_eigensolver(kind):
    pd = _herm_small(kind)
    n = pd.vsize
    Hm = np.stack([dense.dense_hxv(pd, np.eye(n)[i]) for i in range(n)], axis=1)
    evals, evecs = np.linalg.eigh(0.5 * (Hm + Hm.T))
    H = PivotMatrix2(pd)
    v, r = lanczos(H, max_iter=min(n, 120), tol_eigval=1e-13, tol_state=1e-9)
    print(kind, r, evals[:2])
    assert abs(r["energy"] - evals[0]) <= 1e-10 * max(1.0, abs(evals[0]))
    assert r["residual"] <= 1e-6
    assert abs(abs(v @ evecs[:, 0]) - 1.0) <= 1e-8  # same eigenvector (non-degenerate ground state)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_lanczos_mixed_refined_to_fp64(prec):
    pd = _herm_small("hub", seed=5)
    n = pd.vsize
    Hm = np.stack([dense.dense_hxv(pd, np.eye(n)[i]) for i in range(n)], axis=1)
    e0 = np.linalg.eigvalsh(0.5 * (Hm + Hm.T))[0]
    H = PivotMatrix2(pd, precision=prec)
    v, r = lanczos(H, precision=prec, max_iter=min(n, 120), tol_eigval=1e-6, tol_state=1e-4)
    print(r, e0)
    assert abs(r["energy_mixed"] - e0) <= 1e-4 * abs(e0)   # the mixed solve's vector, before the corrections
    assert abs(r["energy"] - e0) <= 1e-9 * max(1.0, abs(e0))  # after FP64 refinement


def test_lanczos_workload_vs_scipy_on_oracle():
    """Hubbard M=512 symmetric pivot: device FP64 Lanczos vs ARPACK (scipy) driving the oracle's HxV."""
    from scipy.sparse.linalg import LinearOperator, eigsh
    pd = synth.make_workload("hubbard_u1xu1_L48_M512", hermitian=True)
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    op = LinearOperator((pd.vsize, pd.vsize), matvec=lambda x: pivots.oracle_hxv(lib, h, np.asarray(x).ravel(), 16),
                        dtype=np.float64)
    e_ref = eigsh(op, k=1, which="SA", tol=1e-12, v0=np.ones(pd.vsize))[0][0]
    lib.oracle_pivot2_destroy(h)
    H = PivotMatrix2(pd)
    v, r = lanczos(H, max_iter=300, tol_eigval=1e-12, tol_state=1e-6)
    print(r, e_ref)
    assert abs(r["energy"] - e_ref) <= 1e-9 * abs(e_ref)
    # mixed precision + FP64 refinement reaches |dE| <= 1e-9 relative as well
    for prec in ("fp32", "bf16x3"):
        Hm = PivotMatrix2(pd, precision=prec)
        v2, r2 = lanczos(Hm, precision=prec, max_iter=300, tol_eigval=1e-7, tol_state=1e-4)
        print(prec, r2)
        assert abs(r2["energy"] - e_ref) <= 1e-9 * abs(e_ref)


def test_vector_algebra_device():
    from vmps_amd.pivot import DeviceVector
    pd = _herm_small("hub")
    H = PivotMatrix2(pd)
    lib, plan = H._lib, H._plan
    xh, yh = synth.random_vector(pd, 1), synth.random_vector(pd, 2)
    x, y = DeviceVector(H, xh), DeviceVector(H, yh)
    r = ctypes.c_double()
    assert lib.vmps_vec_dot(plan, x.ptr, y.ptr, ctypes.byref(r)) == 0
    assert abs(r.value - xh @ yh) <= 1e-12 * np.abs(xh).sum() * np.abs(yh).max()
    assert lib.vmps_vec_axpby(plan, 0.5, x.ptr, -2.0, y.ptr) == 0
    assert np.abs(y.download() - (0.5 * xh - 2.0 * yh)).max() <= 1e-13
    assert lib.vmps_vec_norm(plan, x.ptr, ctypes.byref(r)) == 0
    assert abs(r.value - np.linalg.norm(xh)) <= 1e-12 * np.linalg.norm(xh)
    assert lib.vmps_vec_fill_random(plan, x.ptr, 42) == 0
    xr = x.download()
    assert abs(np.linalg.norm(xr) - 1.0) < 1e-12 and np.abs(xr).max() < 1.0
    # the same seed gives the same vector
    assert lib.vmps_vec_fill_random(plan, y.ptr, 42) == 0
    assert np.array_equal(xr, y.download())


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
def test_lanczos_invariant_start_vector(precision):
    """A start vector that spans an invariant subspace (here the whole
    one-dimensional space of the two-spin pivot with both spins up) ends the
    iteration by breakdown, also before min_iter, with a finite exact energy."""
    import fullbasis as fb
    pd = fb.chain_pivot(models.U1, models.chain_mpo("HeisenbergU1", 2), (2,), 0, 2, None, None)
    assert pd.vsize == 1
    H = PivotMatrix2(pd)
    v, r = lanczos(H, v0=np.ones(1), precision=precision)
    print(precision, r)
    assert np.isfinite(r["energy"]) and abs(r["energy"] - 0.25) <= 1e-12
    assert np.all(np.isfinite(v))


@pytest.mark.parametrize("kind", ["heis", "hub"])
def test_graph_iteration_matches_eager(kind):
    """Each iteration replayed from one captured HIP graph gives exactly the
    result of the same kernels launched one by one, over repeated solves on one
    plan (graph reuse), a mixed solve with FP64 refinement (one graph per
    precision) and a change of tiles (the graph is captured again)."""
    pd = _herm_small(kind)
    H = PivotMatrix2(pd)
    runs = [dict(precision="fp64"), dict(precision="fp64", seed=7), dict(precision="bf16x3"),
            dict(precision="fp64")]
    for i, kw in enumerate(runs):
        if i == 3:
            H.set_tiles(32 if H.tiles()["wave_edge"] == 16 else 16, H.tiles()["big_min"])
        print(kind, "run", i, kw, flush=True)
        vg, rg = lanczos(H, max_iter=80, tol_eigval=1e-12, tol_state=1e-8, graph=True, **kw)
        ve, re_ = lanczos(H, max_iter=80, tol_eigval=1e-12, tol_state=1e-8, graph=False, **kw)
        assert rg["iterations"] == re_["iterations"] and rg["iterations_refine"] == re_["iterations_refine"]
        assert rg["energy"] == re_["energy"]
        assert np.array_equal(vg, ve)


@pytest.mark.parametrize("prec", ["fp32", "bf16x3"])
def test_mixed_refinement_correction_vs_restart(prec):
    """The correction refinement (FP64 residual, mixed CG correction, FP64
    Rayleigh-Ritz) and the FP64 restart reach the FP64 eigenvalue of the
    M=512 Hubbard pivot; the correction needs at most four FP64 HxV."""
    pd = synth.make_workload("hubbard_u1xu1_L48_M512", hermitian=True)
    H = PivotMatrix2(pd)
    v0 = np.random.default_rng(3).uniform(-1, 1, pd.vsize)
    _, r64 = lanczos(H, v0=v0, max_iter=300, tol_eigval=1e-12, tol_state=1e-7)
    H.set_precision(prec)
    out = {}
    for meth in ("correction", "restart"):
        v, r = lanczos(H, v0=v0, precision=prec, max_iter=300, tol_eigval=1e-12, tol_state=1e-7,
                       refine_tol_eigval=1e-12, refine_tol_state=1e-7, refine_method=meth)
        print(prec, meth, r, flush=True)
        assert r["converged"]
        assert abs(r["energy"] - r64["energy"]) <= 1e-10 * abs(r64["energy"])
        assert r["residual"] <= 1e-6
        out[meth] = r
    assert out["correction"]["hxv_fp64"] <= 4 < out["restart"]["hxv_fp64"]


def test_zero_start_vector_refused():
    from vmps_amd import _lib
    pd = _herm_small("heis")
    H = PivotMatrix2(pd)
    with pytest.raises(_lib.VmpsStatusError) as e:
        lanczos(H, v0=np.zeros(pd.vsize))
    assert "zero" in str(e.value)


def test_partial_reorthogonalisation_keeps_full_accuracy():
    """FP64 solves re-orthogonalise partially (omega recurrence): on the
    Hermitian M=512 Hubbard pivot, 60 iterations from a random start give the
    Ritz value of the fully re-orthogonalised basis (the lowest eigenvalue of
    the tridiagonal matrix of a numpy run with full re-orthogonalisation on
    the oracle's HxV) as the Rayleigh quotient of the returned vector, and
    no fallback to full re-orthogonalisation (one HxV per iteration plus the
    residual's)."""
    pd = synth.make_workload("hubbard_u1xu1_L48_M512", hermitian=True)
    H = PivotMatrix2(pd)
    v0 = np.random.default_rng(1).uniform(-1, 1, pd.vsize)
    _, r = lanczos(H, v0=v0, max_iter=60, tol_eigval=0.0, tol_state=0.0)
    H.close()
    # reference: the same 60 iterations with full re-orthogonalisation on the oracle's HxV
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    w = np.concatenate([np.full(rr * c, pd.sym.coeff_dot(qo)) for (_, _, qo, _, rr, c) in pd.block_offsets()])
    sw = np.sqrt(w)
    try:
        Q = [v0 * sw / np.linalg.norm(v0 * sw)]
        al, be = [], []
        for k in range(60):
            x = pivots.oracle_hxv(lib, h, Q[-1] / sw, 16) * sw
            for p in range(2):
                c = np.array(Q) @ x
                x = x - np.array(Q).T @ c
                al.append(c[-1]) if p == 0 else al.__setitem__(-1, al[-1] + c[-1])
            be.append(np.linalg.norm(x))
            Q.append(x / be[-1])
    finally:
        lib.oracle_pivot2_destroy(h)
    T = np.diag(al) + np.diag(be[:-1], 1) + np.diag(be[:-1], -1)
    e_ref = np.linalg.eigvalsh(T)[0]
    print("partial re-orthogonalisation: energy %.14f reference %.14f hxv %d" % (
        r["energy"], e_ref, r["hxv_calls"]), flush=True)
    assert abs(r["energy"] - e_ref) <= 1e-11 * abs(e_ref)
    assert r["hxv_calls"] <= r["iterations"] + 4  # no second, fully re-orthogonalised solve


def test_partial_reorthogonalisation_falls_back_on_a_nonsymmetric_operator():
    """A random-environment pivot that is not symmetric breaks the omega
    recurrence's assumption; the Ritz value then disagrees with the Rayleigh
    quotient and the solve runs again with full re-orthogonalisation (about
    twice the HxV); graph and eager runs agree bit for bit."""
    pd = synth.make_workload("hubbard_u1xu1_L48_M512", hermitian=False)
    H = PivotMatrix2(pd)
    v0 = np.random.default_rng(1).uniform(-1, 1, pd.vsize)
    _, r1 = lanczos(H, v0=v0, max_iter=60, tol_eigval=0.0, tol_state=0.0)
    _, r2 = lanczos(H, v0=v0, max_iter=60, tol_eigval=0.0, tol_state=0.0, graph=False)
    H.close()
    print("non-symmetric pivot: hxv %d / %d iterations, energy %.10f / %.10f" % (
        r1["hxv_calls"], r1["iterations"], r1["energy"], r2["energy"]), flush=True)
    assert r1["hxv_calls"] >= 2 * r1["iterations"]
    assert r1["energy"] == r2["energy"] and r1["residual"] == r2["residual"]
