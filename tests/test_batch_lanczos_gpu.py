"""One Lanczos solve per pivot of a batch plan (vmps_lanczos_batch), all
pivots iterating together in one captured HIP graph per iteration: every
pivot's eigenpair equals its own single-plan solve."""
import numpy as np
import pytest

from vmps_amd import _lib, models, synth
from vmps_amd.pivot import PivotBatch, PivotMatrix2, lanczos, lanczos_batch

pytestmark = pytest.mark.gpu


def _pivots(seed=11):
    """Symmetric pivots of different sizes and models (U(1) chains and SU(2))."""
    rng = np.random.default_rng(seed)
    out = []
    heis, hub = models.heisenberg_u1_mpo(), models.hubbard_u1xu1_mpo()
    for i, M in enumerate([12, 40, 24, 64, 18, 48]):
        if i % 2 == 0:
            n = int(rng.integers(6, 14))
            out.append(synth.make_pivot(heis, heis, synth.heisenberg_bond(n, M), synth.heisenberg_bond(n + 2, M), rng,
                                        hermitian=True))
        else:
            n = int(rng.integers(3, 7))
            out.append(synth.make_pivot(hub, hub, synth.hubbard_bond(n, M, alpha=1.0, lam=1.2),
                                        synth.hubbard_bond(n + 2, M, alpha=1.0, lam=1.2), rng, hermitian=True))
    return out


def _wdot(pd, x, y):
    w = np.concatenate([np.full(r * c, pd.sym.coeff_dot(qo)) for (_, _, qo, _, r, c) in pd.block_offsets()])
    return float(np.sum(w * x * y))


def test_batch_lanczos_matches_single_solves():
    pds = _pivots()
    B = PivotBatch(pds)
    v0 = np.random.default_rng(5).uniform(-1, 1, B.vsize)
    opts = dict(max_iter=200, tol_eigval=1e-13, tol_state=1e-8)
    vb, rb = lanczos_batch(B, v0=v0, **opts)
    parts = B.split(vb)
    starts = B.split(v0)
    for i, pd in enumerate(pds):
        H = PivotMatrix2(pd)
        vs, rs = lanczos(H, v0=starts[i], **opts)
        print("pivot", i, "vsize", pd.vsize, "batch", rb[i]["energy"], rb[i]["iterations"], "single", rs["energy"],
              rs["iterations"], flush=True)
        assert rb[i]["converged"]
        assert abs(rb[i]["energy"] - rs["energy"]) <= 1e-11 * max(1.0, abs(rs["energy"]))
        assert rb[i]["residual"] <= 1e-6
        ov = _wdot(pd, parts[i], vs)
        assert abs(abs(ov) - 1.0) <= 1e-8
        H.close()


def test_batch_lanczos_graph_equals_eager():
    pds = _pivots(seed=12)[:4]
    B = PivotBatch(pds)
    v0 = np.random.default_rng(6).uniform(-1, 1, B.vsize)
    va, ra = lanczos_batch(B, v0=v0, max_iter=120, tol_eigval=1e-12, tol_state=1e-7, graph=True)
    vb, rbb = lanczos_batch(B, v0=v0, max_iter=120, tol_eigval=1e-12, tol_state=1e-7, graph=False)
    assert np.array_equal(va, vb)
    assert [r["energy"] for r in ra] == [r["energy"] for r in rbb]


def test_batch_lanczos_bf16x3_refined():
    """split-BF16x3 batch solve refined in FP64: each pivot's energy within
    1e-9 of its FP64 solve."""
    pds = _pivots(seed=13)
    B = PivotBatch(pds)
    v0 = np.random.default_rng(7).uniform(-1, 1, B.vsize)
    _, r64 = lanczos_batch(B, v0=v0, max_iter=200, tol_eigval=1e-12, tol_state=1e-7)
    Bm = PivotBatch(pds, precision="bf16x3")
    _, rm = lanczos_batch(Bm, v0=v0, precision="bf16x3", max_iter=200, tol_eigval=1e-12, tol_state=1e-7,
                          refine_tol_eigval=1e-12, refine_tol_state=1e-7)
    for a, b in zip(r64, rm):
        print(a["energy"], b["energy"], b["iterations"], b["iterations_refine"], flush=True)
        assert abs(a["energy"] - b["energy"]) <= 1e-9 * max(1.0, abs(a["energy"]))


def test_single_and_batch_entry_points_refuse_each_others_plans():
    pds = _pivots()[:2]
    B = PivotBatch(pds)
    with pytest.raises(_lib.VmpsStatusError) as e:
        lanczos(B)
    assert e.value.status == 6
    H = PivotMatrix2(pds[0])
    import ctypes
    o = _lib.LanczosOpts()
    H._lib.vmps_lanczos_opts_default(ctypes.byref(o))
    res = (_lib.LanczosResult * 1)()
    v = np.ones(H.vsize)
    st = H._lib.vmps_lanczos_batch(H._plan, v.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(o), res)
    assert st == 6


def test_batch64_bf16x3_hxv_matches_oracle_subset():
    """The 64-pivot uneven batch in split-BF16x3 (one launch per stage) against
    the oracle on a subset of its pivots (relative error of BF16x3)."""
    import oracle
    import pivots
    from vmps_amd.pivot import HxV
    pds = synth.make_batch("batch64_u1_M128_2048")
    B = PivotBatch(pds, precision="bf16x3")
    v = np.random.default_rng(9).standard_normal(B.vsize)
    out = B.split(HxV(B, v))
    vin = B.split(v)
    lib = oracle.load()
    for i in (0, 7, 31, 50, 63):
        print("batch64 pivot", i, "vsize", pds[i].vsize, flush=True)
        h = pivots.build_oracle(lib, pds[i])
        ref = pivots.oracle_hxv(lib, h, vin[i], 16)
        lib.oracle_pivot2_destroy(h)
        assert np.linalg.norm(out[i] - ref) <= 5e-5 * np.linalg.norm(ref)
