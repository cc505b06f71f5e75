"""Tile configurations and the autotuner: every configuration gives the
oracle's HxV (FP64: 1e-12 relative; split-BF16 x3: 5e-5), the autotuner picks
one of its candidates and caches it per block-size histogram."""
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import synth
from vmps_amd.pivot import HxV, PivotMatrix2, clear_tile_cache

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import pivots  # noqa: E402

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
BF16X3_RTOL = 5e-5


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


@pytest.fixture(scope="module")
def case():
    # blocks up to ~150, so that the 128 x 128 split-BF16 tiles are used for some
    pd = pivots.small_heisenberg(seed=3, M=400, nleft=6)
    v = synth.random_vector(pd, 4)
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    try:
        ref = pivots.oracle_hxv(lib, h, v, nthreads=8)
    finally:
        lib.oracle_pivot2_destroy(h)
    return pd, v, ref


@pytest.mark.parametrize("edge", [16, 32])
def test_fp64_wave_edges(case, edge):
    pd, v, ref = case
    print("fp64 wave edge", edge, "vsize", pd.vsize, flush=True)
    H = PivotMatrix2(pd)
    H.set_tiles(edge, H.tiles()["big_min"])
    assert H.tiles()["wave_edge"] == edge
    assert _rel(HxV(H, v), ref) <= FP64_RTOL


@pytest.mark.parametrize("big_min", [96, 64, 0])
def test_bf16x3_big_thresholds(case, big_min):
    pd, v, ref = case
    print("bf16x3 big_min", big_min, flush=True)
    H = PivotMatrix2(pd, precision="bf16x3")
    H.set_tiles(H.tiles()["wave_edge"], big_min)
    assert H.tiles()["big_min"] == big_min
    assert _rel(HxV(H, v), ref) <= BF16X3_RTOL


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_RTOL), ("fp32", 1e-5), ("bf16x3", BF16X3_RTOL)])
@pytest.mark.parametrize("which", ["heis", "lr40"])
def test_explicit_stage_b(case, precision, tol, which):
    """U = sum c T formed by its own kernel, stage C reading U: same result as
    the fused combine (and the oracle), on the D_W = 5 chain and on the
    long-range D_W = 40 SU(2) pivot with many terms per output block."""
    if which == "heis":
        pd, v, ref = case
    else:
        pd = synth.make_workload("heisenberg_su2_lr40_L60_M1024", M=96)
        v = synth.random_vector(pd, 8)
        lib = oracle.load()
        h = pivots.build_oracle(lib, pd)
        try:
            ref = pivots.oracle_hxv(lib, h, v, nthreads=8)
        finally:
            lib.oracle_pivot2_destroy(h)
    print("explicit stage B", which, precision, flush=True)
    H = PivotMatrix2(pd, precision=precision)
    fused = HxV(H, v)
    H.set_tiles(explicit_u=1)
    assert H.tiles()["explicit_u"] == 1
    out = HxV(H, v)
    assert _rel(out, ref) <= tol
    assert _rel(out, fused) <= (1e-15 if precision == "fp64" else 2 * tol)
    H.set_tiles(explicit_u=0)  # and back
    assert _rel(HxV(H, v), fused) == 0.0


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_RTOL), ("fp32", 1e-5), ("bf16x3", BF16X3_RTOL)])
def test_autotune_picks_candidate_and_caches(case, precision, tol):
    pd, v, ref = case
    print("autotune", precision, flush=True)
    H = PivotMatrix2(pd, precision=precision)
    cfg = H.autotune(iters=3)
    if cfg["from_cache"] == 0:
        assert cfg["n_tried"] >= 2 and cfg["ms"] > 0
    if precision == "bf16x3":
        assert cfg["big_min"] in (96, 192, 0)
    else:
        assert cfg["wave_edge"] in (16, 32)
    assert _rel(HxV(H, v), ref) <= tol
    # a second plan of the same shape takes the cached choice without timing
    H2 = PivotMatrix2(pd, precision=precision)
    cfg2 = H2.autotune(iters=3)
    assert cfg2["from_cache"] == 1
    assert (cfg2["wave_edge"], cfg2["big_min"], cfg2["explicit_u"]) == (cfg["wave_edge"], cfg["big_min"],
                                                                      cfg["explicit_u"])
    assert _rel(HxV(H2, v), ref) <= tol
    # after the cache is cleared, the next plan of that shape is timed again
    clear_tile_cache()
    H3 = PivotMatrix2(pd, precision=precision)
    cfg3 = H3.autotune(iters=3)
    assert cfg3["from_cache"] == 0 and cfg3["n_tried"] >= 2
    assert _rel(HxV(H3, v), ref) <= tol


def test_set_tiles_rejects_bad_edge(case):
    pd, _, _ = case
    H = PivotMatrix2(pd)
    with pytest.raises(Exception):
        H.set_tiles(24, 96)


@pytest.mark.parametrize("precision,tol", [("fp64", FP64_RTOL), ("fp32", 1e-5)])
@pytest.mark.parametrize("big_min64,xu", [(96, 0), (64, 1), (128, 1)])
def test_fp64_fp32_big_tiles(case, precision, tol, big_min64, xu):
    """LDS-staged 128 x 128 FP64 / FP32 tiles for the problems >= big_min64
    (with the fused and with the explicit stage B) against the oracle; FP64 is
    bit-identical to the wave tiles only up to summation order."""
    pd, v, ref = case
    print("big tiles", precision, big_min64, xu, flush=True)
    H = PivotMatrix2(pd, precision=precision)
    H.set_tiles(big_min64=big_min64, explicit_u=xu)
    cfg = H.tiles()
    assert cfg["big_min64"] == big_min64 and cfg["explicit_u"] == xu
    assert _rel(HxV(H, v), ref) <= tol
