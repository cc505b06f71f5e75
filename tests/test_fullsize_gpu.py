"""The workload shapes at their full sizes: against the oracle where it
finishes in seconds (SU(2)xU(1) ladder M=2048, long-range SU(2) D_W=40
M=1024), through size-independent properties where it does not (D_W=40
M=4096: linearity, independence of the tile configuration, agreement of the
mixed modes with FP64), and the 64-pivot batch against its pivots' own plans."""
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import synth
from vmps_amd.pivot import HxV, PivotBatch, PivotMatrix2

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import pivots  # noqa: E402

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12
FP32_RTOL = 1e-5
BF16X3_RTOL = 5e-5


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _oracle(pd, v):
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd)
    try:
        return pivots.oracle_hxv(lib, h, v, nthreads=16)
    finally:
        lib.oracle_pivot2_destroy(h)


@pytest.mark.parametrize("name", ["hubbard_su2xu1_ladder_L64_M2048", "heisenberg_su2_lr40_L60_M1024"])
def test_full_size_vs_oracle(name):
    pd = synth.make_workload(name)
    v = synth.random_vector(pd, 3)
    print("case", name, "vsize", pd.vsize, flush=True)
    ref = _oracle(pd, v)
    for precision, tol in (("fp64", FP64_RTOL), ("bf16x3", BF16X3_RTOL)):
        H = PivotMatrix2(pd, precision=precision)
        H.autotune(iters=2)
        rel = _rel(HxV(H, v), ref)
        print(precision, "tiles", H.tiles(), "rel err", rel, flush=True)
        assert rel <= tol
        H.close()


def test_dw40_m4096_properties():
    name = "heisenberg_su2_lr40_L60_M4096"
    pd = synth.make_workload(name)
    x = synth.random_vector(pd, 4)
    y = synth.random_vector(pd, 5)
    print("case", name, "vsize", pd.vsize, flush=True)
    H = PivotMatrix2(pd)
    hx, hy, hxy = HxV(H, x), HxV(H, y), HxV(H, x + 2.0 * y)
    assert _rel(hxy, hx + 2.0 * hy) <= 1e-13                 # linearity
    assert np.array_equal(HxV(H, x), hx)                      # deterministic
    for cfg in (dict(wave_edge=16, big_min64=128, explicit_u=1), dict(wave_edge=32, big_min64=0, explicit_u=0)):
        H.set_tiles(**cfg)
        rel = _rel(HxV(H, x), hx)
        print("fp64 tiles", cfg, "rel vs default", rel, flush=True)
        assert rel <= 1e-13                                   # same sums, other order
    H.close()
    for precision, tol in (("fp32", FP32_RTOL), ("bf16x3", BF16X3_RTOL)):
        Hm = PivotMatrix2(pd, precision=precision)
        Hm.autotune(iters=2)
        rel = _rel(HxV(Hm, x), hx)
        print(precision, "rel vs fp64", rel, flush=True)
        assert rel <= tol
        Hm.close()


def test_batch64_equals_single_plans():
    pds = synth.make_batch("batch64_u1_M128_2048")
    vs = [synth.random_vector(pd, 10 + i) for i, pd in enumerate(pds)]
    B = PivotBatch(pds)
    B.autotune(iters=2)
    out = B.split(HxV(B, np.concatenate(vs)))
    worst = 0.0
    for i, (pd, v) in enumerate(zip(pds, vs)):
        H = PivotMatrix2(pd)
        worst = max(worst, _rel(out[i], HxV(H, v)))
        H.close()
    print("batch vs single plans, worst rel", worst, flush=True)
    assert worst <= 1e-13
