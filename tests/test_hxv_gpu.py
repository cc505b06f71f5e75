"""GPU parity of the HIP HxV against the CPU oracle (FP64 drop-in: 1e-12
relative; FP32 mixed: 1e-5 relative), through the C ABI."""
import ctypes
import sys

import numpy as np
import pytest

import oracle
from vmps_amd import _lib, synth
from vmps_amd.pivot import HxV, PivotMatrix2, dot

sys.path.insert(0, __file__.rsplit("/", 1)[0])
import pivots  # noqa: E402

pytestmark = pytest.mark.gpu

FP64_RTOL = 1e-12   # ||Vout_gpu - Vout_oracle||_2 / ||Vout_oracle||_2
FP32_RTOL = 1e-5
BF16X3_RTOL = 5e-5  # hi/lo split keeps ~16 mantissa bits per operand


def _rel(a, b):
    return np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300)


def _oracle_out(pd, vin, nthreads=0, projectors=None, Epenalty=0.0):
    lib = oracle.load()
    h = pivots.build_oracle(lib, pd, projectors=projectors, Epenalty=Epenalty)
    try:
        return pivots.oracle_hxv(lib, h, vin, nthreads=nthreads)
    finally:
        lib.oracle_pivot2_destroy(h)


@pytest.fixture(scope="module")
def heis():
    return pivots.small_heisenberg()


@pytest.fixture(scope="module")
def hub():
    return pivots.small_hubbard()


@pytest.mark.parametrize("which", ["heis", "hub"])
def test_fp64_small_matches_oracle(which, request):
    pd = request.getfixturevalue(which)
    print("case", which, "vsize", pd.vsize, flush=True)
    H = PivotMatrix2(pd)
    for seed in (1, 2, 3):
        v = synth.random_vector(pd, seed)
        out = HxV(H, v)
        ref = _oracle_out(pd, v)
        assert _rel(out, ref) <= FP64_RTOL


@pytest.mark.parametrize("name", ["heisenberg_u1_L32_M64", "hubbard_u1xu1_L48_M512"])
def test_fp64_workloads_match_oracle(name):
    pd = synth.make_workload(name)
    print("workload", name, "vsize", pd.vsize, flush=True)
    H = PivotMatrix2(pd)
    v = synth.random_vector(pd, 11)
    out = HxV(H, v)
    ref = _oracle_out(pd, v, nthreads=16)
    rel = _rel(out, ref)
    print("rel err", rel)
    assert rel <= FP64_RTOL
    # elementwise: every block element within 1e-12 of the block scale
    assert np.max(np.abs(out - ref)) <= FP64_RTOL * np.max(np.abs(ref))


def test_fp64_deterministic_and_linear(hub):
    H = PivotMatrix2(hub)
    x = synth.random_vector(hub, 5)
    y = synth.random_vector(hub, 6)
    hx1, hx2 = HxV(H, x), HxV(H, x)
    assert np.array_equal(hx1, hx2)  # owner-computes, no atomics
    hy = HxV(H, y)
    hz = HxV(H, 0.75 * x - 1.5 * y)
    assert _rel(hz, 0.75 * hx1 - 1.5 * hy) <= FP64_RTOL


def test_dot_weighted(hub):
    H = PivotMatrix2(hub)
    x = synth.random_vector(hub, 7)
    y = synth.random_vector(hub, 8)
    lib = oracle.load()
    h = pivots.build_oracle(lib, hub)
    r = ctypes.c_double()
    P = ctypes.POINTER(ctypes.c_double)
    oracle.check(lib.oracle_dot(h, x.ctypes.data_as(P), y.ctypes.data_as(P), ctypes.byref(r)))
    lib.oracle_pivot2_destroy(h)
    assert abs(dot(H, x, y) - r.value) <= 1e-12 * abs(r.value)
    assert abs(dot(H, x, y) - float(x @ y)) <= 1e-12 * abs(float(x @ y))  # U(1): all weights 1


def test_blockstructure_passthrough(hub):
    """The caller's own qlhs/qrhs/factor_cgcs tables drive the device plan."""
    lib = oracle.load()
    h = pivots.build_oracle(lib, hub)
    from vmps_amd.pivot import get_blockstructure
    bs = get_blockstructure(lib, h, prefix="oracle_")
    lib.oracle_pivot2_destroy(h)
    H = PivotMatrix2(hub, blockstructure=bs)
    v = synth.random_vector(hub, 9)
    assert _rel(HxV(H, v), _oracle_out(hub, v)) <= FP64_RTOL


def test_excited_state_penalty(hub):
    rng = np.random.default_rng(3)
    A0 = [rng.standard_normal(hub.vsize) for _ in range(2)]
    H = PivotMatrix2(hub, projectors=A0, Epenalty=4.5)
    v = synth.random_vector(hub, 10)
    ref = _oracle_out(hub, v, projectors=A0, Epenalty=4.5)
    assert _rel(HxV(H, v), ref) <= FP64_RTOL


def test_fp32_mixed_error(hub):
    H = PivotMatrix2(hub, precision="fp32")
    v = synth.random_vector(hub, 12)
    rel = _rel(HxV(H, v), _oracle_out(hub, v))
    print("fp32 rel err", rel)
    assert rel <= FP32_RTOL


@pytest.mark.parametrize("name", [None, "hubbard_u1xu1_L48_M512", "heisenberg_u1_L64_M2048"])
def test_bf16x3_mixed_error(hub, name):
    pd = hub if name is None else synth.make_workload(name)
    H = PivotMatrix2(pd, precision="bf16x3")
    v = synth.random_vector(pd, 12)
    ref = _oracle_out(pd, v, nthreads=16)
    out = HxV(H, v)
    rel = _rel(out, ref)
    print("bf16x3 rel err", rel)
    assert rel <= BF16X3_RTOL
    # switching precision on the same plan gives the FP64 drop-in again
    H.set_precision("fp64")
    assert _rel(HxV(H, v), ref) <= FP64_RTOL


def test_nonfinite_input_rejected(hub):
    H = PivotMatrix2(hub)
    v = synth.random_vector(hub, 13)
    v[7] = np.nan
    with pytest.raises(_lib.VmpsStatusError) as e:
        HxV(H, v)
    assert e.value.status == 5


def test_stats_consistent(hub):
    H = PivotMatrix2(hub)
    st = H.stats()
    assert st["flops_alg"] == pytest.approx(st["flops_A"] + st["flops_C"])
    assert st["flops_alg"] <= st["flops_ref"]
    lib = oracle.load()
    h = pivots.build_oracle(lib, hub)
    f = ctypes.c_double()
    oracle.check(lib.oracle_flops_ref(h, ctypes.byref(f)))
    lib.oracle_pivot2_destroy(h)
    assert st["flops_ref"] == pytest.approx(f.value, rel=1e-12)


_CLASS = {0: "PivotMatrix0", 1: "PivotMatrix1", 2: "PivotMatrix2"}


def _matrix(pd, **kw):
    from vmps_amd import pivot
    return getattr(pivot, _CLASS[pd.nsite])(pd, **kw)


@pytest.mark.parametrize("kind", ["heis1", "hub1", "heis0", "hub0"])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_RTOL), ("bf16x3", BF16X3_RTOL)])
def test_one_and_zero_site_match_oracle(kind, precision, tol):
    """PivotMatrix1 / PivotMatrix0 HxV on the device vs the oracle's literal
    one-site OxV and zero-site HxV."""
    pd = pivots.ALL_KINDS[kind]()
    print("case", kind, precision, "vsize", pd.vsize, flush=True)
    H = _matrix(pd, precision=precision)
    for seed in (1, 2):
        v = synth.random_vector(pd, seed)
        assert _rel(HxV(H, v), _oracle_out(pd, v)) <= tol


@pytest.mark.parametrize("kind", ["heis2", "hub2", "hub1"])
def test_sum_of_terms_matches_oracle(kind):
    """Pivots with three terms (PivotMatrix2::Terms / PivotMatrix1::Terms):
    device result == sum of the oracle's per-term results."""
    pd = pivots.with_terms(pivots.ALL_KINDS[kind]())
    print("case", kind, "terms", 1 + len(pd.extra), flush=True)
    H = _matrix(pd)
    v = synth.random_vector(pd, 3)
    ref = pivots.oracle_hxv_terms(oracle.load(), pd, v)
    assert _rel(HxV(H, v), ref) <= FP64_RTOL


@pytest.mark.parametrize("name,M", [("hubbard_su2xu1_ladder_L64_M2048", 256), ("heisenberg_su2_lr40_L60_M1024", 96)])
@pytest.mark.parametrize("precision,tol", [("fp64", FP64_RTOL), ("bf16x3", BF16X3_RTOL)])
def test_su2_workloads_match_oracle(name, M, precision, tol):
    """The SU(2) workload shapes (snake ladder SU(2) x U(1), D_W = 6; long-range
    SU(2) Heisenberg, D_W = 40) at reduced bond dimension against the oracle."""
    pd = synth.make_workload(name, M=M)
    print("case", name, "M", M, precision, "vsize", pd.vsize, flush=True)
    H = PivotMatrix2(pd, precision=precision)
    v = synth.random_vector(pd, 21)
    rel = _rel(HxV(H, v), _oracle_out(pd, v, nthreads=16))
    print("rel err", rel)
    assert rel <= tol
