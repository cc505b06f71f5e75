"""Environments resident on the device (vmps_denv): a contraction or a pivot
fed from device environments gives bit for bit what the same call gives when
fed from host memory (the packed operands are the same numbers, gathered on
the device instead of the host), and the layout survives the round trip."""
import copy
import ctypes
import os
import sys

import numpy as np
import pytest

from vmps_amd import _lib, env, models, synth
from vmps_amd.pivot import HxV, PivotBatch, PivotMatrix0, PivotMatrix1, PivotMatrix2

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import pivots  # noqa: E402
from envs import random_env, random_site_tensor  # noqa: E402
from test_su2_cpu import su2_pivot, su2_pivot1  # noqa: E402

KINDS = dict(pivots.ALL_KINDS)
KINDS["terms"] = lambda: pivots.with_terms(pivots.small_hubbard())
KINDS["su2_2"] = su2_pivot
KINDS["su2_1"] = su2_pivot1
KINDS["su2_lr40"] = lambda: synth.make_workload("heisenberg_su2_lr40_L60_M1024", M=96)

pytestmark = pytest.mark.gpu

MATRIX = {0: PivotMatrix0, 1: PivotMatrix1, 2: PivotMatrix2}


def _norm(sectors):
    out = []
    for qin, qout, qmid, blocks in sectors:
        out.append((tuple(qin), tuple(qout), tuple(qmid),
                    [None if b is None or b.size == 0 else np.asarray(b) for b in blocks]))
    return out


def _same_env(a, b):
    a, b = _norm(a), _norm(b)
    assert len(a) == len(b)
    for (qi, qo, qm, ba), (qi2, qo2, qm2, bb) in zip(a, b):
        assert (qi, qo, qm) == (qi2, qo2, qm2)
        assert len(ba) == len(bb)
        for x, y in zip(ba, bb):
            assert (x is None) == (y is None)
            if x is not None:
                assert x.shape == y.shape and np.array_equal(x, y)


def _on_device(pd):
    """pd with every environment replaced by a DeviceEnv of the same blocks."""
    nq = pd.sym.nq
    dd = copy.copy(pd)
    dd.L = env.DeviceEnv.from_host(pd.L, nq)
    dd.R = env.DeviceEnv.from_host(pd.R, nq)
    dd.extra = [synth.PivotTerm(t.W12, t.W34, env.DeviceEnv.from_host(t.L, nq), env.DeviceEnv.from_host(t.R, nq))
                for t in pd.extra]
    return dd


def test_round_trip():
    pd = pivots.small_hubbard()
    print("case round trip", flush=True)
    d = env.DeviceEnv.from_host(pd.L, pd.sym.nq)
    _same_env(d.to_host(), pd.L)
    d.close()


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("kind", ["heis2", "hub2", "hub1", "heis0", "terms", "su2_2", "su2_1", "su2_lr40"])
def test_pivot_device_envs_match_host(kind, precision):
    pd = KINDS[kind]()
    print("case", kind, precision, "vsize", pd.vsize, flush=True)
    cls = MATRIX[pd.nsite]
    v = synth.random_vector(pd, 3)
    H = cls(pd, precision=precision)
    ref = HxV(H, v)
    H.close()
    Hd = cls(_on_device(pd), precision=precision)
    got = HxV(Hd, v)
    Hd.close()
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("precision", ["fp64", "bf16x3"])
@pytest.mark.parametrize("model", ["hub", "su2"])
def test_contract_device_matches_host(model, precision):
    rng = np.random.default_rng(11)
    if model == "hub":
        site = models.hubbard_u1xu1_mpo()
        left = synth.hubbard_bond(3, 20, alpha=1.0, lam=1.2)
        right = synth.hubbard_bond(4, 20, alpha=1.0, lam=1.2)
    else:  # SU(2): the contraction carries the recoupling factors
        site = models.heisenberg_su2_mpo()
        left, right = synth.heisenberg_su2_bond(5, 16), synth.heisenberg_su2_bond(6, 16)
    Abra = random_site_tensor(site, left, right, rng)
    Aket = random_site_tensor(site, left, right, rng)
    nq = site.sym.nq
    print("case contract", model, precision, flush=True)
    Lold = random_env(site, left, rng, "L")
    host = env.contract_L(site, Lold, Abra, Aket, precision=precision)
    for old in (Lold, env.DeviceEnv.from_host(Lold, nq)):
        _same_env(env.contract_L(site, old, Abra, Aket, precision=precision, device=True).to_host(), host)
        _same_env(env.contract_L(site, old, Abra, Aket, precision=precision), host)
    Rold = random_env(site, right, rng, "R")
    host = env.contract_R(site, Rold, Abra, Aket, precision=precision)
    dold = env.DeviceEnv.from_host(Rold, nq)
    _same_env(env.contract_R(site, dold, Abra, Aket, precision=precision, device=True).to_host(), host)


def test_sweep_chain_stays_on_device():
    """L grows one site on the device and feeds a pivot without a host copy:
    same HxV as the host-fed chain."""
    rng = np.random.default_rng(12)
    site = models.heisenberg_u1_mpo()
    b0, b1 = synth.heisenberg_bond(4, 12), synth.heisenberg_bond(5, 12)
    A = random_site_tensor(site, b0, b1, rng)
    L0 = random_env(site, b0, rng, "L")
    print("case chain", flush=True)
    L1_host = env.contract_L(site, L0, A, A)
    L1_dev = env.contract_L(site, env.DeviceEnv.from_host(L0, site.sym.nq), A, A, device=True)
    _same_env(L1_dev.to_host(), L1_host)
    pd = synth.make_pivot(site, site, b1, synth.heisenberg_bond(7, 12), rng)
    if not any(blk is not None for s in L1_host for blk in s[3]):
        pytest.skip("empty environment")
    # the pivot's own random L is replaced by the contracted one (same bond)
    pd.L = L1_host
    v = synth.random_vector(pd, 5)
    try:
        H = PivotMatrix2(pd)
    except _lib.VmpsStatusError:
        pytest.skip("contracted environment does not match the synthetic pivot's mid sectors")
    ref = HxV(H, v)
    H.close()
    dd = copy.copy(pd)
    dd.L = L1_dev
    Hd = PivotMatrix2(dd)
    assert np.array_equal(HxV(Hd, v), ref)
    Hd.close()


def test_queued_updates_feed_the_next_update_and_a_pivot():
    """A device update returns with its kernels queued; the next update that
    reads it and a pivot plan that gathers it, made at once (no download in
    between), wait for it: the same results as the host-fed chain."""
    rng = np.random.default_rng(13)
    site = models.heisenberg_u1_mpo()
    b0, b1, b2 = synth.heisenberg_bond(4, 12), synth.heisenberg_bond(5, 12), synth.heisenberg_bond(6, 12)
    A1 = random_site_tensor(site, b0, b1, rng)
    A2 = random_site_tensor(site, b1, b2, rng)
    L0 = random_env(site, b0, rng, "L")
    print("case queued chain", flush=True)
    L1_host = env.contract_L(site, L0, A1, A1)
    L2_host = env.contract_L(site, L1_host, A2, A2)
    L1_dev = env.contract_L(site, env.DeviceEnv.from_host(L0, site.sym.nq), A1, A1, device=True)
    L2_dev = env.contract_L(site, L1_dev, A2, A2, device=True)
    _same_env(L2_dev.to_host(), L2_host)
    if not any(blk is not None for s in L1_host for blk in s[3]):
        pytest.skip("empty environment")
    pd = synth.make_pivot(site, site, b1, synth.heisenberg_bond(7, 12), rng)
    pd.L = L1_host
    v = synth.random_vector(pd, 5)
    try:
        H = PivotMatrix2(pd)
    except _lib.VmpsStatusError:
        pytest.skip("contracted environment does not match the synthetic pivot's mid sectors")
    ref = HxV(H, v)
    H.close()
    dd = copy.copy(pd)
    dd.L = env.contract_L(site, env.DeviceEnv.from_host(L0, site.sym.nq), A1, A1, device=True)  # still queued
    Hd = PivotMatrix2(dd)
    assert np.array_equal(HxV(Hd, v), ref)
    Hd.close()


def test_errors():
    lib = _lib.load()
    pd = pivots.small_heisenberg()
    dL = env.DeviceEnv.from_host(pd.L, pd.sym.nq)
    # a batch plan takes host environments only
    dd = copy.copy(pd)
    dd.L = dL
    with pytest.raises(_lib.VmpsStatusError):
        PivotBatch([pd, dd])
    # one environment given twice, or host sectors after a device one
    h = synth.feed(lib, "vmps_", pd, _lib.check)
    try:
        st = lib.vmps_pivot2_add_env_device(h, 0, dL.handle)
        assert st == 6  # VMPS_ERR_STATE: L already has sectors
    finally:
        lib.vmps_pivot2_destroy(h)
    # the result of a device contraction is not on the host
    site = models.heisenberg_u1_mpo()
    rng = np.random.default_rng(13)
    b0, b1 = synth.heisenberg_bond(4, 10), synth.heisenberg_bond(5, 10)
    A = random_site_tensor(site, b0, b1, rng)
    L0 = random_env(site, b0, rng, "L")
    d = env.contract_L(site, L0, A, A, device=True)
    n = [ctypes.c_int64() for _ in range(3)]
    assert lib.vmps_denv_size(d.handle, *[ctypes.byref(x) for x in n]) == 0
    d.close()
