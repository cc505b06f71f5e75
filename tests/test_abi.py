"""The C ABI library loads, exports every declared symbol, and reports errors
as status codes (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

from vmps_amd import _lib, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if fn.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", fn)).read()
            names |= set(re.findall(r"\b(vmps_[a-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_all_declared_symbols():
    lib = _lib.load()
    names = declared_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert lib.vmps_abi_version() == 4


def test_bad_arguments_give_status_and_message():
    lib = _lib.load()
    h = ctypes.c_void_p()
    kinds = (ctypes.c_int * 1)(7)
    q = (ctypes.c_int * 2)(0, 0)
    st = lib.vmps_pivot2_create(1, kinds, 1, q, 1, q, 1, q, 1, q, ctypes.byref(h))
    assert st == 1 and b"symmetry" in lib.vmps_last_error()
    st = lib.vmps_pivot2_create(3, kinds, 1, q, 1, q, 1, q, 1, q, ctypes.byref(h))
    assert st == 1


def test_layout_errors():
    lib = _lib.load()
    from vmps_amd import models
    W = models.heisenberg_u1_mpo()
    pd = synth.make_pivot(W, W, [((1,), 2)], [((1,), 2)], np.random.default_rng(0))
    h = synth.feed(lib, "vmps_", pd, _lib.check)
    q = (ctypes.c_int * 1)(1)
    off = ctypes.c_int64()
    # duplicate block in one component
    blocks0 = pd.V[0]
    if blocks0:
        qi = (ctypes.c_int * 1)(*blocks0[0][0])
        qo = (ctypes.c_int * 1)(*blocks0[0][1])
        assert lib.vmps_pivot2_add_V_block(h, 0, qi, qo, 2, 2, ctypes.byref(off)) == 2
    # component out of range
    assert lib.vmps_pivot2_add_V_block(h, 99, q, q, 1, 1, ctypes.byref(off)) == 1
    # block structure read before precalc
    nl, nr = ctypes.c_int64(), ctypes.c_int64()
    assert lib.vmps_pivot2_get_blockstructure(h, ctypes.byref(nl), ctypes.byref(nr), None, None, None, None) == 6
    # bulk forms: a sector already fed is a duplicate, a bad W index an argument error,
    # with the status of the failing item
    from vmps_amd import pack
    bk = pack.tripod(pd.L[:1], 1)
    assert lib.vmps_pivot2_add_env_sectors(h, 0, bk.count, *bk.args()) == 2
    assert b"duplicate" in lib.vmps_last_error()
    s1, s2, k, qi, qo, rows, cols, r, c, v = W.blocks[0]
    bw = pack.w_blocks([(s1, s2, 99, qi, qo, rows, cols, r, c, v)], 1)
    assert lib.vmps_pivot2_add_W_blocks(h, 12, bw.count, *bw.args()) == 1
    comp = next(i for i, blocks in enumerate(pd.V) if blocks)
    bv = pack.v_blocks([[] for _ in range(comp)] + [pd.V[comp][:1]], 1)
    assert lib.vmps_pivot2_add_V_blocks(h, bv.count, *bv.args(), None) == 2
    # device environments: argument checks come before any device call
    assert lib.vmps_pivot2_add_env_device(h, 0, None) == 1
    lib.vmps_pivot2_destroy(h)
    d = ctypes.c_void_p()
    bk = pack.tripod(pd.L, 1)
    assert lib.vmps_denv_create(1, bk.count, *bk.args(), None) == 1
    assert lib.vmps_denv_create(0, bk.count, *bk.args(), ctypes.byref(d)) == 1
    assert not d.value
    e = ctypes.c_void_p()
    kinds = (ctypes.c_int * 1)(1)  # U(1)
    _lib.check(lib.vmps_env_create(0, 1, kinds, 1, q, 1, q, ctypes.byref(e)))
    assert lib.vmps_env_contract_device(e, 0, None) == 1
    assert lib.vmps_env_add_old_device(e, None) == 1
    lib.vmps_env_destroy(e)


def test_bulk_builders_equal_single_item_calls():
    """The engine's bulk feeding and the one-call-per-item feeding give the
    same vector layout and block structure."""
    lib = _lib.load()
    from vmps_amd.pivot import get_blockstructure
    from vmps_amd import models
    W = models.heisenberg_u1_mpo()
    rng = np.random.default_rng(3)
    pd = synth.make_pivot(W, W, [((1,), 3), ((-1,), 2), ((3,), 1)], [((1,), 2), ((-1,), 3), ((-3,), 1)], rng)
    hb = synth.feed(lib, "vmps_", pd, _lib.check)

    class Single:  # the same library seen under a prefix that selects the per-item path
        def __getattr__(self, name):
            return getattr(lib, name.replace("single_", "vmps_", 1))

    hs = synth.feed(Single(), "single_", pd, _lib.check)
    for h in (hb, hs):
        _lib.check(lib.vmps_pivot2_precalc(h))
    sizes = []
    for h in (hb, hs):
        n = ctypes.c_int64()
        _lib.check(lib.vmps_pivot2_vsize(h, ctypes.byref(n)))
        sizes.append(n.value)
    assert sizes[0] == sizes[1] == pd.vsize
    bb, bs = get_blockstructure(lib, hb), get_blockstructure(lib, hs)
    for x, y in zip(bb, bs):
        assert np.array_equal(np.asarray(x), np.asarray(y))
    for h in (hb, hs):
        lib.vmps_pivot2_destroy(h)


def test_set_blockstructure_rejects_bad_tables():
    lib = _lib.load()
    from vmps_amd import models
    W = models.heisenberg_u1_mpo()
    pd = synth.make_pivot(W, W, [((1,), 2)], [((1,), 2)], np.random.default_rng(0))
    h = synth.feed(lib, "vmps_", pd, _lib.check)
    qlhs = (ctypes.c_int64 * 2)(0, 0)
    rb = (ctypes.c_int64 * 2)(1, 0)  # not monotone
    qrhs = (ctypes.c_int64 * 12)()
    fac = (ctypes.c_double * 1)(1.0)
    assert lib.vmps_pivot2_set_blockstructure(h, 1, qlhs, rb, qrhs, fac) == 1
    lib.vmps_pivot2_destroy(h)
