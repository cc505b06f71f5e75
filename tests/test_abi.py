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
    assert lib.vmps_abi_version() == 2


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
