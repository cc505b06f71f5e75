"""vmps_block_svd_noalloc: the block SVD of the two-site split with its work
arrays in the library's fixed device scratch (no device allocation), checked
against vmps_block_svd on the same inputs, twice in a row with the second
call smaller, and its size limit refused on the host."""
import ctypes

import numpy as np
import pytest

from vmps_amd import _lib

P_dbl = ctypes.POINTER(ctypes.c_double)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_int = ctypes.POINTER(ctypes.c_int)


def _blocks(rng, shapes):
    return [rng.standard_normal(s) for s in shapes]


def _call(fn, mats):
    rows = np.array([m.shape[0] for m in mats], dtype=np.int64)
    cols = np.array([m.shape[1] for m in mats], dtype=np.int64)
    k = np.minimum(rows, cols)
    A = np.concatenate([m.ravel(order="F") for m in mats] + [np.zeros(1)])
    U = np.full(int((rows * k).sum()) + 1, np.nan)
    S = np.full(int(k.sum()) + 1, np.nan)
    Vt = np.full(int((k * cols).sum()) + 1, np.nan)
    sw = ctypes.c_int(-1)
    st = fn(len(mats), rows.ctypes.data_as(P_i64), cols.ctypes.data_as(P_i64), A.ctypes.data_as(P_dbl),
            U.ctypes.data_as(P_dbl), S.ctypes.data_as(P_dbl), Vt.ctypes.data_as(P_dbl), ctypes.byref(sw))
    return st, U[:-1], S[:-1], Vt[:-1], sw.value


def _compare(mats):
    lib = _lib.load()
    st_a, Ua, Sa, Va, swa = _call(lib.vmps_block_svd, mats)
    st_b, Ub, Sb, Vb, swb = _call(lib.vmps_block_svd_noalloc, mats)
    assert st_a == 0 and st_b == 0, (st_a, st_b, lib.vmps_last_error())
    assert swa == swb
    # same Jacobi kernel and the same arithmetic in the final scaling: the
    # factors agree to rounding (bit-identical in practice)
    for a, b in ((Ua, Ub), (Sa, Sb), (Va, Vb)):
        assert np.all(np.isfinite(b))
        assert np.max(np.abs(a - b)) <= 1e-14 * max(1.0, np.max(np.abs(a)))
    # and they are an SVD of every block
    uo = so = vo = 0
    for m in mats:
        r, c = m.shape
        k = min(r, c)
        u = Ub[uo:uo + r * k].reshape((r, k), order="F")
        s = Sb[so:so + k]
        vt = Vb[vo:vo + k * c].reshape((k, c), order="F")
        assert np.all(np.diff(s) <= 0)
        assert np.allclose(s, np.linalg.svd(m, compute_uv=False), rtol=1e-12, atol=1e-13)
        assert np.linalg.norm(u @ (s[:, None] * vt) - m) <= 1e-12 * max(1.0, np.linalg.norm(m))
        uo, so, vo = uo + r * k, so + k, vo + k * c


@pytest.mark.gpu
def test_noalloc_block_svd_matches_block_svd_twice():
    rng = np.random.default_rng(11)
    first = _blocks(rng, [(48, 48), (60, 17), (9, 40), (1, 5), (0, 3), (130, 96)])
    print("case noalloc first call: %d blocks" % len(first), flush=True)
    _compare(first)
    second = _blocks(rng, [(20, 7), (5, 5), (3, 11)])
    print("case noalloc second call: %d blocks" % len(second), flush=True)
    _compare(second)


def test_noalloc_block_svd_refuses_oversize_on_host():
    """Beyond the 64 MiB scratch (or k > 8192): VMPS_ERR_ARG before any device work."""
    lib = _lib.load()
    rows = np.array([3000], dtype=np.int64)
    cols = np.array([3000], dtype=np.int64)
    d = np.zeros(1)
    i = ctypes.c_int()
    st = lib.vmps_block_svd_noalloc(1, rows.ctypes.data_as(P_i64), cols.ctypes.data_as(P_i64), d.ctypes.data_as(P_dbl),
                                    d.ctypes.data_as(P_dbl), d.ctypes.data_as(P_dbl), d.ctypes.data_as(P_dbl),
                                    ctypes.byref(i))
    assert st == 1 and b"scratch" in lib.vmps_last_error()
    rows = np.array([9000], dtype=np.int64)
    cols = np.array([9000], dtype=np.int64)
    st = lib.vmps_block_svd_noalloc(1, rows.ctypes.data_as(P_i64), cols.ctypes.data_as(P_i64), d.ctypes.data_as(P_dbl),
                                    d.ctypes.data_as(P_dbl), d.ctypes.data_as(P_dbl), d.ctypes.data_as(P_dbl),
                                    ctypes.byref(i))
    assert st == 1
