"""Environment updates contract_L / contract_R on the device
(tensors/DmrgContractions.h:34-311), through the C ABI.

Environments are lists of sectors (qin, qout, qmid, [block_a or None]);
MPS tensors are lists over the local index s of [(qin, qout, block)];
the MPO site tensor is a models.MpoSite.  Blocks are 2-D float64 arrays.
"""
import ctypes

import numpy as np

from . import _lib, pack

P_int = ctypes.POINTER(ctypes.c_int)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)


def _iarr(seq):
    a = np.ascontiguousarray(np.array(seq, dtype=np.int32).ravel())
    return a, a.ctypes.data_as(P_int)


def _contract(which, site, old, Abra, Aket, precision="fp64", stats=None):
    lib = _lib.load()
    sym = site.sym
    kinds, pk = _iarr(sym.kinds)
    ql, pql = _iarr(site.qloc)
    qo, pqo = _iarr(site.qop)
    h = ctypes.c_void_p()
    _lib.check(lib.vmps_env_create(which, sym.nq, pk, len(site.qloc), pql, len(site.qop), pqo, ctypes.byref(h)))
    keep = []
    try:
        bk = pack.tripod(old, sym.nq)
        keep.append(bk)
        _lib.check(lib.vmps_env_add_old_sectors(h, bk.count, *bk.args()))
        for bra, A in ((1, Abra), (0, Aket)):
            bk = pack.a_blocks(A, sym.nq)
            keep.append(bk)
            _lib.check(lib.vmps_env_add_A_blocks(h, bra, bk.count, *bk.args()))
        bk = pack.w_blocks(site.blocks, sym.nq)
        keep.append(bk)
        _lib.check(lib.vmps_env_add_W_blocks(h, bk.count, *bk.args()))
        _lib.check(lib.vmps_env_contract(h, _lib.PREC[precision]))
        nsec, nslots, size = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _lib.check(lib.vmps_env_result_size(h, ctypes.byref(nsec), ctypes.byref(nslots), ctypes.byref(size)))
        nq = sym.nq
        qin = np.zeros(max(1, nsec.value * nq), dtype=np.int32)
        qout = np.zeros_like(qin)
        qmid = np.zeros_like(qin)
        adim = np.zeros(max(1, nsec.value), dtype=np.int32)
        rows = np.zeros(max(1, nslots.value), dtype=np.int64)
        cols = np.zeros_like(rows)
        data = np.zeros(max(1, size.value))
        _lib.check(lib.vmps_env_result_get(h, qin.ctypes.data_as(P_int), qout.ctypes.data_as(P_int),
                                           qmid.ctypes.data_as(P_int), adim.ctypes.data_as(P_int),
                                           rows.ctypes.data_as(P_i64), cols.ctypes.data_as(P_i64),
                                           data.ctypes.data_as(P_dbl)))
        if stats is not None:
            st = _lib.PlanStats()
            _lib.check(lib.vmps_env_stats(h, ctypes.byref(st)))
            stats.update(st.as_dict())
    finally:
        lib.vmps_env_destroy(h)
    out, slot, pos = [], 0, 0
    for q in range(nsec.value):
        blocks = []
        for a in range(adim[q]):
            r, c = int(rows[slot]), int(cols[slot])
            if r * c == 0:
                blocks.append(None)
            else:
                blocks.append(np.asfortranarray(data[pos:pos + r * c].reshape((r, c), order="F")))
                pos += r * c
            slot += 1
        out.append((tuple(int(x) for x in qin[q * nq:(q + 1) * nq]), tuple(int(x) for x in qout[q * nq:(q + 1) * nq]),
                    tuple(int(x) for x in qmid[q * nq:(q + 1) * nq]), blocks))
    return out


def contract_L(site, Lold, Abra, Aket, precision="fp64", stats=None):
    """contract_L(Lold, Abra, W, Aket, qloc, qOp, Lnew) on the device."""
    return _contract(0, site, Lold, Abra, Aket, precision, stats)


def contract_R(site, Rold, Abra, Aket, precision="fp64", stats=None):
    """contract_R(Rold, Abra, W, Aket, qloc, qOp, Rnew) on the device."""
    return _contract(1, site, Rold, Abra, Aket, precision, stats)
