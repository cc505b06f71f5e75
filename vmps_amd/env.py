"""Environment updates contract_L / contract_R on the device
(tensors/DmrgContractions.h:34-311), through the C ABI.

Environments are lists of sectors (qin, qout, qmid, [block_a or None]);
MPS tensors are lists over the local index s of [(qin, qout, block)];
the MPO site tensor is a models.MpoSite.  Blocks are 2-D float64 arrays.
"""
import ctypes
import time

import numpy as np

from . import _lib, pack

P_int = ctypes.POINTER(ctypes.c_int)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)


def _iarr(seq):
    a = np.ascontiguousarray(np.array(seq, dtype=np.int32).ravel())
    return a, a.ctypes.data_as(P_int)


def _read_layout(lib, size_fn, get_fn, h, nq):
    """Sector list (qin, qout, qmid, [block or None]) read through a
    *_size / *_get pair of the C ABI."""
    nsec, nslots, size = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
    _lib.check(size_fn(h, ctypes.byref(nsec), ctypes.byref(nslots), ctypes.byref(size)))
    qin = np.zeros(max(1, nsec.value * nq), dtype=np.int32)
    qout = np.zeros_like(qin)
    qmid = np.zeros_like(qin)
    adim = np.zeros(max(1, nsec.value), dtype=np.int32)
    rows = np.zeros(max(1, nslots.value), dtype=np.int64)
    cols = np.zeros_like(rows)
    data = np.zeros(max(1, size.value))
    _lib.check(get_fn(h, qin.ctypes.data_as(P_int), qout.ctypes.data_as(P_int), qmid.ctypes.data_as(P_int),
                      adim.ctypes.data_as(P_int), rows.ctypes.data_as(P_i64), cols.ctypes.data_as(P_i64),
                      data.ctypes.data_as(P_dbl)))
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


class DeviceEnv:
    """An environment (Tripod) whose blocks stay on the current HIP device
    (vmps_denv): made from host sectors or returned by contract_L / contract_R
    with device=True, and accepted wherever an environment is (the old
    environment of a contraction, L / R of a pivot) without a host copy."""

    def __init__(self, handle, nq):
        self.handle = handle
        self.nq = nq

    @classmethod
    def from_host(cls, sectors, nq):
        lib = _lib.load()
        pk = pack.tripod(sectors, nq)
        h = ctypes.c_void_p()
        _lib.check(lib.vmps_denv_create(nq, pk.count, *pk.args(), ctypes.byref(h)))
        return cls(h, nq)

    def to_host(self):
        lib = _lib.load()
        return _read_layout(lib, lib.vmps_denv_size, lib.vmps_denv_get, self.handle, self.nq)

    def close(self):
        if self.handle is not None and self.handle.value:
            _lib.load().vmps_denv_destroy(self.handle)
        self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _contract(which, site, old, Abra, Aket, precision="fp64", stats=None, device=False, timings=None):
    t0 = time.perf_counter()
    lib = _lib.load()
    sym = site.sym
    kinds, pk = _iarr(sym.kinds)
    ql, pql = _iarr(site.qloc)
    qo, pqo = _iarr(site.qop)
    h = ctypes.c_void_p()
    _lib.check(lib.vmps_env_create(which, sym.nq, pk, len(site.qloc), pql, len(site.qop), pqo, ctypes.byref(h)))
    keep = []
    try:
        if isinstance(old, DeviceEnv):
            _lib.check(lib.vmps_env_add_old_device(h, old.handle))
        else:
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
        t1 = time.perf_counter()
        if device:
            d = ctypes.c_void_p()
            _lib.check(lib.vmps_env_contract_device(h, _lib.PREC[precision], ctypes.byref(d)))
            out = DeviceEnv(d, sym.nq)
        else:
            _lib.check(lib.vmps_env_contract(h, _lib.PREC[precision]))
            out = _read_layout(lib, lib.vmps_env_result_size, lib.vmps_env_result_get, h, sym.nq)
        if stats is not None:
            st = _lib.PlanStats()
            _lib.check(lib.vmps_env_stats(h, ctypes.byref(st)))
            stats.update(st.as_dict())
        if timings is not None:  # host-side hand-over of the blocks, then the contraction itself
            for key, dt in (("env_prepare", t1 - t0), ("env_contract", time.perf_counter() - t1)):
                timings[key] = timings.get(key, 0.0) + dt
    finally:
        lib.vmps_env_destroy(h)
    return out


def contract_L(site, Lold, Abra, Aket, precision="fp64", stats=None, device=False, timings=None):
    """contract_L(Lold, Abra, W, Aket, qloc, qOp, Lnew) on the device; Lold
    may be a DeviceEnv, and device=True returns Lnew as one."""
    return _contract(0, site, Lold, Abra, Aket, precision, stats, device, timings)


def contract_R(site, Rold, Abra, Aket, precision="fp64", stats=None, device=False, timings=None):
    """contract_R(Rold, Abra, W, Aket, qloc, qOp, Rnew) on the device; Rold
    may be a DeviceEnv, and device=True returns Rnew as one."""
    return _contract(1, site, Rold, Abra, Aket, precision, stats, device, timings)
