"""Python face of the engine, mirroring the reference's pivot API.

    H = PivotMatrix2(pivot_data)          # PivotMatrix2 + precalc_blockStructure
    Vout = HxV(H, Vin)                    # HxV(PivotMatrix2, PivotVector, PivotVector&)
    dot(H, V1, V2)                        # dot(PivotVector, PivotVector) with coeff_dot weights

(pivot/DmrgPivotMatrix2.h:29-83, pivot/DmrgPivotVector.h:296-358.)  Vectors
are packed FP64 numpy arrays in the pivot's block order.  All arithmetic runs
in the HIP library; there is no Python or CPU fallback.
"""
import ctypes

import numpy as np

from . import _lib
from . import synth

P_dbl = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(P_dbl)


class PivotMatrix2:
    """Two-site effective Hamiltonian resident on the current HIP device (a
    sum over the terms of `pd`)."""
    NSITE = 2

    def __init__(self, pd, precision="fp64", blockstructure=None, projectors=None, Epenalty=0.0):
        if pd.nsite != self.NSITE:
            raise ValueError("%s needs a %d-site pivot, got %d sites" % (type(self).__name__, self.NSITE, pd.nsite))
        lib = _lib.load()
        self.pd = pd
        self._lib = lib
        h = synth.feed(lib, "vmps_", pd, _lib.check, projectors=projectors, Epenalty=Epenalty)
        try:
            if blockstructure is None:
                _lib.check(lib.vmps_pivot2_precalc(h))
            else:
                set_blockstructure(lib, h, *blockstructure)
            self.blockstructure = get_blockstructure(lib, h)
            vs = ctypes.c_int64()
            _lib.check(lib.vmps_pivot2_vsize(h, ctypes.byref(vs)))
            self.vsize = vs.value
            plan = ctypes.c_void_p()
            _lib.check(lib.vmps_plan_create(h, _lib.PREC[precision], ctypes.byref(plan)))
            self._plan = plan
        finally:
            lib.vmps_pivot2_destroy(h)
        self.precision = precision

    def close(self):
        if getattr(self, "_plan", None):
            self._lib.vmps_plan_destroy(self._plan)
            self._plan = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_precision(self, precision):
        _lib.check(self._lib.vmps_plan_set_precision(self._plan, _lib.PREC[precision]))
        self.precision = precision

    def stats(self):
        st = _lib.PlanStats()
        _lib.check(self._lib.vmps_plan_stats_get(self._plan, ctypes.byref(st)))
        return st.as_dict()

    def tiles(self):
        """Current tile configuration: {'wave_edge', 'big_min', ...}."""
        cfg = _lib.TileConfig()
        _lib.check(self._lib.vmps_plan_get_tiles(self._plan, ctypes.byref(cfg)))
        return cfg.as_dict()

    def set_tiles(self, wave_edge=None, big_min=None, explicit_u=None, big_min64=None):
        """Set the tile configuration; parameters left None keep their value."""
        cfg = _lib.TileConfig()
        _lib.check(self._lib.vmps_plan_get_tiles(self._plan, ctypes.byref(cfg)))
        for k, v in (("wave_edge", wave_edge), ("big_min", big_min), ("explicit_u", explicit_u),
                     ("big_min64", big_min64)):
            if v is not None:
                setattr(cfg, k, int(v))
        _lib.check(self._lib.vmps_plan_set_tiles(self._plan, ctypes.byref(cfg)))

    def autotune(self, iters=10):
        """Time the candidate tile configurations of this plan's precision and
        keep the fastest (cached per block-size histogram)."""
        cfg = _lib.TileConfig()
        _lib.check(self._lib.vmps_plan_autotune(self._plan, int(iters), ctypes.byref(cfg)))
        return cfg.as_dict()

    def time_stages(self, warmup=3, iters=20):
        ms = (ctypes.c_double * 4)()
        _lib.check(self._lib.vmps_plan_time_stages(self._plan, warmup, iters, ms))
        return list(ms)

    def time_hxv(self, warmup=3, iters=20):
        ms = ctypes.c_double()
        _lib.check(self._lib.vmps_plan_time_hxv(self._plan, warmup, iters, ctypes.byref(ms)))
        return ms.value


class PivotMatrix1(PivotMatrix2):
    """One-site effective Hamiltonian (PivotMatrix1, pivot/DmrgPivotMatrix1.h:24-66)."""
    NSITE = 1


class PivotMatrix0(PivotMatrix2):
    """Zero-site effective Hamiltonian (PivotMatrix0, pivot/DmrgPivotMatrix0.h:10-27)."""
    NSITE = 0


class PivotBatch(PivotMatrix2):
    """Independent pivots applied together: one plan over the concatenated
    vectors, one launch per stage for the whole batch.  `offsets[i]` is the
    start of pivot i's vector."""

    def __init__(self, pds, precision="fp64"):
        lib = _lib.load()
        self._lib = lib
        self.pds = list(pds)
        handles = []
        try:
            for pd in self.pds:
                h = synth.feed(lib, "vmps_", pd, _lib.check)
                handles.append(h)
                _lib.check(lib.vmps_pivot2_precalc(h))
            arr = (ctypes.c_void_p * len(handles))(*[h.value for h in handles])
            off = np.zeros(len(handles) + 1, dtype=np.int64)
            plan = ctypes.c_void_p()
            _lib.check(lib.vmps_plan_create_batch(arr, len(handles), _lib.PREC[precision], ctypes.byref(plan),
                                                  off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))))
            self._plan = plan
        finally:
            for h in handles:
                lib.vmps_pivot2_destroy(h)
        self.offsets = off
        self.vsize = int(off[-1])
        self.precision = precision

    def split(self, v):
        return [v[self.offsets[i]:self.offsets[i + 1]] for i in range(len(self.pds))]


class DeviceVector:
    """A PivotVector resident in device memory (the plan's layout)."""

    def __init__(self, H, host=None):
        self.H = H
        p = ctypes.c_void_p()
        _lib.check(H._lib.vmps_vec_alloc(H._plan, ctypes.byref(p)))
        self.ptr = p.value
        if host is not None:
            self.upload(host)

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=np.float64)
        assert host.size == self.H.vsize
        _lib.check(self.H._lib.vmps_vec_upload(self.H._plan, self.ptr, _ptr(host)))

    def download(self):
        out = np.empty(self.H.vsize)
        _lib.check(self.H._lib.vmps_vec_download(self.H._plan, self.ptr, _ptr(out)))
        return out

    def free(self):
        if self.ptr:
            _lib.check(self.H._lib.vmps_vec_free(self.H._plan, self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def clear_tile_cache():
    """Forget the autotuner's cached tile choices (process-wide), so that the
    next plan of each shape is timed again."""
    _lib.check(_lib.load().vmps_autotune_clear_cache())


def HxV(H, vin):
    vin = np.ascontiguousarray(vin, dtype=np.float64)
    if vin.size != H.vsize:
        raise ValueError("vector size %d does not match the pivot (%d)" % (vin.size, H.vsize))
    vout = np.empty_like(vin)
    _lib.check(H._lib.vmps_hxv(H._plan, _ptr(vin), _ptr(vout)))
    return vout


def dot(H, x, y):
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    r = ctypes.c_double()
    _lib.check(H._lib.vmps_dot(H._plan, _ptr(x), _ptr(y), ctypes.byref(r)))
    return r.value


REFINE = {"correction": 0, "restart": 1}


def _lanczos_opts(H, precision, refine_fp64, max_iter, tol_eigval, tol_state, refine_tol_eigval, refine_tol_state,
                  seed, graph, refine_method):
    o = _lib.LanczosOpts()
    H._lib.vmps_lanczos_opts_default(ctypes.byref(o))
    o.max_iter, o.tol_eigval, o.tol_state = max_iter, tol_eigval, tol_state
    o.precision = _lib.PREC[precision]
    o.refine_fp64 = 1 if refine_fp64 else 0
    o.refine_tol_eigval, o.refine_tol_state = refine_tol_eigval, refine_tol_state
    o.seed = seed
    o.graph = 1 if graph else 0
    o.refine_method = REFINE[refine_method]
    return o


def lanczos_device(H, dv, precision="fp64", refine_fp64=True, max_iter=100, tol_eigval=1e-8, tol_state=1e-5,
                   refine_tol_eigval=1e-12, refine_tol_state=1e-7, seed=1, graph=True, refine_method="correction"):
    """lanczos() on a DeviceVector: dv holds the start vector and receives
    the eigenvector.  Returns the result dict."""
    o = _lanczos_opts(H, precision, refine_fp64, max_iter, tol_eigval, tol_state, refine_tol_eigval,
                      refine_tol_state, seed, graph, refine_method)
    r = _lib.LanczosResult()
    _lib.check(H._lib.vmps_lanczos_device(H._plan, dv.ptr, ctypes.byref(o), ctypes.byref(r)))
    return r.as_dict()


def contract_aa2(H, dv, blocks1, blocks2, out_off, i1, i2, factor):
    """contract_AA2 on the device into the DeviceVector dv
    (vmps_contract_aa2_device): dv = sum over terms t of factor[t] *
    blocks1[i1[t]] . blocks2[i2[t]] at out_off[t], zero elsewhere."""
    from . import pack
    p1, keep1 = pack._block_ptrs(blocks1)
    p2, keep2 = pack._block_ptrs(blocks2)
    r1 = np.array([b.shape[0] for b in blocks1] or [0], dtype=np.int64)
    c1 = np.array([b.shape[1] for b in blocks1] or [0], dtype=np.int64)
    r2 = np.array([b.shape[0] for b in blocks2] or [0], dtype=np.int64)
    c2 = np.array([b.shape[1] for b in blocks2] or [0], dtype=np.int64)
    oo = np.array(out_off or [0], dtype=np.int64)
    t1 = np.array(i1 or [0], dtype=np.int32)
    t2 = np.array(i2 or [0], dtype=np.int32)
    f = np.array(factor or [0.0], dtype=np.float64)
    PP = ctypes.POINTER(ctypes.c_void_p)
    P64 = ctypes.POINTER(ctypes.c_int64)
    P32 = ctypes.POINTER(ctypes.c_int)
    _lib.check(H._lib.vmps_contract_aa2_device(
        H._plan, len(blocks1), r1.ctypes.data_as(P64), c1.ctypes.data_as(P64), p1.ctypes.data_as(PP),
        len(blocks2), r2.ctypes.data_as(P64), c2.ctypes.data_as(P64), p2.ctypes.data_as(PP), len(out_off),
        oo.ctypes.data_as(P64), t1.ctypes.data_as(P32), t2.ctypes.data_as(P32), _ptr(f), dv.ptr))
    del keep1, keep2


def vec_norm(H, dv):
    """norm(V) of a DeviceVector (coeff_dot weights)."""
    r = ctypes.c_double()
    _lib.check(H._lib.vmps_vec_norm(H._plan, dv.ptr, ctypes.byref(r)))
    return r.value


def vec_fill_random(H, dv, seed):
    _lib.check(H._lib.vmps_vec_fill_random(H._plan, dv.ptr, seed))


def lanczos(H, v0=None, precision="fp64", refine_fp64=True, max_iter=100, tol_eigval=1e-8, tol_state=1e-5,
            refine_tol_eigval=1e-12, refine_tol_state=1e-7, seed=1, graph=True, refine_method="correction",
            overwrite_v0=False):
    """Lowest eigenpair of the pivot (the LanczosSolver::edgeState call of
    DmrgSolver::iteration_two).  A mixed-precision solve is refined to the
    FP64 tolerances by corrections ("correction") or an FP64 restart
    ("restart").  Returns (eigenvector, result dict); with `overwrite_v0` a
    contiguous float64 v0 receives the eigenvector in place (no copy)."""
    o = _lanczos_opts(H, precision, refine_fp64, max_iter, tol_eigval, tol_state, refine_tol_eigval,
                      refine_tol_state, seed, graph, refine_method)
    r = _lib.LanczosResult()
    if v0 is None:  # GaussianRandomVector::fill: uniform(-1, 1) entries
        v0 = np.random.default_rng(seed).uniform(-1.0, 1.0, H.vsize)
    v = np.ascontiguousarray(v0, dtype=np.float64)
    if v is v0 and not overwrite_v0:
        v = v.copy()
    _lib.check(H._lib.vmps_lanczos(H._plan, _ptr(v), ctypes.byref(o), ctypes.byref(r)))
    return v, r.as_dict()


def lanczos_batch(H, v0=None, precision="fp64", refine_fp64=True, max_iter=100, tol_eigval=1e-8, tol_state=1e-5,
                  refine_tol_eigval=1e-12, refine_tol_state=1e-7, seed=1, graph=True):
    """One lowest eigenpair per pivot of a PivotBatch, all solved together
    (the per-site solves of VumpsSolver, VUMPS/VumpsSolver.h:1168-1211).
    Returns (batch eigenvector, list of per-pivot result dicts)."""
    if not isinstance(H, PivotBatch):
        raise TypeError("lanczos_batch needs a PivotBatch")
    o = _lib.LanczosOpts()
    H._lib.vmps_lanczos_opts_default(ctypes.byref(o))
    o.max_iter, o.tol_eigval, o.tol_state = max_iter, tol_eigval, tol_state
    o.precision = _lib.PREC[precision]
    o.refine_fp64 = 1 if refine_fp64 else 0
    o.refine_tol_eigval, o.refine_tol_state = refine_tol_eigval, refine_tol_state
    o.seed = seed
    o.graph = 1 if graph else 0
    res = (_lib.LanczosResult * len(H.pds))()
    if v0 is None:
        v0 = np.random.default_rng(seed).uniform(-1.0, 1.0, H.vsize)
    v = np.ascontiguousarray(v0, dtype=np.float64).copy()
    _lib.check(H._lib.vmps_lanczos_batch(H._plan, _ptr(v), ctypes.byref(o), res))
    return v, [r.as_dict() for r in res]


def get_blockstructure(lib, h, prefix="vmps_"):
    fn = getattr(lib, prefix + "pivot2_get_blockstructure")
    nl, nr = ctypes.c_int64(), ctypes.c_int64()
    _lib.check(fn(h, ctypes.byref(nl), ctypes.byref(nr), None, None, None, None))
    qlhs = np.zeros(2 * nl.value, dtype=np.int64)
    rb = np.zeros(nl.value + 1, dtype=np.int64)
    qrhs = np.zeros(12 * nr.value, dtype=np.int64)
    fac = np.zeros(nr.value, dtype=np.float64)
    P64 = ctypes.POINTER(ctypes.c_int64)
    _lib.check(fn(h, ctypes.byref(nl), ctypes.byref(nr), qlhs.ctypes.data_as(P64), rb.ctypes.data_as(P64),
                  qrhs.ctypes.data_as(P64), fac.ctypes.data_as(P_dbl)))
    return qlhs.reshape(-1, 2), rb, qrhs.reshape(-1, 12), fac


def set_blockstructure(lib, h, qlhs, rhs_begin, qrhs, factors, prefix="vmps_"):
    fn = getattr(lib, prefix + "pivot2_set_blockstructure")
    P64 = ctypes.POINTER(ctypes.c_int64)
    qlhs = np.ascontiguousarray(qlhs, dtype=np.int64).ravel()
    rb = np.ascontiguousarray(rhs_begin, dtype=np.int64)
    qrhs = np.ascontiguousarray(qrhs, dtype=np.int64).ravel()
    fac = np.ascontiguousarray(factors, dtype=np.float64)
    _lib.check(fn(h, len(rb) - 1, qlhs.ctypes.data_as(P64), rb.ctypes.data_as(P64), qrhs.ctypes.data_as(P64),
                  fac.ctypes.data_as(P_dbl)))
