"""ORACLE -- test infrastructure, not part of the product.

Loader of the CPU FP64 restatement of the reference's zero-, one- and two-site HxV
(hxv_oracle.c).  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, always as the checker or the timed CPU baseline.
"""
import ctypes
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))

P_int = ctypes.POINTER(ctypes.c_int)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)
c_void = ctypes.c_void_p

SIGNATURES = {
    "oracle_last_error": (ctypes.c_char_p, []),
    "oracle_pivot2_create": (ctypes.c_int, [ctypes.c_int, P_int, ctypes.c_int, P_int, ctypes.c_int, P_int,
                                            ctypes.c_int, P_int, ctypes.c_int, P_int, ctypes.POINTER(c_void)]),
    "oracle_pivot1_create": (ctypes.c_int, [ctypes.c_int, P_int, ctypes.c_int, P_int, ctypes.c_int, P_int,
                                            ctypes.POINTER(c_void)]),
    "oracle_pivot0_create": (ctypes.c_int, [ctypes.c_int, P_int, ctypes.POINTER(c_void)]),
    "oracle_pivot2_destroy": (None, [c_void]),
    "oracle_pivot2_local_basis": (ctypes.c_int, [c_void, P_int, P_int]),
    "oracle_pivot2_add_env": (ctypes.c_int, [c_void, ctypes.c_int, P_int, P_int, P_int, ctypes.c_int, P_i64, P_i64,
                                             P_dbl]),
    "oracle_pivot2_add_W": (ctypes.c_int, [c_void, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P_int,
                                           P_int, ctypes.c_int, ctypes.c_int, ctypes.c_int64, P_int, P_int, P_dbl]),
    "oracle_pivot2_add_V_block": (ctypes.c_int, [c_void, ctypes.c_int, P_int, P_int, ctypes.c_int64,
                                                 ctypes.c_int64, P_i64]),
    "oracle_pivot2_vsize": (ctypes.c_int, [c_void, P_i64]),
    "oracle_pivot2_precalc": (ctypes.c_int, [c_void]),
    "oracle_pivot2_set_blockstructure": (ctypes.c_int, [c_void, ctypes.c_int64, P_i64, P_i64, P_i64, P_dbl]),
    "oracle_pivot2_get_blockstructure": (ctypes.c_int, [c_void, P_i64, P_i64, P_i64, P_i64, P_i64, P_dbl]),
    "oracle_pivot2_set_projectors": (ctypes.c_int, [c_void, ctypes.c_double, ctypes.c_int, P_dbl]),
    "oracle_hxv": (ctypes.c_int, [c_void, P_dbl, P_dbl]),
    "oracle_hxv_omp": (ctypes.c_int, [c_void, P_dbl, P_dbl, ctypes.c_int]),
    "oracle_dot": (ctypes.c_int, [c_void, P_dbl, P_dbl, P_dbl]),
    "oracle_flops_ref": (ctypes.c_int, [c_void, P_dbl]),
    "oracle_dgemm_gflops": (ctypes.c_double, [ctypes.c_int, ctypes.c_int]),
    "oracle_hxv_range": (ctypes.c_int, [c_void, P_dbl, P_dbl, ctypes.c_int64, ctypes.c_int64]),
    "oracle_nlhs": (ctypes.c_int64, [c_void]),
    "oracle_flops_ref_range": (ctypes.c_int, [c_void, ctypes.c_int64, ctypes.c_int64, P_dbl]),
    "oracle_wigner6j": (ctypes.c_double, [ctypes.c_int] * 6),
    "oracle_wigner9j": (ctypes.c_double, [ctypes.c_int] * 9),
    "oracle_coeff_HPsi": (ctypes.c_double, [ctypes.c_int, P_int, P_int]),
    "oracle_coeff_buildR": (ctypes.c_double, [ctypes.c_int, P_int, P_int]),
    "oracle_coeff_buildL": (ctypes.c_double, [ctypes.c_int, P_int, P_int]),
    "oracle_coeff_Apair": (ctypes.c_double, [ctypes.c_int, P_int, P_int]),
}

_lib = None


def _has_avx512():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    fl = line.split()
                    return "avx512f" in fl and "avx512dq" in fl and "avx512bw" in fl and "avx512vl" in fl
    except OSError:
        pass
    return False


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def load():
    global _lib
    if _lib is not None:
        return _lib
    name = "liboracle_hxv_v4.so" if _has_avx512() else "liboracle_hxv_v3.so"
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    for fn, (res, args) in SIGNATURES.items():
        f = getattr(lib, fn)
        f.restype = res
        f.argtypes = args
    lib.variant = name
    _lib = lib
    return lib


def check(status):
    if status != 0:
        raise RuntimeError("oracle status %d: %s" % (status, load().oracle_last_error().decode()))
