"""ctypes binding of libvmps_hxv.so (the C ABI declared in include/vmps_hxv.h).

The library is loaded from the package directory only.  A missing library is
an error: there is no CPU fallback for the device path.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvmps_hxv.so")

c_int = ctypes.c_int
c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
P_int = ctypes.POINTER(ctypes.c_int)
P_i64 = ctypes.POINTER(ctypes.c_int64)
P_dbl = ctypes.POINTER(ctypes.c_double)
c_void = ctypes.c_void_p

VMPS_OK = 0
STATUS_NAMES = {
    0: "VMPS_OK", 1: "VMPS_ERR_ARG", 2: "VMPS_ERR_LAYOUT", 3: "VMPS_ERR_OOM", 4: "VMPS_ERR_HIP",
    5: "VMPS_ERR_NONFINITE", 6: "VMPS_ERR_STATE", 7: "VMPS_ERR_INTERNAL", 8: "VMPS_ERR_NOCONV",
}
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}


class PlanStats(ctypes.Structure):
    _fields_ = [
        ("n_terms", c_i64), ("n_T", c_i64), ("n_U", c_i64), ("n_out", c_i64),
        ("n_L", c_i64), ("n_R", c_i64), ("n_V", c_i64),
        ("flops_alg", c_dbl), ("flops_ref", c_dbl), ("bytes_alg_fp64", c_dbl),
        ("L_elems", c_i64), ("R_elems", c_i64), ("V_elems", c_i64), ("T_elems", c_i64), ("U_elems", c_i64),
        ("max_block_dim", c_i64), ("flops_A", c_dbl), ("flops_C", c_dbl),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class LanczosOpts(ctypes.Structure):
    _fields_ = [
        ("max_iter", c_int), ("min_iter", c_int), ("tol_eigval", c_dbl), ("tol_state", c_dbl),
        ("precision", c_int), ("refine_fp64", c_int), ("refine_tol_eigval", c_dbl), ("refine_tol_state", c_dbl),
        ("seed", ctypes.c_uint64), ("graph", c_int), ("refine_method", c_int),
    ]


class LanczosResult(ctypes.Structure):
    _fields_ = [
        ("energy", c_dbl), ("residual", c_dbl), ("iterations", c_int), ("iterations_refine", c_int),
        ("hxv_calls", c_int), ("converged", c_int), ("energy_mixed", c_dbl), ("residual_mixed", c_dbl),
        ("ms", c_dbl), ("hxv_fp64", c_int), ("refine_steps", c_int), ("refine_cg", c_int),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class TileConfig(ctypes.Structure):
    _fields_ = [("wave_edge", c_int), ("big_min", c_int), ("big_min64", c_int), ("explicit_u", c_int), ("ms", c_dbl),
                ("from_cache", c_int), ("n_tried", c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class VmpsStatusError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s: %s" % (STATUS_NAMES.get(status, status), msg))
        self.status = status


# name -> (restype, argtypes)
SIGNATURES = {
    "vmps_abi_version": (c_int, []),
    "vmps_last_error": (ctypes.c_char_p, []),
    "vmps_plan_create_batch": (c_int, [ctypes.POINTER(c_void), c_int, c_int, ctypes.POINTER(c_void), P_i64]),
    "vmps_coupling_6j": (c_dbl, [c_int] * 6),
    "vmps_coupling_9j": (c_dbl, [c_int] * 9),
    "vmps_pivot1_create": (c_int, [c_int, P_int, c_int, P_int, c_int, P_int, ctypes.POINTER(c_void)]),
    "vmps_pivot0_create": (c_int, [c_int, P_int, ctypes.POINTER(c_void)]),
    "vmps_pivot_add_term": (c_int, [c_void, c_int, P_int, c_int, P_int, P_int]),
    "vmps_pivot_select_term": (c_int, [c_void, c_int]),
    "vmps_pivot_info": (c_int, [c_void, P_int, P_int]),
    "vmps_pivot2_create": (c_int, [c_int, P_int, c_int, P_int, c_int, P_int, c_int, P_int, c_int, P_int,
                                   ctypes.POINTER(c_void)]),
    "vmps_pivot2_destroy": (None, [c_void]),
    "vmps_pivot2_local_basis": (c_int, [c_void, P_int, P_int]),
    "vmps_pivot2_add_env": (c_int, [c_void, c_int, P_int, P_int, P_int, c_int, P_i64, P_i64, P_dbl]),
    "vmps_pivot2_add_W": (c_int, [c_void, c_int, c_int, c_int, c_int, P_int, P_int, c_int, c_int, c_i64,
                                  P_int, P_int, P_dbl]),
    "vmps_pivot2_add_V_block": (c_int, [c_void, c_int, P_int, P_int, c_i64, c_i64, P_i64]),
    "vmps_pivot2_add_env_sectors": (c_int, [c_void, c_int, c_int, P_int, P_int, P_int, P_int, P_i64, P_i64,
                                            ctypes.POINTER(c_void)]),
    "vmps_pivot2_add_W_blocks": (c_int, [c_void, c_int, c_int, P_int, P_int, P_int, P_int, P_i64, P_int, P_int,
                                         P_dbl]),
    "vmps_pivot2_add_V_blocks": (c_int, [c_void, c_int, P_int, P_int, P_int, P_i64, P_i64, P_i64]),
    "vmps_pivot2_vsize": (c_int, [c_void, P_i64]),
    "vmps_pivot2_precalc": (c_int, [c_void]),
    "vmps_pivot2_set_blockstructure": (c_int, [c_void, c_i64, P_i64, P_i64, P_i64, P_dbl]),
    "vmps_pivot2_get_blockstructure": (c_int, [c_void, P_i64, P_i64, P_i64, P_i64, P_i64, P_dbl]),
    "vmps_pivot2_set_projectors": (c_int, [c_void, c_dbl, c_int, P_dbl]),
    "vmps_pivot2_stats": (c_int, [c_void, ctypes.POINTER(PlanStats)]),
    "vmps_plan_create": (c_int, [c_void, c_int, ctypes.POINTER(c_void)]),
    "vmps_plan_destroy": (None, [c_void]),
    "vmps_plan_stats_get": (c_int, [c_void, ctypes.POINTER(PlanStats)]),
    "vmps_plan_set_precision": (c_int, [c_void, c_int]),
    "vmps_hxv": (c_int, [c_void, P_dbl, P_dbl]),
    "vmps_hxv_device": (c_int, [c_void, c_void, c_void, c_void]),
    "vmps_dot": (c_int, [c_void, P_dbl, P_dbl, P_dbl]),
    "vmps_plan_time_hxv": (c_int, [c_void, c_int, c_int, P_dbl]),
    "vmps_plan_time_stages": (c_int, [c_void, c_int, c_int, P_dbl]),
    "vmps_plan_get_tiles": (c_int, [c_void, ctypes.POINTER(TileConfig)]),
    "vmps_block_svd": (c_int, [c_int, P_i64, P_i64, P_dbl, P_dbl, P_dbl, P_dbl, P_int]),
    "vmps_block_svd_noalloc": (c_int, [c_int, P_i64, P_i64, P_dbl, P_dbl, P_dbl, P_dbl, P_int]),
    "vmps_block_eigh_split": (c_int, [c_int, P_i64, P_i64, P_dbl, c_int, P_dbl, P_dbl, P_dbl]),
    "vmps_plan_set_tiles": (c_int, [c_void, ctypes.POINTER(TileConfig)]),
    "vmps_plan_autotune": (c_int, [c_void, c_int, ctypes.POINTER(TileConfig)]),
    "vmps_autotune_clear_cache": (c_int, []),
    "vmps_env_create": (c_int, [c_int, c_int, P_int, c_int, P_int, c_int, P_int, ctypes.POINTER(c_void)]),
    "vmps_env_destroy": (None, [c_void]),
    "vmps_env_add_old": (c_int, [c_void, P_int, P_int, P_int, c_int, P_i64, P_i64, P_dbl]),
    "vmps_env_add_A": (c_int, [c_void, c_int, c_int, P_int, P_int, c_i64, c_i64, P_dbl]),
    "vmps_env_add_W": (c_int, [c_void, c_int, c_int, c_int, P_int, P_int, c_int, c_int, c_i64, P_int, P_int, P_dbl]),
    "vmps_env_add_old_sectors": (c_int, [c_void, c_int, P_int, P_int, P_int, P_int, P_i64, P_i64,
                                         ctypes.POINTER(c_void)]),
    "vmps_env_add_A_blocks": (c_int, [c_void, c_int, c_int, P_int, P_int, P_int, P_i64, P_i64, ctypes.POINTER(c_void)]),
    "vmps_env_add_W_blocks": (c_int, [c_void, c_int, P_int, P_int, P_int, P_int, P_i64, P_int, P_int, P_dbl]),
    "vmps_env_contract": (c_int, [c_void, c_int]),
    "vmps_env_result_size": (c_int, [c_void, P_i64, P_i64, P_i64]),
    "vmps_env_result_get": (c_int, [c_void, P_int, P_int, P_int, P_int, P_i64, P_i64, P_dbl]),
    "vmps_env_stats": (c_int, [c_void, ctypes.POINTER(PlanStats)]),
    "vmps_denv_create": (c_int, [c_int, c_int, P_int, P_int, P_int, P_int, P_i64, P_i64, ctypes.POINTER(c_void),
                                 ctypes.POINTER(c_void)]),
    "vmps_denv_destroy": (None, [c_void]),
    "vmps_denv_size": (c_int, [c_void, P_i64, P_i64, P_i64]),
    "vmps_denv_get": (c_int, [c_void, P_int, P_int, P_int, P_int, P_i64, P_i64, P_dbl]),
    "vmps_env_contract_device": (c_int, [c_void, c_int, ctypes.POINTER(c_void)]),
    "vmps_env_add_old_device": (c_int, [c_void, c_void]),
    "vmps_pivot2_add_env_device": (c_int, [c_void, c_int, c_void]),
    "vmps_vec_alloc": (c_int, [c_void, ctypes.POINTER(c_void)]),
    "vmps_vec_free": (c_int, [c_void, c_void]),
    "vmps_vec_upload": (c_int, [c_void, c_void, P_dbl]),
    "vmps_vec_download": (c_int, [c_void, c_void, P_dbl]),
    "vmps_vec_dot": (c_int, [c_void, c_void, c_void, P_dbl]),
    "vmps_vec_norm": (c_int, [c_void, c_void, P_dbl]),
    "vmps_vec_axpby": (c_int, [c_void, c_dbl, c_void, c_dbl, c_void]),
    "vmps_vec_fill_random": (c_int, [c_void, c_void, ctypes.c_uint64]),
    "vmps_block_eigh_split_gather": (c_int, [c_void, c_i64, c_int, P_i64, P_i64, c_int, P_int, P_i64, P_i64, P_i64,
                                             P_i64, P_i64, P_i64, P_dbl, c_int, P_dbl, P_dbl, P_dbl]),
    "vmps_contract_aa2_device": (c_int, [c_void, c_int, P_i64, P_i64, ctypes.POINTER(c_void), c_int, P_i64, P_i64,
                                         ctypes.POINTER(c_void), c_int, P_i64, P_int, P_int, P_dbl, c_void]),
    "vmps_lanczos_opts_default": (None, [ctypes.POINTER(LanczosOpts)]),
    "vmps_lanczos": (c_int, [c_void, P_dbl, ctypes.POINTER(LanczosOpts), ctypes.POINTER(LanczosResult)]),
    "vmps_lanczos_device": (c_int, [c_void, c_void, ctypes.POINTER(LanczosOpts), ctypes.POINTER(LanczosResult)]),
    "vmps_lanczos_batch": (c_int, [c_void, P_dbl, ctypes.POINTER(LanczosOpts), ctypes.POINTER(LanczosResult)]),
}

_lib = None


def load():
    """Load the library (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("libvmps_hxv.so is not built (run `python -m vmps_amd.build`); "
                           "the HIP engine has no CPU fallback")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status):
    if status != VMPS_OK:
        msg = load().vmps_last_error()
        raise VmpsStatusError(status, msg.decode() if msg else "")
