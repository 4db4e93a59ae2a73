"""ctypes binding of the native library libebem_hip.so (C ABI of include/ebem.h).

The library is the only compute path: if it is missing or cannot be loaded this
module raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libebem_hip.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

# symbol -> (restype, argtypes); every symbol declared in include/ebem.h
SIGNATURES = {
    "ebem_last_error": (C.c_char_p, []),
    "ebem_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "ebem_destroy": (C.c_int, [_vp]),
    "ebem_set_mesh": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _ip, C.c_int]),
    "ebem_set_problem": (C.c_int, [_vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "ebem_set_rule": (C.c_int, [_vp, C.c_int]),
    "ebem_set_inner_rule": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int]),
    "ebem_inner_rule_tables": (C.c_int, [C.c_int, C.c_int, C.c_int, _dp, _dp]),
    "ebem_num_blocks": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "ebem_assemble": (C.c_int, [_vp, C.c_int]),
    "ebem_store_info": (C.c_int, [_vp, _i64p, _i64p, _i64p, _i64p]),
    "ebem_assembly_stats": (C.c_int, [_vp, _i64p, _i64p]),
    "ebem_assembly_kp_stats": (C.c_int, [_vp, _i64p]),
    "ebem_pp_kp_stats": (C.c_int, [_vp, _i64p]),
    "ebem_store_lag_blocks": (C.c_int, [_vp, _i64p]),
    "ebem_get_block": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_get_block_rows": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "ebem_rhs": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_k_num_blocks": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "ebem_assemble_k": (C.c_int, [_vp]),
    "ebem_k_store_info": (C.c_int, [_vp, _i64p, _i64p, _i64p]),
    "ebem_get_k_block_rows": (C.c_int, [_vp, C.c_int, C.c_int, C.c_int, _dp]),
    "ebem_dirichlet_datum": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_beta_k": (C.c_int, [_vp, _dp, _dp]),
    "ebem_rhs_dirichlet": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_pp_eval_direct": (C.c_int, [_vp, _dp, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "ebem_neumann_datum": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_beta_v": (C.c_int, [_vp, _dp, _dp]),
    "ebem_mot_solve_neumann": (C.c_int, [_vp, C.c_int, _dp]),
    "ebem_mot_solve": (C.c_int, [_vp, _dp, _dp]),
    "ebem_pp_eval": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "ebem_pp_prepare": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp]),
    "ebem_node_integral": (C.c_int, [_vp, _dp, C.c_int, C.c_double, _dp]),
    "ebem_set_pp_rule": (C.c_int, [_vp, C.c_int, C.c_int]),
    "ebem_pp_eval_numeric": (C.c_int, [_vp, _dp, C.c_int, _dp, C.c_int, _dp, _dp]),
    "ebem_timings": (C.c_int, [_vp, _dp, C.c_int]),
    "ebem_analiticIntg_CC": (C.c_int, [_vp] + [C.c_double] * 7 + [_dp, _dp, _dp]),
    "ebem_analiticIntg_SC_P": (C.c_int, [_vp] + [C.c_double] * 5 + [_dp, _dp, _dp]),
    "ebem_analiticIntg_SC_SP": (C.c_int, [_vp] + [C.c_double] * 5 + [_dp, _dp, _dp]),
    "ebem_analiticIntg_TC": (C.c_int, [_vp] + [C.c_double] * 5 + [_dp, _dp, _dp]),
    "ebem_analiticIntg_TR_P": (C.c_int, [_vp] + [C.c_double] * 3 + [_dp, _dp, _dp]),
    "ebem_analiticIntg_TR_SP": (C.c_int, [_vp] + [C.c_double] * 3 + [_dp, _dp, _dp]),
    "ebem_analitic_intg": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp, _dp]),
    "ebem_analitic_intg_noalloc": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp, _dp]),
    "ebem_set_option": (C.c_int, [_vp, C.c_int, C.c_double]),
    "ebem_sur_create": (C.c_int, [C.POINTER(_vp), C.c_int]),
    "ebem_sur_destroy": (C.c_int, [_vp]),
    "ebem_sur_add_samples": (C.c_int, [_vp, _vp, C.c_int64, _ip, _ip, _ip]),
    "ebem_sur_num_samples": (C.c_int, [_vp, _i64p]),
    "ebem_sur_get_samples": (C.c_int, [_vp, C.c_int64, C.c_int64, _fp, _fp]),
    "ebem_sur_init": (C.c_int, [_vp, C.c_int64, C.c_uint64]),
    "ebem_sur_train": (C.c_int, [_vp, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp]),
    "ebem_sur_loss": (C.c_int, [_vp, C.c_int64, C.c_int64, _dp]),
    "ebem_sur_grad": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _dp]),
    "ebem_sur_predict": (C.c_int, [_vp, C.c_int64, _fp, _fp]),
    "ebem_sur_params": (C.c_int, [_vp, C.c_int, _fp, _fp]),
    "ebem_attach_surrogate": (C.c_int, [_vp, _vp]),
    "ebem_surrogate_stats": (C.c_int, [_vp, _i64p, _i64p, _dp]),
}

_lib = None


class EbemError(RuntimeError):
    """A native call returned a negative EBEM_E* code."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


def load() -> C.CDLL:
    """Load libebem_hip.so (built by elastodynamicebem_amd.build) or raise."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"native library missing: {LIB_PATH} (run __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc != 0:
        msg = load().ebem_last_error().decode(errors="replace")
        raise EbemError(rc, msg)


def dptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def fptr(a: np.ndarray | None):
    if a is None:
        return None
    return a.ctypes.data_as(_fp)
