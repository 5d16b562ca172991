"""ctypes binding of libmstm_hip.so (the C-ABI of include/mstm_hip.h).

The library is loaded from the package directory only; if it is missing the
import of any solver entry point raises: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmstm_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "mstm_hip.h")

FP64, FP32 = 0, 1
FP64_FP32 = 2  # setup_direct: both tables in one pass
SOLVE_PLAIN, SOLVE_MIXED, SOLVE_LEARNED, SOLVE_FGMRES, SOLVE_LU = 0, 1, 2, 3, 4


class MstmHipError(RuntimeError):
    pass


_lib = None

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
_fp = C.POINTER(C.c_float)

_SIGNATURES = {
    "mstm_hip_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "mstm_hip_destroy": (C.c_int, [_vp]),
    "mstm_hip_last_error": (C.c_char_p, [_vp]),
    "mstm_hip_set_spheres": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, C.c_double, C.c_double]),
    "mstm_hip_set_mie": (C.c_int, [_vp, _ip, _dp]),
    "mstm_hip_get_orders": (C.c_int, [_vp, _ip, _ip]),
    "mstm_hip_number_eqns": (C.c_int, [_vp]),
    "mstm_hip_get_mie": (C.c_int, [_vp, _dp]),
    "mstm_hip_setup_direct": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_node_selection": (C.c_int, [_vp, _dp, _dp, _dp, C.c_int, _dp, C.c_int, C.c_int, C.c_int, _ip, _ip]),
    "mstm_hip_fft_grid": (C.c_int, [_vp, _dp, _ip, _ip]),
    "mstm_hip_setup_fft": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_sphereinteraction": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, _ip, C.c_int, C.c_int, _ip, C.c_int]),
    "mstm_hip_sphereinteraction_noalloc": (C.c_int, [_vp, C.c_int, _dp, _dp, _ip, _ip, C.c_int, C.c_int, _ip,
                                                     C.c_int]),
    "mstm_hip_cbicg": (C.c_int, [_vp, C.c_int, C.c_double, _dp, _dp, C.c_int, _ip, _dp]),
    "mstm_hip_fixedorsoln": (C.c_int, [_vp, C.c_double, C.c_double, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                       _dp, _dp, _dp, _ip, _ip, _ip]),
    "mstm_hip_clear": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_translation_orders": (C.c_int, [_vp, _dp, C.c_double, C.c_int, _ip, _ip]),
    "mstm_hip_set_translation_orders": (C.c_int, [_vp, _dp, _ip, C.c_int]),
    "mstm_hip_merge_to_common_origin": (C.c_int, [_vp, _dp, _dp]),
    "mstm_hip_scattering_matrix": (C.c_int, [_vp, C.c_int, _dp, C.c_int, _dp, _dp, _dp]),
    "mstm_hip_asymmetry": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp]),
    "mstm_hip_fosmexpansion": (C.c_int, [_vp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
    "mstm_hip_set_inner_solver": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_set_refinement_cycles": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_set_graphs": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_direct_solver": (C.c_int, [_vp, _dp, _dp, C.c_int, C.c_int, _dp]),
    "mstm_hip_tmatrix_solution": (C.c_int, [_vp, C.c_double, C.c_int, C.c_double, C.c_int, _dp, C.c_int64, _dp, _dp,
                                            _dp, _ip, _ip, _dp, _ip]),
    "mstm_hip_stats": (C.c_int, [_vp, _dp, C.c_int]),
    "mstm_hip_stats_names": (C.c_char_p, []),
    "mstm_hip_lig_nparams": (C.c_int, [C.c_int, C.c_int, C.c_int]),
    "mstm_hip_set_learned": (C.c_int, [_vp, _fp, C.c_int, C.c_int, C.c_int, C.c_int]),
    "mstm_hip_lig_features": (C.c_int, [_vp, _fp, _fp]),
    "mstm_hip_lig_samples": (C.c_int, [_vp, _dp, C.c_int, _fp, _fp]),
    "mstm_hip_lig_guess": (C.c_int, [_vp, _dp]),
    "mstm_hip_trainer_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int, C.c_int, C.c_int, C.c_int, C.c_uint]),
    "mstm_hip_trainer_destroy": (C.c_int, [_vp]),
    "mstm_hip_trainer_last_error": (C.c_char_p, [_vp]),
    "mstm_hip_trainer_params": (C.c_int, [_vp, _fp, C.c_int, C.c_int]),
    "mstm_hip_trainer_set_data": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _fp]),
    "mstm_hip_trainer_epoch": (C.c_int, [_vp, C.c_int, C.c_float, C.c_uint, _dp]),
    "mstm_hip_trainer_loss": (C.c_int, [_vp, C.c_int, _dp]),
    "mstm_hip_trainer_grad": (C.c_int, [_vp, C.c_int, _fp, _fp, _fp, _fp, _fp, _dp]),
    "mstm_hip_trainer_forward": (C.c_int, [_vp, C.c_int, _fp, _fp]),
    "mstm_hip_batch_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int]),
    "mstm_hip_batch_destroy": (C.c_int, [_vp]),
    "mstm_hip_batch_last_error": (C.c_char_p, [_vp]),
    "mstm_hip_batch_set": (C.c_int, [_vp, C.c_int, _ip, _dp, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp]),
    "mstm_hip_batch_setup": (C.c_int, [_vp, C.c_int]),
    "mstm_hip_batch_orders": (C.c_int, [_vp, _ip]),
    "mstm_hip_batch_solve": (C.c_int, [_vp, C.c_double, C.c_int, C.c_int, _ip, _dp]),
    "mstm_hip_batch_vectors": (C.c_int, [_vp, C.c_int, _dp]),
    "mstm_hip_batch_product": (C.c_int, [_vp, C.c_int, C.c_int, _dp, _dp]),
    "mstm_hip_batch_qeff": (C.c_int, [_vp, _dp]),
    "mstm_hip_batch_stats": (C.c_int, [_vp, _dp, C.c_int]),
    "mstm_hip_batch_stats_names": (C.c_char_p, []),
    "mstm_hip_batch_tune_table": (C.c_int, [_vp, _dp, C.c_int]),
}


def header_symbols() -> list[str]:
    """Every function declared in include/mstm_hip.h."""
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*|void)\s*(mstm_hip_\w+)\s*\(", txt, re.M)))


def load_library(path: str | None = None):
    global _lib
    if _lib is not None:
        return _lib
    path = path or LIB_PATH
    if not os.path.exists(path):
        raise MstmHipError(f"{path} is missing: build it with `python -m mstm_amd.build` (no CPU fallback exists)")
    L = C.CDLL(path)
    for name, (res, args) in _SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L
