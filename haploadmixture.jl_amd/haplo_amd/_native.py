"""ctypes bindings of the native libraries (``lib/libhaplo_hip.so``, ``lib/libhaplo_sim.so``).

The HIP library is the product path: if it is missing, every entry point that needs it raises
``NativeLibraryError`` -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_DIR = os.path.join(PKG_ROOT, "lib")
REPO_ROOT = os.path.dirname(PKG_ROOT)
HEADER = os.path.join(REPO_ROOT, "include", "haplo_hip.h")
HIP_LIB = os.path.join(LIB_DIR, "libhaplo_hip.so")
SIM_LIB = os.path.join(LIB_DIR, "libhaplo_sim.so")


class NativeLibraryError(RuntimeError):
    pass


class HipError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hk error {code}: {msg}")
        self.code = code


_i64 = ctypes.c_int64
_int = ctypes.c_int
_dbl = ctypes.c_double
_vp = ctypes.c_void_p


class FitOpts(ctypes.Structure):
    """``hk_fit_opts`` (include/haplo_hip.h)."""

    _fields_ = [
        ("em_iters", _int), ("n_iter", _int), ("rtol", _dbl), ("Qn", _int),
        ("fix_q", _int), ("fix_p", _int), ("q_penalty", _int), ("p_penalty", _int),
        ("rho", _dbl), ("eps", _dbl), ("verbose", _int), ("profile", _int), ("use_graph", _int),
        ("em_fix_q", _int), ("guarantee_increase", _int), ("qn_mode", _int), ("batch_rounds", _int),
    ]


class FitResult(ctypes.Structure):
    """``hk_fit_result`` (include/haplo_hip.h)."""

    _fields_ = [
        ("ll", _dbl), ("ll_init", _dbl), ("iters", _int), ("converged", _int), ("qp_failures", _int),
        ("seconds", _dbl),
        ("t_qpass_ms", _dbl), ("t_ppass_ms", _dbl), ("t_ll_ms", _dbl), ("t_qstep_ms", _dbl),
        ("t_pstep_ms", _dbl), ("t_em_ms", _dbl), ("t_qn_ms", _dbl),
        ("n_qpass", _int), ("n_ppass", _int), ("n_ll", _int), ("n_qstep", _int), ("n_pstep", _int),
        ("n_em", _int), ("n_qn", _int), ("qp_iter_limit", _int),
    ]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


_hip = None
_sim = None


def header_symbols(path: str = HEADER):
    """Function names declared in include/haplo_hip.h."""
    txt = open(path).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(hk_[a-z0-9_]+)\s*\(", txt)))


def hip():
    """Load libhaplo_hip.so (raises NativeLibraryError if it has not been built)."""
    global _hip
    if _hip is not None:
        return _hip
    if not os.path.exists(HIP_LIB):
        raise NativeLibraryError(
            f"{HIP_LIB} not found: build it with `python haploadmixture.jl_amd/build.py` "
            "(there is no CPU fallback for the HIP engine)")
    L = ctypes.CDLL(HIP_LIB)
    L.hk_ctx_create.argtypes = [ctypes.POINTER(_vp), _i64, _i64, _int, _vp, _i64, _i64, _int]
    L.hk_ctx_destroy.argtypes = [_vp]
    L.hk_ctx_create_shared.argtypes = [ctypes.POINTER(_vp), _vp, _int, _int]
    L.hk_fit_batch.argtypes = [ctypes.POINTER(_vp), _int, ctypes.POINTER(FitOpts), _vp, ctypes.POINTER(FitResult), _vp]
    L.hk_autotune.argtypes = [_vp, _int, ctypes.POINTER(_int), ctypes.POINTER(_int)]
    L.hk_last_error.argtypes = [_vp]
    L.hk_last_error.restype = ctypes.c_char_p
    L.hk_create_error.restype = ctypes.c_char_p
    L.hk_set_params.argtypes = [_vp, _vp, _vp]
    L.hk_get_params.argtypes = [_vp, _vp, _vp]
    L.hk_loglik.argtypes = [_vp, ctypes.POINTER(_dbl), _vp]
    L.hk_em.argtypes = [_vp, _int, _vp, _vp]
    L.hk_update_q.argtypes = [_vp, _vp, _vp]
    L.hk_update_p.argtypes = [_vp, _vp, _vp]
    L.hk_double_missing.argtypes = [_vp, _vp]
    L.hk_decode.argtypes = [_vp, _vp]
    L.hk_count_double_missing_file.argtypes = [_vp, _vp, _i64, _i64]
    L.hk_buf_upload.argtypes = [_vp, _int, _vp]
    L.hk_buf_download.argtypes = [_vp, _int, _vp]
    L.hk_buf_copy.argtypes = [_vp, _int, _int]
    L.hk_loglik_buf.argtypes = [_vp, _int, _int, ctypes.POINTER(_dbl)]
    L.hk_em_buf.argtypes = [_vp, _int]
    L.hk_update_q_buf.argtypes = [_vp]
    L.hk_update_p_buf.argtypes = [_vp]
    L.hk_loglik_noalloc.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.hk_q_step.argtypes = [_vp, _int, _dbl, _vp]
    L.hk_p_step.argtypes = [_vp, _vp, _int, _dbl, _vp]
    L.hk_fit_opts_default.argtypes = [ctypes.POINTER(FitOpts)]
    L.hk_fit_opts_default.restype = None
    L.hk_fit.argtypes = [_vp, ctypes.POINTER(FitOpts), _dbl, _vp, ctypes.POINTER(FitResult)]
    L.hk_fit_ex.argtypes = [_vp, ctypes.POINTER(FitOpts), _dbl, _vp, _vp, ctypes.POINTER(FitResult)]
    L.hk_set_tuning.argtypes = [_vp, _int, _int]
    L.hk_get_tuning.argtypes = [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]
    L.hk_set_pass_shape.argtypes = [_vp, _int, _int]
    L.hk_get_pass_shape.argtypes = [_vp, ctypes.POINTER(_int), ctypes.POINTER(_int)]
    L.hk_sync.argtypes = [_vp]
    L.hk_time_pass.argtypes = [_vp, _int, _int, ctypes.POINTER(_dbl)]
    L.hk_copy_bench.argtypes = [_vp, _i64, _int, ctypes.POINTER(_dbl)]
    L.hk_features.argtypes = [_vp, _int, ctypes.c_uint64, _vp]
    L.hk_skfr.argtypes = [_vp, _int, _int, _int, _int, _int, ctypes.c_uint64, _vp, _vp, _vp, ctypes.POINTER(_int),
                          ctypes.POINTER(_dbl), ctypes.POINTER(_int), _vp]
    L.hk_skfr_path.argtypes = [_vp, _int, _int, _vp, _int, _int, _int, ctypes.c_uint64, _vp, _vp, _vp, _vp, _vp]
    L.hk_mlp_create.argtypes = [ctypes.POINTER(_vp), _int, _int, _int, ctypes.c_uint64]
    L.hk_mlp_destroy.argtypes = [_vp]
    L.hk_mlp_last_error.argtypes = [_vp]
    L.hk_mlp_last_error.restype = ctypes.c_char_p
    L.hk_mlp_num_params.argtypes = [_vp, ctypes.POINTER(_i64), _vp]
    L.hk_mlp_get_params.argtypes = [_vp, _vp, _vp]
    L.hk_mlp_set_params.argtypes = [_vp, _vp]
    L.hk_mlp_load_data.argtypes = [_vp, _vp, _vp, _i64]
    L.hk_mlp_forward.argtypes = [_vp, _vp, _i64, _vp]
    L.hk_mlp_predict.argtypes = [_vp, _vp, _int, ctypes.c_uint64, _vp]
    L.hk_mlp_grad.argtypes = [_vp, _i64, _i64, _int, ctypes.POINTER(ctypes.c_float), _vp]
    L.hk_mlp_adam.argtypes = [_vp, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _int]
    _hip = L
    return L


def sim():
    global _sim
    if _sim is not None:
        return _sim
    if not os.path.exists(SIM_LIB):
        raise NativeLibraryError(f"{SIM_LIB} not found: build it with `python haploadmixture.jl_amd/build.py`")
    L = ctypes.CDLL(SIM_LIB)
    L.hk_sim_bed.argtypes = [_i64, _i64, _int, _dbl, _dbl, _dbl, ctypes.c_uint64, _vp, _vp, _vp]
    L.hk_sim_init.argtypes = [_i64, _i64, _int, ctypes.c_uint64, _vp, _vp]
    L.hk_sim_bed_world.argtypes = [_i64, _i64, _int, _dbl, _dbl, _dbl, ctypes.c_uint64, ctypes.c_uint64, _vp, _vp, _vp]
    L.hk_sim_threads.restype = _int
    _sim = L
    return L


def ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(_vp)
