"""ctypes binding of libdts_amd.so (the C ABI declared in include/dts_capi.h).

Loading fails loudly when the library has not been built: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libdts_amd.so")

DTS_OK = 0

c_int, c_float, c_double, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_void_p
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class FilterParams(ctypes.Structure):
    _fields_ = [("sigma_x", c_float), ("sigma_y", c_float), ("sigma_r", c_float)]


class OptimizeParams(ctypes.Structure):
    _fields_ = [("num_iterations", c_int), ("loss", c_int), ("lambda_", c_float),
                ("sigma_z", c_float), ("step_size", c_float),
                ("photo_consis_lambda", c_float), ("DT_lambda", c_float)]


class SolveOptions(ctypes.Structure):
    _fields_ = [("filter_options", FilterParams), ("optim_options", OptimizeParams),
                ("conf_params", FilterParams), ("conf_sigma0", c_float),
                ("overwrite_target_above_conf", c_float)]


class HfbsGridParams(ctypes.Structure):
    _fields_ = [("sigma_x", c_float), ("sigma_y", c_float), ("sigma_l", c_float),
                ("lambda_", c_float), ("alpha", c_float), ("beta", c_float),
                ("bistoch_iter_max", c_int), ("optim_iter_max", c_int),
                ("neg_exp_beta", c_float), ("grid_x", c_int), ("grid_y", c_int),
                ("grid_l", c_int)]


class BatchFrame(ctypes.Structure):
    """Mirror of dts_batch_frame (include/dts_capi.h)."""
    _fields_ = [("width", c_int), ("height", c_int), ("color_space", c_int),
                ("rgba", c_void_p), ("target", c_void_p), ("confidence", c_void_p),
                ("filter_options", FilterParams), ("optim_options", OptimizeParams),
                ("overwrite_target_above_conf", c_float)]


_P = ctypes.POINTER
_SIGNATURES = {
    "dts_last_error": (ctypes.c_char_p, []),
    "dts_version": (c_int, []),
    "dts_create": (c_int, [c_int, c_int, c_int, _P(c_void_p)]),
    "dts_destroy": (c_int, [c_void_p]),
    "dts_get_stream": (c_int, [c_void_p, _P(c_void_p)]),
    "dts_set_stream": (c_int, [c_void_p, c_void_p]),
    "dts_synchronize": (c_int, [c_void_p]),
    "dts_set_image_dim": (c_int, [c_void_p, c_int, c_int]),
    "dts_get_image_dim": (c_int, [c_void_p, _ip, _ip]),
    "dts_clear_all": (c_int, [c_void_p]),
    "dts_init_frame": (c_int, [c_void_p, c_int, c_void_p, _fp, _fp]),
    "dts_init_guide": (c_int, [c_void_p, c_int, c_void_p]),
    "dts_init_frame_device": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "dts_init_right_frame": (c_int, [c_void_p, c_float, c_int, c_void_p]),
    "dts_process_right_frame": (c_int, [c_void_p]),
    "dts_compute_differential": (c_int, [c_void_p, _P(FilterParams)]),
    "dts_integrate": (c_int, [c_void_p]),
    "dts_filter": (c_int, [c_void_p, _P(FilterParams), _fp, _fp, c_int]),
    "dts_filter_device": (c_int, [c_void_p, _P(FilterParams), c_void_p, c_void_p, c_int]),
    "dts_optimize": (c_int, [c_void_p, _P(OptimizeParams), c_float]),
    "dts_optimize_stereo": (c_int, [c_void_p, _P(OptimizeParams), c_float]),
    "dts_optimize_checkpointed": (c_int, [c_void_p, _P(OptimizeParams), c_float, c_int]),
    "dts_optimize_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "dts_rf_filter": (c_int, [c_void_p, _P(FilterParams), c_void_p, c_void_p, c_int, c_int,
                              c_int]),
    "dts_rf_filter_adjoint": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int]),
    "dts_compute_confidence": (c_int, [c_void_p, _P(FilterParams), c_float, c_int]),
    "dts_solve": (c_int, [c_void_p, _P(SolveOptions), c_int]),
    "dts_solve_noalloc": (c_int, [c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                  _P(SolveOptions), c_int, c_void_p, c_void_p]),
    "dts_download": (c_int, [c_void_p, c_int, c_void_p]),
    "dts_result_device": (c_int, [c_void_p, _P(c_void_p)]),
    "dts_debug_bounds": (c_int, [c_void_p, c_float, c_int, _ip, _ip]),
    "dts_debug_ct": (c_int, [c_void_p, c_int, c_void_p]),
    "dts_debug_stream": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 ctypes.c_size_t, c_int, _fp]),
    "dts_set_graph_mode": (c_int, [c_void_p, c_int]),
    "dts_timer_start": (c_int, [c_void_p]),
    "dts_timer_stop": (c_int, [c_void_p, _fp]),
    "dts_profile_passes": (c_int, [c_void_p, c_int]),
    "dts_profile_read": (c_int, [c_void_p, _P(c_double), _ip, _P(c_double), _ip]),
    "dts_profile_read_fused": (c_int, [c_void_p, _P(c_double), _ip]),
    "dts_set_fused": (c_int, [c_void_p, c_int]),
    "dts_batch_create": (c_int, [c_int, ctypes.c_size_t, _P(c_void_p)]),
    "dts_batch_destroy": (c_int, [c_void_p]),
    "dts_batch_set_frames": (c_int, [c_void_p, _P(BatchFrame), c_int, c_int]),
    "dts_batch_set_tuning": (c_int, [c_void_p, c_int, c_int]),
    "dts_batch_get_tuning": (c_int, [c_void_p, _P(c_int), _P(c_int)]),
    "dts_batch_set_fused": (c_int, [c_void_p, c_int]),
    "dts_batch_set_extent_planning": (c_int, [c_void_p, c_int]),
    "dts_batch_get_frame_extents": (c_int, [c_void_p, c_int, _ip]),
    "dts_batch_get_fused": (c_int, [c_void_p, _P(c_int), _P(c_int)]),
    "dts_batch_set_fused_waves": (c_int, [c_void_p, c_int]),
    "dts_batch_get_fused_waves": (c_int, [c_void_p, _P(c_int)]),
    "dts_batch_autotune": (c_int, [c_void_p, c_int, _P(c_float)]),
    "dts_batch_solve": (c_int, [c_void_p]),
    "dts_batch_set_graph_mode": (c_int, [c_void_p, c_int]),
    "dts_batch_download": (c_int, [c_void_p, c_int, _fp]),
    "dts_batch_result_device": (c_int, [c_void_p, c_int, _P(c_void_p)]),
    "dts_batch_synchronize": (c_int, [c_void_p]),
    "dts_batch_timer_start": (c_int, [c_void_p]),
    "dts_batch_timer_stop": (c_int, [c_void_p, _fp]),
    "hfbs_create": (c_int, [c_int, c_int, c_int, c_int, c_int, _P(c_void_p)]),
    "hfbs_destroy": (c_int, [c_void_p]),
    "hfbs_get_stream": (c_int, [c_void_p, _P(c_void_p)]),
    "hfbs_set_stream": (c_int, [c_void_p, c_void_p]),
    "hfbs_synchronize": (c_int, [c_void_p]),
    "hfbs_set_image_dim": (c_int, [c_void_p, c_int, c_int]),
    "hfbs_set_params": (c_int, [c_void_p, _P(HfbsGridParams)]),
    "hfbs_default_params": (c_int, [_P(HfbsGridParams)]),
    "hfbs_init_frame": (c_int, [c_void_p, c_int, c_void_p, _fp, _fp]),
    "hfbs_init_frame_device": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    "hfbs_clear_all": (c_int, [c_void_p]),
    "hfbs_optimize": (c_int, [c_void_p, _ip]),
    "hfbs_download": (c_int, [c_void_p, c_int, _fp]),
    "hfbs_result_device": (c_int, [c_void_p, _P(c_void_p)]),
    "hfbs_debug_grid_dims": (c_int, [c_void_p, _ip, _ip, _ip]),
    "hfbs_debug_grid": (c_int, [c_void_p, c_int, _fp]),
    "hfbs_set_graph_mode": (c_int, [c_void_p, c_int]),
    "hfbs_timer_start": (c_int, [c_void_p]),
    "hfbs_timer_stop": (c_int, [c_void_p, _fp]),
}

_lib = None


class DtsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dts error {code}: {msg}")
        self.code = code


def load():
    """Load libdts_amd.so; raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m domain_transform_solver_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in _SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def exported_symbols():
    return list(_SIGNATURES)


def check(rc: int) -> int:
    if rc != DTS_OK:
        raise DtsError(rc, load().dts_last_error().decode())
    return rc
