"""FP64 CPU oracle for the DTS / HFBS hot path (ctypes wrapper).

TEST INFRASTRUCTURE ONLY.  The product path (the HIP library and the package
``domain_transform_solver_amd``) never imports this module; only the tests,
``__graft_entry__.smoke()`` and the CPU-baseline leg of ``bench.py`` use it,
and only as the reference results the GPU results are compared with (and, in
``bench.py``, as the timed CPU baseline).

The wrapper mirrors the reference's class flow
(src/domain_transform_solver.h:34-85, src/domain_transform_stereo.h:36-57,
src/hfbs.h:99-150) so a test can drive the oracle and the GPU library with the
same call sequence.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# DTS_ORACLE_LIB selects another build of the same sources (the sanitizer
# build of `make san`, loaded with libasan preloaded by the test that runs it).
_LIB_PATH = os.environ.get("DTS_ORACLE_LIB") or os.path.join(_HERE, "liboracle_dts.so")
_lib = None

RGB, YCBCR, YYY = 0, 1, 2
L1, L2, CHARBONNIER = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)
_i64p = ctypes.POINTER(ctypes.c_int64)
_u8p = ctypes.POINTER(ctypes.c_uint8)


class HfbsParams(ctypes.Structure):
    """Mirror of HFBSGridParams minus the derived grid sizes (src/hfbs.h:18-60)."""

    _fields_ = [
        ("sigma_x", ctypes.c_float),
        ("sigma_y", ctypes.c_float),
        ("sigma_l", ctypes.c_float),
        ("lambda_", ctypes.c_float),
        ("alpha", ctypes.c_float),
        ("beta", ctypes.c_float),
        ("bistoch_iter_max", ctypes.c_int),
        ("optim_iter_max", ctypes.c_int),
        ("neg_exp_beta", ctypes.c_float),
    ]


def hfbs_params(sigma_x=8.0, sigma_y=8.0, sigma_l=4.0, lambda_=8.0, alpha=1.0,
                beta=0.9, bistoch_iter_max=20, optim_iter_max=140,
                neg_exp_beta=0.9) -> HfbsParams:
    """Defaults are the reference's HFBSGridParams (src/hfbs.h:20-52)."""
    return HfbsParams(sigma_x, sigma_y, sigma_l, lambda_, alpha, beta,
                      bistoch_iter_max, optim_iter_max, neg_exp_beta)


def build(force: bool = False) -> str:
    if os.environ.get("DTS_ORACLE_LIB"):
        return _LIB_PATH
    if force or not os.path.exists(_LIB_PATH):
        subprocess.check_call(["make", "-s", "-B" if force else "all"], cwd=_HERE)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        vp = ctypes.c_void_p
        f = ctypes.c_float
        i = ctypes.c_int
        sig = {
            "orc_dts_create": (vp, [i, i]),
            "orc_dts_destroy": (None, [vp]),
            "orc_dts_set_image_dim": (i, [vp, i, i]),
            "orc_dts_set_threads": (None, [i]),
            "orc_dts_clear_all": (i, [vp]),
            "orc_dts_init_frame": (i, [vp, i, _u8p, _fp, _fp]),
            "orc_dts_init_right_frame": (i, [vp, f, i, _u8p]),
            "orc_dts_compute_differential": (i, [vp, f, f, f]),
            "orc_dts_integrate": (i, [vp]),
            "orc_dts_process_right_frame": (i, [vp]),
            "orc_dts_filter": (i, [vp, f, f, f, _dp, _dp, i]),
            "orc_dts_optimize": (i, [vp, i, i, f, f, f, f]),
            "orc_dts_optimize_stereo": (i, [vp, i, i, f, f, f, f, f]),
            "orc_dts_optimize_adjoint": (i, [vp, i, i, f, f, f, f, _dp, _dp, _dp]),
            "orc_dts_set_target_conf_f64": (i, [vp, _dp, _dp]),
            "orc_dts_set_guide_f64": (i, [vp, _dp, f, f, f]),
            "orc_dts_compute_confidence": (i, [vp, f, f, f, f, i]),
            "orc_dts_result": (_dp, [vp]),
            "orc_dts_confidence": (_dp, [vp]),
            "orc_dts_target": (_dp, [vp]),
            "orc_dts_dHdx": (_dp, [vp]),
            "orc_dts_dVdy": (_dp, [vp]),
            "orc_dts_ctH_fx": (_i64p, [vp]),
            "orc_dts_ctV_fx": (_i64p, [vp]),
            "orc_dts_color": (_u8p, [vp]),
            "orc_line_bounds": (None, [_i64p, i, f, _ip, _ip]),
            "orc_dts_bounds": (i, [vp, f, i, _ip, _ip]),
            "orc_to_fx": (ctypes.c_int64, [ctypes.c_double]),
            "orc_sigma_fx": (ctypes.c_int64, [f]),
            "orc_hfbs_optimize": (i, [_u8p, i, i, _fp, _fp,
                                      ctypes.POINTER(HfbsParams), i, i, i, _dp, _ip]),
            "orc_hfbs_grids": (i, [_u8p, i, i, _fp, _fp, ctypes.POINTER(HfbsParams),
                                   _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
            "orc_hfbs_solve": (i, [i, i, i, ctypes.POINTER(HfbsParams), _dp, _dp, _dp, _dp,
                                   _dp]),
            "orc_hfbs_slice": (i, [_u8p, i, i, ctypes.POINTER(HfbsParams), i, i, i, _dp, _dp,
                                   _dp]),
            "orc_laplace_line": (None, [_dp, _dp, i, i, ctypes.c_double]),
            "orc_convert_color": (None, [_u8p, i, i, i]),
            "orc_rf_sigma_i": (ctypes.c_double, [ctypes.c_double, i, i]),
            "orc_rf_line": (None, [_dp, _dp, _dp, i]),
            "orc_rf_filter": (i, [_dp, i, i, f, f, f, i, _dp, _dp]),
            "orc_rf_filter_adjoint": (i, [_dp, i, i, f, f, f, i, _dp, _dp, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def set_threads(n: int) -> None:
    lib().orc_dts_set_threads(int(n))


def _ptr(a: np.ndarray | None, ctype):
    if a is None:
        return None
    return a.ctypes.data_as(ctypes.POINTER(ctype))


def _f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


class OracleDTS:
    """FP64 restatement of DomainTransformSolver / DomainTransformStereo."""

    def __init__(self, max_w: int, max_h: int):
        self._h = lib().orc_dts_create(int(max_w), int(max_h))
        if not self._h:
            raise ValueError("invalid max dims")
        self.w, self.h = int(max_w), int(max_h)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_dts_destroy(h)
            self._h = None

    def set_image_dim(self, w: int, h: int) -> int:
        r = lib().orc_dts_set_image_dim(self._h, int(w), int(h))
        if r == 0:
            self.w, self.h = int(w), int(h)
        return r

    def clear_all(self):
        lib().orc_dts_clear_all(self._h)

    def init_frame(self, color_space: int, rgba: np.ndarray, target: np.ndarray,
                   confidence: np.ndarray | None = None):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        assert rgba.size == self.w * self.h * 4
        t = _f32(target)
        c = _f32(confidence)
        lib().orc_dts_init_frame(self._h, color_space, _ptr(rgba, ctypes.c_uint8),
                                 _ptr(t, ctypes.c_float), _ptr(c, ctypes.c_float))

    def init_right_frame(self, doff: float, color_space: int, rgba: np.ndarray):
        rgba = np.ascontiguousarray(rgba, dtype=np.uint8)
        lib().orc_dts_init_right_frame(self._h, float(doff), color_space,
                                       _ptr(rgba, ctypes.c_uint8))

    def compute_differential(self, sigma_x, sigma_y, sigma_r):
        lib().orc_dts_compute_differential(self._h, sigma_x, sigma_y, sigma_r)

    def integrate(self):
        lib().orc_dts_integrate(self._h)

    def process_right_frame(self):
        lib().orc_dts_process_right_frame(self._h)

    def filter(self, sigma_x, sigma_y, sigma_r, inp: np.ndarray, num_iter: int = 3):
        x = np.ascontiguousarray(inp, dtype=np.float64)
        out = np.empty((self.h, self.w), np.float64)
        lib().orc_dts_filter(self._h, sigma_x, sigma_y, sigma_r, _ptr(x, ctypes.c_double),
                             _ptr(out, ctypes.c_double), int(num_iter))
        return out

    def optimize(self, num_iter, loss, lam, sigma_z, step, overwrite=-1.0):
        lib().orc_dts_optimize(self._h, int(num_iter), int(loss), lam, sigma_z, step,
                               overwrite)

    def optimize_adjoint(self, num_iter, loss, lam, sigma_z, step, overwrite, ybar):
        """(tbar, cbar): derivative of <ybar, result> w.r.t. target and
        confidence (re-runs the forward; the result is left in the state)."""
        yb = np.ascontiguousarray(ybar, dtype=np.float64)
        tb = np.empty((self.h, self.w), np.float64)
        cb = np.empty((self.h, self.w), np.float64)
        lib().orc_dts_optimize_adjoint(self._h, int(num_iter), int(loss), lam, sigma_z, step,
                                       overwrite, _ptr(yb, ctypes.c_double),
                                       _ptr(tb, ctypes.c_double), _ptr(cb, ctypes.c_double))
        return tb, cb

    def set_target_conf_f64(self, target, confidence):
        t = np.ascontiguousarray(target, dtype=np.float64)
        c = np.ascontiguousarray(confidence, dtype=np.float64)
        lib().orc_dts_set_target_conf_f64(self._h, _ptr(t, ctypes.c_double),
                                          _ptr(c, ctypes.c_double))

    def set_guide_f64(self, guide, sigma_x, sigma_y, sigma_r):
        """Differentials and ct from a real-valued RGBA guide (H x W x 4)."""
        g = np.ascontiguousarray(guide, dtype=np.float64)
        lib().orc_dts_set_guide_f64(self._h, _ptr(g, ctypes.c_double), sigma_x, sigma_y,
                                    sigma_r)

    def optimize_stereo(self, num_iter, loss, lam, sigma_z, step, photo_lambda,
                        overwrite=-1.0):
        lib().orc_dts_optimize_stereo(self._h, int(num_iter), int(loss), lam, sigma_z,
                                      step, photo_lambda, overwrite)

    def compute_confidence(self, sigma_x, sigma_y, sigma_r, sigma0, left_clear):
        lib().orc_dts_compute_confidence(self._h, sigma_x, sigma_y, sigma_r, sigma0,
                                         int(left_clear))

    def _plane(self, fn, dtype, ch=1):
        p = fn(self._h)
        n = self.w * self.h * ch
        arr = np.ctypeslib.as_array(p, shape=(n,)).copy()
        return arr.astype(dtype, copy=False).reshape(
            (self.h, self.w) if ch == 1 else (self.h, self.w, ch))

    def result(self):
        return self._plane(lib().orc_dts_result, np.float64)

    def confidence(self):
        return self._plane(lib().orc_dts_confidence, np.float64)

    def target(self):
        return self._plane(lib().orc_dts_target, np.float64)

    def dHdx(self):
        return self._plane(lib().orc_dts_dHdx, np.float64)

    def dVdy(self):
        return self._plane(lib().orc_dts_dVdy, np.float64)

    def ctH_fx(self):
        return self._plane(lib().orc_dts_ctH_fx, np.int64)

    def ctV_fx(self):
        return self._plane(lib().orc_dts_ctV_fx, np.int64)

    def color(self):
        return self._plane(lib().orc_dts_color, np.uint8, 4)

    def bounds(self, sigma: float, vertical: bool):
        lo = np.empty((self.h, self.w), np.int32)
        up = np.empty((self.h, self.w), np.int32)
        lib().orc_dts_bounds(self._h, sigma, int(bool(vertical)), _ptr(lo, ctypes.c_int),
                             _ptr(up, ctypes.c_int))
        return lo, up


def line_bounds(ct_fx: np.ndarray, sigma: float):
    ct = np.ascontiguousarray(ct_fx, dtype=np.int64)
    L = ct.size
    lo = np.empty(L, np.int32)
    up = np.empty(L, np.int32)
    lib().orc_line_bounds(_ptr(ct, ctypes.c_int64), L, sigma, _ptr(lo, ctypes.c_int),
                          _ptr(up, ctypes.c_int))
    return lo, up


def to_fx(d: float) -> int:
    return int(lib().orc_to_fx(float(d)))


def sigma_fx(s: float) -> int:
    return int(lib().orc_sigma_fx(float(s)))


def convert_color(rgba: np.ndarray, color_space: int) -> np.ndarray:
    a = np.ascontiguousarray(rgba, dtype=np.uint8).copy()
    h, w = a.shape[:2]
    lib().orc_convert_color(_ptr(a, ctypes.c_uint8), w, h, color_space)
    return a


def hfbs_optimize(rgba_converted: np.ndarray, target: np.ndarray,
                  confidence: np.ndarray | None, params: HfbsParams,
                  max_grid=(1 << 30, 1 << 30, 1 << 30)):
    """One HFBS::Optimize on an already colour-converted guide.

    Returns (out float64 HxW, grid_dims (gx, gy, gl), fell_back)."""
    g = np.ascontiguousarray(rgba_converted, dtype=np.uint8)
    h, w = g.shape[:2]
    t = _f32(target)
    c = _f32(confidence)
    out = np.empty((h, w), np.float64)
    dims = np.zeros(3, np.int32)
    r = lib().orc_hfbs_optimize(_ptr(g, ctypes.c_uint8), w, h, _ptr(t, ctypes.c_float),
                                _ptr(c, ctypes.c_float), ctypes.byref(params),
                                int(max_grid[0]), int(max_grid[1]), int(max_grid[2]),
                                _ptr(out, ctypes.c_double), _ptr(dims, ctypes.c_int))
    return out, tuple(int(v) for v in dims), bool(r)


def hfbs_grids(rgba_converted, target, confidence, params: HfbsParams):
    """Intermediate HFBS grids, each shaped (gl, gy, gx)."""
    g = np.ascontiguousarray(rgba_converted, dtype=np.uint8)
    h, w = g.shape[:2]
    gx = int(np.ceil(np.float32(w) / np.float32(params.sigma_x)))
    gy = int(np.ceil(np.float32(h) / np.float32(params.sigma_y)))
    gl = int(np.ceil(255.0 / float(np.float32(params.sigma_l))))
    names = ["b", "splat_c", "splat_sum", "splat_norm", "inv_diag_a", "z_init", "z_final"]
    arrs = {k: np.zeros((gl, gy, gx), np.float64) for k in names}
    t = _f32(target)
    c = _f32(confidence)
    lib().orc_hfbs_grids(_ptr(g, ctypes.c_uint8), w, h, _ptr(t, ctypes.c_float),
                         _ptr(c, ctypes.c_float), ctypes.byref(params),
                         *[_ptr(arrs[k], ctypes.c_double) for k in names])
    return arrs


def hfbs_solve(b, splat_norm, inv_diag_a, z_init, params: HfbsParams) -> np.ndarray:
    """params.optim_iter_max heavy-ball steps from z_init (zero momentum) on
    grids shaped (gl, gy, gx)."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (b, splat_norm, inv_diag_a, z_init)]
    gl, gy, gx = arrs[0].shape
    assert all(a.shape == (gl, gy, gx) for a in arrs)
    out = np.empty_like(arrs[0])
    r = lib().orc_hfbs_solve(gx, gy, gl, ctypes.byref(params),
                             *[_ptr(a, ctypes.c_double) for a in arrs],
                             _ptr(out, ctypes.c_double))
    assert r == 0
    return out


def hfbs_slice(rgba_converted, splat_sum, z, params: HfbsParams) -> np.ndarray:
    """Splat-count weighted trilinear slice of grid z (gl, gy, gx) at the
    pixels of a colour-converted guide (H x W x 4)."""
    g = np.ascontiguousarray(rgba_converted, dtype=np.uint8)
    h, w = g.shape[:2]
    S = np.ascontiguousarray(splat_sum, dtype=np.float64)
    Z = np.ascontiguousarray(z, dtype=np.float64)
    gl, gy, gx = S.shape
    assert Z.shape == S.shape
    out = np.empty((h, w), np.float64)
    r = lib().orc_hfbs_slice(_ptr(g, ctypes.c_uint8), w, h, ctypes.byref(params), gx, gy, gl,
                             _ptr(S, ctypes.c_double), _ptr(Z, ctypes.c_double),
                             _ptr(out, ctypes.c_double))
    assert r == 0
    return out


def laplace_line(src: np.ndarray, beta: float) -> np.ndarray:
    s = np.ascontiguousarray(src, dtype=np.float64)
    d = np.empty_like(s)
    lib().orc_laplace_line(_ptr(s, ctypes.c_double), _ptr(d, ctypes.c_double), s.size, 1,
                           float(beta))
    return d


# ---- RF-DT (recursive-filter domain transform; not in the reference) -------------------

def rf_sigma_i(sigma: float, i: int, n: int) -> float:
    return float(lib().orc_rf_sigma_i(float(sigma), int(i), int(n)))


def rf_line(x: np.ndarray, wt: np.ndarray) -> np.ndarray:
    """One line, causal then anticausal recursion; wt[k] joins k-1 and k."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    wt = np.ascontiguousarray(wt, dtype=np.float64)
    z = np.empty_like(x)
    lib().orc_rf_line(_ptr(x, ctypes.c_double), _ptr(wt, ctypes.c_double),
                      _ptr(z, ctypes.c_double), x.size)
    return z


def _guide_f64(guide: np.ndarray) -> np.ndarray:
    g = np.ascontiguousarray(guide, dtype=np.float64)
    assert g.ndim == 3 and g.shape[2] == 4
    return g


def rf_filter(guide, sigma_x, sigma_y, sigma_r, num_iter, inp) -> np.ndarray:
    """RF-DT of inp (H x W) on a real RGBA guide (H x W x 4, the colour-converted
    guide for 8-bit input)."""
    g = _guide_f64(guide)
    h, w = g.shape[:2]
    x = np.ascontiguousarray(inp, dtype=np.float64)
    out = np.empty((h, w), np.float64)
    r = lib().orc_rf_filter(_ptr(g, ctypes.c_double), w, h, sigma_x, sigma_y, sigma_r,
                            int(num_iter), _ptr(x, ctypes.c_double), _ptr(out, ctypes.c_double))
    assert r == 0
    return out


def rf_filter_adjoint(guide, sigma_x, sigma_y, sigma_r, num_iter, inp, ybar):
    """(xbar, gbar): derivatives of <ybar, rf_filter(...)> with respect to the
    input (H x W) and the guide (H x W x 4)."""
    g = _guide_f64(guide)
    h, w = g.shape[:2]
    x = np.ascontiguousarray(inp, dtype=np.float64)
    yb = np.ascontiguousarray(ybar, dtype=np.float64)
    xb = np.empty((h, w), np.float64)
    gb = np.empty((h, w, 4), np.float64)
    r = lib().orc_rf_filter_adjoint(_ptr(g, ctypes.c_double), w, h, sigma_x, sigma_y, sigma_r,
                                    int(num_iter), _ptr(x, ctypes.c_double),
                                    _ptr(yb, ctypes.c_double), _ptr(xb, ctypes.c_double),
                                    _ptr(gb, ctypes.c_double))
    assert r == 0
    return xb, gb
