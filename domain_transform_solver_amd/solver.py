"""Python mirror of the reference's solver classes over the C ABI.

Method names, argument meaning and call order follow the reference's C++ API
(src/domain_transform_filter.h:19-70, src/domain_transform_solver.h:34-85,
src/domain_transform_stereo.h:36-57, src/hfbs.h:99-150,
src/edge_aware_solver.h:13-90) so that tests read like the reference's own
usage.  Invalid dims keep the previous size (the reference prints and
ignores them); every other failure raises DtsError.  All compute runs in
libdts_amd.so on the GPU.
"""
from __future__ import annotations

import ctypes
import hashlib
import json
import os
from dataclasses import dataclass, field

import numpy as np

from . import _lib
from ._lib import FilterParams, HfbsGridParams, OptimizeParams, SolveOptions, check

RGB, YCbCr, YYY = 0, 1, 2
L1, L2, CHARBONNIER = 0, 1, 2
COLOR_IMAGE, DIFFERENTIAL, INTEGRAL, TARGET, CONFIDENCE, OPTIMIZED_QUANTITY = range(6)


@dataclass
class DomainFilterParams:
    """src/domain_transform_filter_params.h:8-31 (sigmas[0] used by confidence)."""
    sigma_x: float = 8.0
    sigma_y: float = 8.0
    sigma_r: float = 0.2
    sigmas: list = field(default_factory=lambda: [16.0])

    def c(self) -> FilterParams:
        return FilterParams(self.sigma_x, self.sigma_y, self.sigma_r)


@dataclass
class DomainOptimizeParams:
    """src/domain_transform_optimize.h:30-50 (same defaults)."""
    num_iterations: int = 3
    loss: int = L2
    lambda_: float = 1.0
    sigma_z: float = 100.0
    step_size: float = 0.99
    photo_consis_lambda: float = 1.0
    DT_lambda: float = 0.2

    def c(self) -> OptimizeParams:
        return OptimizeParams(self.num_iterations, self.loss, self.lambda_, self.sigma_z,
                              self.step_size, self.photo_consis_lambda, self.DT_lambda)


@dataclass
class DtsOptions:
    """src/edge_aware_solver.h:13-21."""
    filter_options: DomainFilterParams = field(default_factory=DomainFilterParams)
    optim_options: DomainOptimizeParams = field(default_factory=DomainOptimizeParams)
    conf_params: DomainFilterParams = field(default_factory=DomainFilterParams)
    color_space: int = RGB
    image_dims: tuple = (100, 200)
    overwrite_target_above_conf: float = -1.0

    def c(self) -> SolveOptions:
        return SolveOptions(self.filter_options.c(), self.optim_options.c(),
                            self.conf_params.c(), float(self.conf_params.sigmas[0]),
                            float(self.overwrite_target_above_conf))


def _u8(a):
    a = np.ascontiguousarray(a, dtype=np.uint8)
    return a, a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    if a is None:
        return None, None
    a = np.ascontiguousarray(a, dtype=np.float32)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


class DomainTransformSolver:
    """DomainTransformFilter + DomainTransformSolver on one C-ABI handle."""

    _flags = 0

    def __init__(self, max_width: int, max_height: int):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.dts_create(int(max_width), int(max_height), self._flags,
                                   ctypes.byref(h)))
        self._h = h
        self.width, self.height = int(max_width), int(max_height)

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dts_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def SetImageDim(self, width: int, height: int) -> bool:
        rc = self._lib.dts_set_image_dim(self._h, int(width), int(height))
        if rc == 0:
            self.width, self.height = int(width), int(height)
        return rc == 0

    def ClearAll(self):
        check(self._lib.dts_clear_all(self._h))

    def InitFrame(self, color_space, rgba, target, confidence=None):
        g, gp = _u8(rgba)
        assert g.size == 4 * self.width * self.height, "guide size != image dims"
        t, tp = _f32(target)
        c, cp = _f32(confidence)
        check(self._lib.dts_init_frame(self._h, int(color_space), gp, tp, cp))

    def ComputeColorSpaceDifferential(self, p: DomainFilterParams):
        fp = p.c()
        check(self._lib.dts_compute_differential(self._h, ctypes.byref(fp)))

    def IntegrateColorDifferentials(self):
        check(self._lib.dts_integrate(self._h))

    def Filter(self, p: DomainFilterParams, inp, num_iter: int = 3):
        x, xp = _f32(inp)
        out = np.empty((self.height, self.width), np.float32)
        fp = p.c()
        check(self._lib.dts_filter(self._h, ctypes.byref(fp), xp,
                                   out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                   int(num_iter)))
        return out

    def Optimize(self, p: DomainOptimizeParams, overwrite_target_above_conf: float = -1.0):
        op = p.c()
        check(self._lib.dts_optimize(self._h, ctypes.byref(op),
                                     float(overwrite_target_above_conf)))

    def OptimizeCheckpointed(self, p: DomainOptimizeParams,
                             overwrite_target_above_conf: float = -1.0, ckpt_every: int = 0):
        """Optimize that keeps checkpoints for OptimizeAdjoint (same result)."""
        op = p.c()
        check(self._lib.dts_optimize_checkpointed(self._h, ctypes.byref(op),
                                                  float(overwrite_target_above_conf),
                                                  int(ckpt_every)))

    def OptimizeAdjoint(self, ybar, with_guide: bool = False):
        """(tbar, cbar[, gbar]): derivatives of <ybar, result> with respect to
        the target, the confidence (and the guide, zero almost everywhere)."""
        yb = np.ascontiguousarray(ybar, np.float32)
        tb = np.empty((self.height, self.width), np.float32)
        cb = np.empty((self.height, self.width), np.float32)
        gb = np.empty((self.height, self.width, 4), np.float32) if with_guide else None
        check(self._lib.dts_optimize_adjoint(self._h, yb.ctypes.data, tb.ctypes.data,
                                             cb.ctypes.data,
                                             None if gb is None else gb.ctypes.data, 0))
        return (tb, cb, gb) if with_guide else (tb, cb)

    def OptimizeAdjointDevice(self, ybar_ptr: int, tbar_ptr: int, cbar_ptr: int,
                              gbar_ptr: int | None = None):
        """OptimizeAdjoint on device pointers (the handle's stream)."""
        check(self._lib.dts_optimize_adjoint(self._h, ybar_ptr, tbar_ptr, cbar_ptr,
                                             gbar_ptr, 1))

    def RFFilter(self, p: DomainFilterParams, inp, num_iter: int = 3, keep_tape: bool = False):
        """Recursive-filter domain transform of inp on the current guide (not in
        the reference; see dts_rf_filter).  keep_tape keeps what RFFilterAdjoint
        needs."""
        x = np.ascontiguousarray(inp, np.float32)
        if x.shape != (self.height, self.width):
            raise ValueError(f"input shape {x.shape} != {(self.height, self.width)}")
        out = np.empty((self.height, self.width), np.float32)
        fp = p.c()
        check(self._lib.dts_rf_filter(self._h, ctypes.byref(fp), x.ctypes.data, out.ctypes.data,
                                      int(num_iter), int(bool(keep_tape)), 0))
        return out

    def RFFilterDevice(self, p: DomainFilterParams, in_ptr: int, out_ptr: int,
                       num_iter: int = 3, keep_tape: bool = False):
        """RFFilter on device pointers (the handle's stream)."""
        fp = p.c()
        check(self._lib.dts_rf_filter(self._h, ctypes.byref(fp), in_ptr, out_ptr, int(num_iter),
                                      int(bool(keep_tape)), 1))

    def RFFilterAdjoint(self, ybar, with_guide: bool = True):
        """(xbar, gbar): derivatives of <ybar, RFFilter output> with respect to
        the input and the guide (H x W x 4) of the last RFFilter with keep_tape."""
        yb = np.ascontiguousarray(ybar, np.float32)
        xb = np.empty((self.height, self.width), np.float32)
        gb = np.empty((self.height, self.width, 4), np.float32) if with_guide else None
        check(self._lib.dts_rf_filter_adjoint(self._h, yb.ctypes.data, xb.ctypes.data,
                                              None if gb is None else gb.ctypes.data, 0))
        return xb, gb

    def RFFilterAdjointDevice(self, ybar_ptr: int, xbar_ptr: int, gbar_ptr: int | None = None):
        """RFFilterAdjoint on device pointers (the handle's stream)."""
        check(self._lib.dts_rf_filter_adjoint(self._h, ybar_ptr, xbar_ptr, gbar_ptr, 1))

    def ComputeConfidence(self, conf_params: DomainFilterParams, left_side_clear_width: int):
        fp = conf_params.c()
        check(self._lib.dts_compute_confidence(self._h, ctypes.byref(fp),
                                               float(conf_params.sigmas[0]),
                                               int(left_side_clear_width)))

    def Solve(self, opts: DtsOptions, compute_confidence: bool):
        o = opts.c()
        check(self._lib.dts_solve(self._h, ctypes.byref(o), int(bool(compute_confidence))))

    def Download(self, image_type: int):
        if image_type == COLOR_IMAGE:
            out = np.empty((self.height, self.width, 4), np.uint8)
        else:
            out = np.empty((self.height, self.width), np.float32)
        check(self._lib.dts_download(self._h, int(image_type),
                                     out.ctypes.data_as(ctypes.c_void_p)))
        return out

    def Synchronize(self):
        check(self._lib.dts_synchronize(self._h))

    def SetGraphMode(self, enable: bool):
        check(self._lib.dts_set_graph_mode(self._h, int(bool(enable))))

    def SetFused(self, enable: bool):
        """One launch per Optimize iteration (no overwrite) where the windows
        fit; False keeps the two line passes per iteration."""
        check(self._lib.dts_set_fused(self._h, int(bool(enable))))

    # --- test helpers -------------------------------------------------
    def debug_bounds(self, sigma: float, vertical: bool):
        lo = np.empty((self.height, self.width), np.int32)
        up = np.empty((self.height, self.width), np.int32)
        check(self._lib.dts_debug_bounds(self._h, float(sigma), int(bool(vertical)),
                                         lo.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                         up.ctypes.data_as(ctypes.POINTER(ctypes.c_int))))
        return lo, up

    def debug_ct(self, vertical: bool):
        out = np.empty((self.height, self.width), np.int64)
        check(self._lib.dts_debug_ct(self._h, int(bool(vertical)),
                                     out.ctypes.data_as(ctypes.c_void_p)))
        return out


class DomainTransformStereo(DomainTransformSolver):
    """src/domain_transform_stereo.h:36-57."""

    _flags = 1

    def InitRightFrame(self, center_offset: float, color_space: int, rgba):
        g, gp = _u8(rgba)
        check(self._lib.dts_init_right_frame(self._h, float(center_offset),
                                             int(color_space), gp))

    def ProcessRightFrame(self):
        check(self._lib.dts_process_right_frame(self._h))

    def Optimize(self, p: DomainOptimizeParams, overwrite_target_above_conf: float = -1.0):
        op = p.c()
        check(self._lib.dts_optimize_stereo(self._h, ctypes.byref(op),
                                            float(overwrite_target_above_conf)))


@dataclass
class HFBSGridParams:
    """src/hfbs.h:18-60 (same defaults)."""
    sigma_x: float = 8.0
    sigma_y: float = 8.0
    sigma_l: float = 4.0
    lambda_: float = 8.0
    alpha: float = 1.0
    beta: float = 0.9
    bistoch_iter_max: int = 20
    optim_iter_max: int = 140
    neg_exp_beta: float = 0.9

    def c(self) -> HfbsGridParams:
        return HfbsGridParams(self.sigma_x, self.sigma_y, self.sigma_l, self.lambda_,
                              self.alpha, self.beta, self.bistoch_iter_max,
                              self.optim_iter_max, self.neg_exp_beta, 1, 1, 1)


class HFBS:
    """src/hfbs.h:99-150."""

    def __init__(self, max_width, max_height, max_grid_x, max_grid_y, max_grid_l):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.hfbs_create(int(max_width), int(max_height), int(max_grid_x),
                                    int(max_grid_y), int(max_grid_l), ctypes.byref(h)))
        self._h = h
        self.width, self.height = int(max_width), int(max_height)
        self.fell_back = False

    def close(self):
        if getattr(self, "_h", None):
            self._lib.hfbs_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def SetImageDim(self, width, height) -> bool:
        rc = self._lib.hfbs_set_image_dim(self._h, int(width), int(height))
        if rc == 0:
            self.width, self.height = int(width), int(height)
        return rc == 0

    def SetFilterParams(self, p: HFBSGridParams):
        cp = p.c()
        check(self._lib.hfbs_set_params(self._h, ctypes.byref(cp)))

    def ClearAll(self):
        check(self._lib.hfbs_clear_all(self._h))

    def InitFrame(self, color_space, rgba, target, confidence=None):
        g, gp = _u8(rgba)
        t, tp = _f32(target)
        c, cp = _f32(confidence)
        check(self._lib.hfbs_init_frame(self._h, int(color_space), gp, tp, cp))

    def Optimize(self):
        fb = ctypes.c_int(0)
        check(self._lib.hfbs_optimize(self._h, ctypes.byref(fb)))
        self.fell_back = bool(fb.value)

    def Download(self, image_type):
        out = np.empty((self.height, self.width), np.float32)
        check(self._lib.hfbs_download(self._h, int(image_type),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def Synchronize(self):
        check(self._lib.hfbs_synchronize(self._h))

    def SetGraphMode(self, enable: bool):
        check(self._lib.hfbs_set_graph_mode(self._h, int(bool(enable))))

    def grid_dims(self):
        gx, gy, gl = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        check(self._lib.hfbs_debug_grid_dims(self._h, ctypes.byref(gx), ctypes.byref(gy),
                                             ctypes.byref(gl)))
        return gx.value, gy.value, gl.value

    def debug_grid(self, which: int):
        gx, gy, gl = self.grid_dims()
        out = np.empty((gl, gy, gx), np.float32)
        check(self._lib.hfbs_debug_grid(self._h, int(which),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out


def edge_aware_optimize(rgba, target, confidence, opts: DtsOptions, solver):
    """EdgeAwareOptimize (src/edge_aware_solver.h:60-90): returns (result, ms)."""
    solver.SetImageDim(*opts.image_dims)
    solver.ClearAll()
    solver.InitFrame(opts.color_space, rgba, target, confidence)
    lib = _lib.load()
    check(lib.dts_timer_start(solver.handle))
    solver.Solve(opts, confidence is None)
    ms = ctypes.c_float()
    check(lib.dts_timer_stop(solver.handle, ctypes.byref(ms)))
    return solver.Download(OPTIMIZED_QUANTITY), float(ms.value)


@dataclass
class BatchFrameSpec:
    """One frame of a batched solve: InitFrame inputs plus its DtsOptions."""
    rgba: np.ndarray
    target: np.ndarray
    confidence: np.ndarray | None
    filter_options: DomainFilterParams
    optim_options: DomainOptimizeParams
    color_space: int = YCbCr
    overwrite_target_above_conf: float = -1.0


class BatchSolver:
    """Many frames of uneven size solved in one graph replay (dts_batch_*)."""

    def __init__(self, max_frames: int, max_total_pixels: int):
        self._lib = _lib.load()
        h = ctypes.c_void_p()
        check(self._lib.dts_batch_create(int(max_frames), int(max_total_pixels),
                                         ctypes.byref(h)))
        self._h = h
        self.shapes = []

    def close(self):
        if getattr(self, "_h", None):
            self._lib.dts_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def SetFrames(self, frames: list[BatchFrameSpec]):
        keep = []
        arr = (_lib.BatchFrame * len(frames))()
        for i, f in enumerate(frames):
            g = np.ascontiguousarray(f.rgba, np.uint8)
            t = np.ascontiguousarray(f.target, np.float32)
            c = None if f.confidence is None else np.ascontiguousarray(f.confidence, np.float32)
            keep += [g, t, c]
            h, w = t.shape
            arr[i].width, arr[i].height, arr[i].color_space = w, h, int(f.color_space)
            arr[i].rgba = g.ctypes.data
            arr[i].target = t.ctypes.data
            arr[i].confidence = None if c is None else c.ctypes.data
            arr[i].filter_options = f.filter_options.c()
            arr[i].optim_options = f.optim_options.c()
            arr[i].overwrite_target_above_conf = float(f.overwrite_target_above_conf)
        check(self._lib.dts_batch_set_frames(self._h, arr, len(frames), 0))
        self.shapes = [f.target.shape for f in frames]
        self._sigmas = [float(f.optim_options.sigma_z) for f in frames]
        self._iters = [int(f.optim_options.num_iterations) for f in frames]

    def Solve(self):
        check(self._lib.dts_batch_solve(self._h))

    def SetExtentPlanning(self, enable: bool):
        """Plan the passes by each frame's measured window extents (default)
        or by the sigma-based caps; applies from the next SetFrames."""
        check(self._lib.dts_batch_set_extent_planning(self._h, int(bool(enable))))

    def FrameExtents(self, frame: int) -> tuple[int, int, int, int]:
        """Extents the frame's passes are planned for: (p - l, u - p) of its
        horizontal, then vertical, windows."""
        e = (ctypes.c_int * 4)()
        check(self._lib.dts_batch_get_frame_extents(self._h, int(frame), e))
        return tuple(e)

    def SetGraphMode(self, enable: bool):
        check(self._lib.dts_batch_set_graph_mode(self._h, int(bool(enable))))

    def SetTuning(self, x_blocks_per_cu: int, y_register_tiles: bool):
        check(self._lib.dts_batch_set_tuning(self._h, int(x_blocks_per_cu),
                                             int(bool(y_register_tiles))))

    def SetFused(self, max_chunks: int):
        """One launch per iteration for the frames without overwrite whose
        window span fits max_chunks (2, 3, 4) chunks of 64 columns; 0: none."""
        check(self._lib.dts_batch_set_fused(self._h, int(max_chunks)))

    def SetFusedWaves(self, waves: int):
        """Wavefronts per fused tile: 4 or 8."""
        check(self._lib.dts_batch_set_fused_waves(self._h, int(waves)))

    def FramesFused(self) -> int:
        mc, nf = ctypes.c_int(), ctypes.c_int()
        check(self._lib.dts_batch_get_fused(self._h, ctypes.byref(mc), ctypes.byref(nf)))
        return nf.value

    def GetTuning(self) -> dict:
        xb, yr = ctypes.c_int(), ctypes.c_int()
        mc, nf = ctypes.c_int(), ctypes.c_int()
        check(self._lib.dts_batch_get_tuning(self._h, ctypes.byref(xb), ctypes.byref(yr)))
        check(self._lib.dts_batch_get_fused(self._h, ctypes.byref(mc), ctypes.byref(nf)))
        fw = ctypes.c_int()
        check(self._lib.dts_batch_get_fused_waves(self._h, ctypes.byref(fw)))
        return {"x_blocks_per_cu": xb.value, "y_register_tiles": bool(yr.value),
                "fused_max_chunks": mc.value, "frames_fused": nf.value,
                "fused_waves": fw.value}

    def signature(self) -> str:
        """Key of the current frame set for the tuning cache: sizes, window
        sigmas, iteration counts and the window extents the passes are
        planned for."""
        parts = [f"{w}x{h}:{sz:.3g}:{k}:{self.FrameExtents(i)}" for i, ((h, w), sz, k) in
                 enumerate(zip(self.shapes, self._sigmas, self._iters))]
        return hashlib.sha1(";".join(parts).encode()).hexdigest()[:16]

    def Autotune(self, cache_path: str | None = None, reps: int = 2, device: str = "") -> dict:
        """Picks the fastest tuning for the current frames by timing the
        candidates (dts_batch_autotune); with cache_path, results are kept in
        a JSON file keyed by device and frame-set signature."""
        key = f"{device}|{self.signature()}"
        cache = {}
        if cache_path and os.path.exists(cache_path):
            with open(cache_path) as f:
                cache = json.load(f)
        if key in cache:
            t = cache[key]
            self.SetTuning(t["x_blocks_per_cu"], t["y_register_tiles"])
            self.SetFused(t.get("fused_max_chunks", 0))
            self.SetFusedWaves(t.get("fused_waves", 4))
            return dict(t, cached=True)
        ms = ctypes.c_float()
        check(self._lib.dts_batch_autotune(self._h, int(reps), ctypes.byref(ms)))
        t = dict(self.GetTuning(), ms=ms.value)
        if cache_path:
            cache[key] = t
            with open(cache_path, "w") as f:
                json.dump(cache, f, indent=1)
        return dict(t, cached=False)

    def Synchronize(self):
        check(self._lib.dts_batch_synchronize(self._h))

    def Download(self, frame: int):
        h, w = self.shapes[frame]
        out = np.empty((h, w), np.float32)
        check(self._lib.dts_batch_download(self._h, int(frame),
                                           out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out
