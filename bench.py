#!/usr/bin/env python3
"""Benchmark of the DTS / HFBS solve on one MI355X.

A step is one full Solve (src/edge_aware_solver.h:37-48 for DTS,
src/hfbs.cu:586-743 for HFBS) over one synthetic frame already resident in
HBM: guide differential, integration, window bounds and the K solver
iterations.  The headline workload is the 1920x1080 colorization solve
(DTS, Charbonnier, overwrite 0.85, 20 iterations); HFBS and the other
workloads are reported under "extra".

Prints ONE JSON line.  Metric: Mpix*iter/s = W*H*K / t_step / 1e6.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "DTS/HFBS solve Mpix·iter/s at 1080p, % HBM peak; batched & fwd+adjoint vs plain"


def dts_bytes_per_iter(overwrite: bool, stereo: bool = False) -> int:
    """Algorithmic bytes per pixel per iteration of the metric's definition:
    X half reads v, t, c, bounds and writes v' (20 B); Y half reads v',
    bounds and writes v'' (12 B); +4 B when the overwrite is active (the
    reference re-reads the target for it; here it is fused into the X pass
    and costs no extra traffic); +12 B for the stereo photometric term."""
    return 32 + (4 if overwrite else 0) + (12 if stereo else 0)


# Algorithmic bytes per pixel of one launch of each pass kernel as built:
# the horizontal pass reads v, t, c (4 B each) and the narrow bounds
# (2 B, extents <= 254) and writes v'; the vertical pass reads v' and the
# wide bounds (4 B) and writes v''.  The metric's own count (above) keeps
# 4-byte bounds for both halves.
X_PASS_BYTES = 18
Y_PASS_BYTES = 12
# One iteration as one launch (k_xypass): v, t, c and the vertical bounds
# read, the narrow horizontal bounds read, v'' written; the intermediate v'
# stays on chip (the target and confidence of the recomputed halo columns
# are cache hits of the same pixels).
XY_FUSED_BYTES = 22


def _opts_colorization(S, K=20):
    from domain_transform_solver_amd import scenes
    p = scenes.colorization_params()
    return S.DtsOptions(
        S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
        S.DomainOptimizeParams(K, p["loss"], p["lam"], p["sigma_z"], p["step"]),
        S.DomainFilterParams(8.0, 8.0, 0.2, [16.0]), p["color_space"], (1920, 1080),
        p["overwrite"])


def time_solves(lib, handle, solve, steps, warmup, sync):
    for _ in range(warmup):
        solve()
    sync()
    t0 = time.perf_counter()
    lib.dts_timer_start(handle) if lib is not None else None
    for _ in range(steps):
        solve()
    ev_ms = ctypes.c_float(0.0)
    if lib is not None:
        lib.dts_timer_stop(handle, ctypes.byref(ev_ms))
    sync()
    t1 = time.perf_counter()
    return (t1 - t0) * 1e3 / steps, ev_ms.value / steps


def bench_dts_colorization(S, args, extra):
    import torch
    from domain_transform_solver_amd import _lib, scenes
    lib = _lib.load()
    W, H, K = 1920, 1080, 20
    sc = scenes.colorization_scene(W, H, seed=2000)
    opts = _opts_colorization(S, K)
    gs = S.DomainTransformSolver(W, H)
    gs.InitFrame(opts.color_space, sc.rgba, sc.target_u, sc.confidence)
    gs.SetGraphMode(not args.no_graph)

    def solve():
        gs.Solve(opts, False)

    def sync():
        gs.Synchronize()
        torch.cuda.synchronize()

    ms_host, ms_ev = time_solves(lib, gs.handle, solve, args.steps, args.warmup, sync)
    P = W * H
    value = P * K / (ms_host * 1e-3) / 1e6
    bytes_solve = P * (12 + K * dts_bytes_per_iter(True))
    # Per-launch time of the iteration kernels (one fused launch per
    # iteration, or the two passes), events on the handle's stream.
    gs.SetGraphMode(False)
    lib.dts_profile_passes(gs.handle, 1)
    nprof = 10
    for _ in range(nprof):
        solve()
    xms, xc, yms, yc = ctypes.c_double(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    fms, fc = ctypes.c_double(), ctypes.c_int()
    lib.dts_profile_read(gs.handle, ctypes.byref(xms), ctypes.byref(xc), ctypes.byref(yms),
                         ctypes.byref(yc))
    lib.dts_profile_read_fused(gs.handle, ctypes.byref(fms), ctypes.byref(fc))
    lib.dts_profile_passes(gs.handle, 0)
    x_avg = xms.value / xc.value if xc.value else None
    y_avg = yms.value / yc.value if yc.value else None
    f_avg = fms.value / fc.value if fc.value else None
    x_gbs = X_PASS_BYTES * P / (x_avg * 1e-3) / 1e9 if x_avg else None
    y_gbs = Y_PASS_BYTES * P / (y_avg * 1e-3) / 1e9 if y_avg else None
    f_gbs = XY_FUSED_BYTES * P / (f_avg * 1e-3) / 1e9 if f_avg else None
    out = gs.Download(S.OPTIMIZED_QUANTITY)
    # The same solve with the host buffers the boundary takes (InitFrame
    # uploads, Download): the PCIe-inclusive rate, reported beside `value`.
    gs.SetGraphMode(not args.no_graph)
    tgt = np.ascontiguousarray(sc.target_u, np.float32)
    conf = np.ascontiguousarray(sc.confidence, np.float32)

    def solve_io():
        gs.InitFrame(opts.color_space, sc.rgba, tgt, conf)
        gs.Solve(opts, False)
        gs.Download(S.OPTIMIZED_QUANTITY)

    nio = max(1, min(args.steps, 20))
    solve_io()
    t0 = time.perf_counter()
    for _ in range(nio):
        solve_io()
    ms_io = (time.perf_counter() - t0) * 1e3 / nio
    extra["dts_colorization_1080p"] = {
        "ms_per_solve_host": ms_host, "ms_per_solve_events": ms_ev,
        "mpix_iter_s": value, "alg_bytes_per_solve": bytes_solve,
        "achieved_GBs_solve": bytes_solve / (ms_host * 1e-3) / 1e9,
        "pct_hbm_solve": 100.0 * bytes_solve / (ms_host * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "x_pass_ms": x_avg, "y_pass_ms": y_avg, "x_pass_GBs": x_gbs, "y_pass_GBs": y_gbs,
        "fused_launch_ms": f_avg, "fused_launch_GBs": f_gbs, "fused_launches": fc.value // nprof,
        "graph_mode": not args.no_graph,
        "ms_per_solve_with_host_buffers": ms_io,
        "mpix_iter_s_with_host_buffers": P * K / (ms_io * 1e-3) / 1e6,
    }
    # The roofline line is for the kernel with the largest share of the
    # solve: the fused iteration when the solve runs one launch per
    # iteration, else the slower of the two passes (both launch K times).
    if f_avg:
        achieved, kname, kpre, alg, avg = (
            f_gbs, "k_xypass<2, 2, true, 64, false, true, 16> (one iteration: Charbonnier + "
            "overwrite, 64x64 tiles of 16 wavefronts)", "k_xypass<2, 2, true, 64, false, true, 16>",
            XY_FUSED_BYTES * P, f_avg)
    elif y_avg >= x_avg:
        achieved, kname, kpre, alg, avg = (y_gbs, "k_ypass<3> (vertical pass)", "k_ypass<3>",
                                           Y_PASS_BYTES * P, y_avg)
    else:
        achieved, kname, kpre, alg, avg = (
            x_gbs, "k_xpass<2, true, false, N> (horizontal pass, Charbonnier + overwrite)",
            "k_xpass<2, true, false", X_PASS_BYTES * P, x_avg)
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": measured_traffic(kpre), "kernel": kname, "alg_bytes_per_launch": alg,
            "avg_launch_us": round(1e3 * avg, 3)}
    return value, ms_host, roof, (sc, opts, out)


def accuracy_vs_oracle(sc, opts, out, extra):
    import oracle
    from domain_transform_solver_amd import scenes
    W, H = 1920, 1080
    o = oracle.OracleDTS(W, H)
    f, op = opts.filter_options, opts.optim_options
    o.init_frame(opts.color_space, sc.rgba, sc.target_u, sc.confidence)
    o.compute_differential(f.sigma_x, f.sigma_y, f.sigma_r)
    o.integrate()
    o.optimize(op.num_iterations, op.loss, op.lambda_, op.sigma_z, op.step_size,
               opts.overwrite_target_above_conf)
    ref = o.result()
    rng = float(sc.target_u.max() - sc.target_u.min())
    err = float(np.abs(out.astype(np.float64) - ref).max())
    pg, po = scenes.psnr(out, sc.gt_u, 0.8), scenes.psnr(ref, sc.gt_u, 0.8)
    extra["dts_colorization_1080p"].update({
        "max_abs_err_vs_fp64_over_range": err / rng, "psnr_gpu": pg, "psnr_oracle": po})


def cpu_baseline_dts(sc, opts, budget_s=10.0):
    """FP64 oracle timed on the host cores on the same 1080p solve; repeats
    until budget_s of wall time has been spent."""
    import oracle
    W, H = 1920, 1080
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.set_threads(threads)
    o = oracle.OracleDTS(W, H)
    f, op = opts.filter_options, opts.optim_options
    o.init_frame(opts.color_space, sc.rgba, sc.target_u, sc.confidence)
    n, t_tot = 0, 0.0
    while n < 1000 and t_tot < budget_s:
        t0 = time.perf_counter()
        o.compute_differential(f.sigma_x, f.sigma_y, f.sigma_r)
        o.integrate()
        o.optimize(op.num_iterations, op.loss, op.lambda_, op.sigma_z, op.step_size,
                   opts.overwrite_target_above_conf)
        t_tot += time.perf_counter() - t0
        n += 1
    val = W * H * op.num_iterations * n / t_tot / 1e6
    return {"value": round(val, 2), "unit": "Mpix·iter/s", "cores": threads, "kind": "port",
            "sample": f"{n} FP64 oracle solves of the 1920x1080 colorization frame "
                      f"(20 iterations each, {t_tot:.1f} s)"}


def hfbs_bytes(P, V, K, Kb):
    """Algorithmic bytes of one HFBS solve: splat reads guide, t, c (12 B/px)
    and slice reads the guide and writes the output (8 B/px); per voxel
    12 B per bistochastization step, 28 B per solver iteration and about
    64 B for the blur, diagonal and initialisation."""
    return 20 * P + V * (12 * Kb + 28 * K + 64)


def bench_hfbs_colorization(S, args, extra):
    from domain_transform_solver_amd import _lib, scenes
    lib = _lib.load()
    W, H, K, Kb = 1920, 1080, 20, 20
    sc = scenes.colorization_scene(W, H, seed=2000)
    gp = S.HFBSGridParams(sigma_x=8, sigma_y=8, sigma_l=8, lambda_=8.0, alpha=1.0, beta=0.9,
                          bistoch_iter_max=Kb, optim_iter_max=K)
    s = S.HFBS(W, H, 240, 135, 32)
    s.SetImageDim(W, H)
    s.SetFilterParams(gp)
    s.InitFrame(0, sc.rgba, sc.target_u, sc.confidence)
    s.SetGraphMode(not args.no_graph)
    for _ in range(args.warmup):
        s.Optimize()
    s.Synchronize()
    t0 = time.perf_counter()
    lib.hfbs_timer_start(s.handle)
    for _ in range(args.steps):
        s.Optimize()
    ev = ctypes.c_float(0.0)
    lib.hfbs_timer_stop(s.handle, ctypes.byref(ev))
    s.Synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / args.steps
    gx, gy, gl = s.grid_dims()
    P, V = W * H, gx * gy * gl
    B = hfbs_bytes(P, V, K, Kb)
    out = s.Download(S.OPTIMIZED_QUANTITY)
    rec = {"ms_per_solve_host": ms, "ms_per_solve_events": ev.value / args.steps,
           "mpix_iter_s": P * K / (ms * 1e-3) / 1e6, "grid": [gx, gy, gl],
           "alg_bytes_per_solve": B, "achieved_GBs_solve": B / (ms * 1e-3) / 1e9,
           "pct_hbm_solve": 100.0 * B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    if not args.no_accuracy:
        import oracle
        prm = oracle.hfbs_params(8, 8, 8, 8.0, 1.0, 0.9, Kb, K, 0.9)
        ref, _, _ = oracle.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, prm)
        rec["max_abs_err_vs_fp64_over_range"] = float(np.abs(out - ref).max()) / 0.8
        rec["psnr_gpu"] = scenes.psnr(out, sc.gt_u, 0.8)
        rec["psnr_oracle"] = scenes.psnr(ref, sc.gt_u, 0.8)
    extra["hfbs_colorization_1080p"] = rec


def bench_batched(S, args, extra):
    """64 frames of uneven size (320x240 .. 1920x1080, factors 4/8/16,
    depth super-resolution presets, L2, K = 10) solved by one dts_batch_solve
    graph replay, against 64 plain per-frame solves (one resident handle per
    frame, all on one stream, each a graph replay).  The batch's passes are
    planned by the frames' measured window extents (SetFrames runs one bounds
    pass for that); the same batch planned by the sigma-based caps is timed
    beside it.  Every timed solve recomputes the bounds and all iterations."""
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    frames, P_tot, PK = batched_frames(S)
    steps = max(1, min(args.steps, 20))
    b = S.BatchSolver(len(frames), P_tot)
    b.SetFrames(frames)
    default_tuning = b.GetTuning()
    b.SetGraphMode(not args.no_graph)
    # Default tuning first, then the autotuned choice (timed below).
    for _ in range(max(1, args.warmup)):
        b.Solve()
    b.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.Solve()
    b.Synchronize()
    ms_default = (time.perf_counter() - t0) * 1e3 / steps
    tuning = b.Autotune(None, reps=2)
    for _ in range(args.warmup):
        b.Solve()
    b.Synchronize()
    t0 = time.perf_counter()
    lib.dts_batch_timer_start(b.handle)
    for _ in range(steps):
        b.Solve()
    ev = ctypes.c_float(0.0)
    lib.dts_batch_timer_stop(b.handle, ctypes.byref(ev))
    b.Synchronize()
    ms_b = (time.perf_counter() - t0) * 1e3 / steps

    # Plain: one handle per frame, every handle on the first handle's stream.
    hs = []
    for f in frames:
        h, w = f.target.shape
        g = S.DomainTransformSolver(w, h)
        g.InitFrame(f.color_space, f.rgba, f.target, f.confidence)
        g.SetGraphMode(not args.no_graph)
        hs.append(g)
    st = ctypes.c_void_p()
    _lib.check(lib.dts_get_stream(hs[0].handle, ctypes.byref(st)))
    for g in hs[1:]:
        _lib.check(lib.dts_set_stream(g.handle, st))
    opts = [S.DtsOptions(f.filter_options, f.optim_options, S.DomainFilterParams(),
                         f.color_space, f.target.shape[::-1]) for f in frames]

    def plain():
        for g, o in zip(hs, opts):
            g.Solve(o, False)

    for _ in range(args.warmup):
        plain()
    hs[0].Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plain()
    hs[0].Synchronize()
    ms_p = (time.perf_counter() - t0) * 1e3 / steps
    diff = max(float(np.abs(b.Download(i) - g.Download(S.OPTIMIZED_QUANTITY)).max())
               for i, g in enumerate(hs))
    planned = [b.FrameExtents(i) for i in range(len(frames))]
    # The same batch planned by the sigma-based caps (no measuring bounds
    # pass in SetFrames), autotuned the same way.
    b.SetExtentPlanning(False)
    b.SetFrames(frames)
    tuning_caps = b.Autotune(None, reps=2)
    for _ in range(args.warmup):
        b.Solve()
    b.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        b.Solve()
    b.Synchronize()
    ms_caps = (time.perf_counter() - t0) * 1e3 / steps
    caps = [b.FrameExtents(i) for i in range(len(frames))]
    B = P_tot * 12 + PK * dts_bytes_per_iter(False)
    extra["dts_batched_64_frames"] = {
        "frames": len(frames), "total_pixels": P_tot, "steps": steps,
        "ms_per_batch_host": ms_b, "ms_per_batch_events": ev.value / steps,
        "mpix_iter_s": PK / (ms_b * 1e-3) / 1e6,
        "alg_bytes_per_batch": B, "achieved_GBs": B / (ms_b * 1e-3) / 1e9,
        "pct_hbm": 100.0 * B / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "ms_per_batch_default_tuning": ms_default, "default_tuning": default_tuning,
        "autotuned": {k: v for k, v in tuning.items() if k != "cached"},
        "ms_64_plain_solves": ms_p, "speedup_vs_plain": ms_p / ms_b,
        "max_abs_diff_batched_vs_plain": diff,
        "planned_extents_max": [max(e[k] for e in planned) for k in range(4)],
        "cap_extents_max": [max(e[k] for e in caps) for k in range(4)],
        "ms_per_batch_planned_by_caps": ms_caps,
        "autotuned_planned_by_caps": {k: v for k, v in tuning_caps.items() if k != "cached"},
    }
    for g in reversed(hs):  # hs[0] owns the shared stream
        g.close()
    b.close()


def bench_depth_sr(S, args, extra, W=320, H=240, f=8):
    """Depth super-resolution, 8x, L2, 10 iterations: at 320x240 the
    reference's own CPU-runnable case, at 1920x1080 the frame size the
    HBM-fraction target is stated on.  One step = one Solve (graph replay):
    differential, integration, bounds and the 10 iterations."""
    from domain_transform_solver_amd import _lib, scenes
    lib = _lib.load()
    sc = scenes.depth_scene(W, H, f, seed=1000)
    p = scenes.depth_sr_params(f)
    opts = S.DtsOptions(S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
                        S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"],
                                               p["step"]),
                        S.DomainFilterParams(), p["color_space"], (W, H))
    g = S.DomainTransformSolver(W, H)
    g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    g.SetGraphMode(not args.no_graph)
    ms, ms_ev = time_solves(lib, g.handle, lambda: g.Solve(opts, False), args.steps,
                            args.warmup, g.Synchronize)
    out = g.Download(S.OPTIMIZED_QUANTITY)
    P, K = W * H, p["num_iter"]
    B = P * 12 + P * K * dts_bytes_per_iter(False)
    # Per-launch time of the one-launch iterations (eager, events), then the
    # same solve with the two line passes per iteration.
    g.SetGraphMode(False)
    lib.dts_profile_passes(g.handle, 1)
    for _ in range(5):
        g.Solve(opts, False)
    fms, fn = ctypes.c_double(), ctypes.c_int()
    lib.dts_profile_read_fused(g.handle, ctypes.byref(fms), ctypes.byref(fn))
    lib.dts_profile_passes(g.handle, 0)
    g.SetFused(False)
    g.SetGraphMode(not args.no_graph)
    ms2, _ = time_solves(lib, g.handle, lambda: g.Solve(opts, False), args.steps,
                         args.warmup, g.Synchronize)
    g.SetFused(True)
    rec = {"ms_per_solve_host": ms, "ms_per_solve_events": ms_ev,
           "mpix_iter_s": P * K / (ms * 1e-3) / 1e6,
           "alg_bytes_per_solve": B,
           "pct_hbm_solve": 100 * B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "fused_iterations": fn.value // 5,
           "fused_launch_us": 1e3 * fms.value / max(fn.value, 1),
           "ms_per_solve_two_pass": ms2,
           "rmse_gpu_m": scenes.rmse(out, sc.gt), "rmse_input_m": scenes.rmse(sc.target, sc.gt)}
    if not args.no_accuracy:
        import oracle
        o = oracle.OracleDTS(W, H)
        o.init_frame(p["color_space"], sc.rgba, sc.target, sc.confidence)
        o.compute_differential(p["sigma_x"], p["sigma_y"], p["sigma_r"])
        o.integrate()
        o.optimize(K, p["loss"], p["lam"], p["sigma_z"], p["step"])
        rng = float(sc.target.max() - sc.target.min())
        rec["max_abs_err_vs_fp64_over_range"] = float(np.abs(out - o.result()).max()) / rng
        rec["rmse_oracle_m"] = scenes.rmse(o.result(), sc.gt)
    extra[f"dts_depth_sr_{W}x{H}"] = rec
    g.close()


def bench_stereo(S, args, extra):
    """Stereo 1280x720: RGB 8/8/0.1, Charbonnier, sigma_z 30, 15 iterations,
    photometric weight 0.01.  One step = differential + integrate + right
    frame + Optimize (the stereo CLI's per-frame sequence)."""
    from domain_transform_solver_amd import scenes
    W, H = 1280, 720
    sc = scenes.stereo_scene(W, H, seed=3000)
    p = scenes.stereo_params()
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"],
                                photo_consis_lambda=p["photo_lambda"])
    g = S.DomainTransformStereo(W, H)
    g.InitFrame(0, sc.rgba, sc.target, sc.confidence)
    g.InitRightFrame(0.0, 0, sc.rgba_right)

    def step():
        g.ComputeColorSpaceDifferential(fp)
        g.IntegrateColorDifferentials()
        g.ProcessRightFrame()
        g.Optimize(op)

    steps = max(1, min(args.steps, 20))
    for _ in range(args.warmup):
        step()
    g.Synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    g.Synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    out = g.Download(S.OPTIMIZED_QUANTITY)
    P, K = W * H, p["num_iter"]
    B = P * (12 + K * dts_bytes_per_iter(False, stereo=True))
    rec = {"ms_per_step_host": ms, "mpix_iter_s": P * K / (ms * 1e-3) / 1e6,
           "alg_bytes_per_step": B, "pct_hbm": 100.0 * B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "bad_pixel_rate_gpu": scenes.bad_pixel_rate(out, sc.gt),
           "bad_pixel_rate_input": scenes.bad_pixel_rate(sc.target, sc.gt)}
    if not args.no_accuracy:
        import oracle
        o = oracle.OracleDTS(W, H)
        o.init_frame(0, sc.rgba, sc.target, sc.confidence)
        o.init_right_frame(0.0, 0, sc.rgba_right)
        o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
        o.integrate()
        o.process_right_frame()
        o.optimize_stereo(K, p["loss"], p["lam"], p["sigma_z"], p["step"], p["photo_lambda"])
        rec["bad_pixel_rate_oracle"] = scenes.bad_pixel_rate(o.result(), sc.gt)
    # The reference stereo CLI's own preset runs a single iteration
    # (src/benchmark_stereo_cli.cpp:242-250); at 15 the Charbonnier fixed step
    # oscillates around the data (in the FP64 oracle as well).
    op1 = S.DomainOptimizeParams(1, p["loss"], p["lam"], p["sigma_z"], p["step"],
                                 photo_consis_lambda=p["photo_lambda"])
    g.Optimize(op1)
    rec["bad_pixel_rate_gpu_1_iteration_preset"] = scenes.bad_pixel_rate(
        g.Download(S.OPTIMIZED_QUANTITY), sc.gt)
    extra["dts_stereo_1280x720"] = rec
    g.close()


def bench_rf(S, args, extra):
    """RF-DT (recursive-filter domain transform, not in the reference) at
    1920x1080, 3 iterations: forward, and forward with tape + adjoint (input
    and guide cotangents), on device pointers."""
    import torch
    from domain_transform_solver_amd import scenes
    W, H, N = 1920, 1080, 3
    sc = scenes.colorization_scene(W, H, seed=2000)
    fp = S.DomainFilterParams(20.0, 20.0, 0.4)
    g = S.DomainTransformSolver(W, H)
    g.InitFrame(0, sc.rgba, sc.target_u, sc.confidence)
    x = torch.from_numpy(np.ascontiguousarray(sc.target_u, np.float32)).cuda()
    y = torch.empty_like(x)
    yb = torch.randn(H, W, device="cuda")
    xb = torch.empty_like(yb)
    gb = torch.empty(H, W, 4, device="cuda")
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 20))

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            fn()
        g.Synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        g.Synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_f = timed(lambda: g.RFFilterDevice(fp, x.data_ptr(), y.data_ptr(), N))

    def fwd_adj():
        g.RFFilterDevice(fp, x.data_ptr(), y.data_ptr(), N, keep_tape=True)
        g.RFFilterAdjointDevice(yb.data_ptr(), xb.data_ptr(), gb.data_ptr())

    ms_fb = timed(fwd_adj)
    rec = {"ms_forward": ms_f, "ms_forward_plus_adjoint": ms_fb,
           "ratio_fwd_adj_over_fwd": ms_fb / ms_f, "iterations": N,
           "mpix_iter_s_forward": W * H * N / (ms_f * 1e-3) / 1e6}
    if not args.no_accuracy:
        import oracle
        guide = sc.rgba.astype(np.float64)
        ref = oracle.rf_filter(guide, 20.0, 20.0, 0.4, N, sc.target_u)
        rec["max_abs_err_vs_fp64_over_max"] = float(np.abs(y.cpu().numpy() - ref).max() /
                                                    np.abs(ref).max())
    extra["rf_dt_1080p"] = rec
    g.close()


def bench_adjoint(S, args, extra):
    """Differentiable solve 1024x768, 8x depth scene, L2, 20 iterations,
    checkpoint every 5: forward (Optimize) vs checkpointed forward + adjoint
    (device pointers, no host copies)."""
    import torch
    from domain_transform_solver_amd import _lib, scenes
    W, H, f, K, C = 1024, 768, 8, 20, 5
    sc = scenes.depth_scene(W, H, f, seed=4100)
    p = scenes.depth_sr_params(f)
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(K, p["loss"], p["lam"], p["sigma_z"], p["step"])
    g = S.DomainTransformSolver(W, H)
    g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    g.ComputeColorSpaceDifferential(fp)
    g.IntegrateColorDifferentials()
    yb = torch.randn(H, W, device="cuda")
    tb = torch.empty_like(yb)
    cb = torch.empty_like(yb)
    torch.cuda.synchronize()
    steps = max(1, min(args.steps, 20))

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            fn()
        g.Synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            fn()
        g.Synchronize()
        return (time.perf_counter() - t0) * 1e3 / steps

    ms_fwd = timed(lambda: g.Optimize(op))

    def fwd_adj():
        g.OptimizeCheckpointed(op, -1.0, C)
        g.OptimizeAdjointDevice(yb.data_ptr(), tb.data_ptr(), cb.data_ptr())

    ms_fb = timed(fwd_adj)
    rec = {"ms_forward": ms_fwd, "ms_forward_plus_adjoint": ms_fb,
           "ratio_fwd_adj_over_fwd": ms_fb / ms_fwd, "checkpoint_every": C, "iterations": K}
    if not args.no_accuracy:
        import oracle
        o = oracle.OracleDTS(W, H)
        o.init_frame(p["color_space"], sc.rgba, sc.target, sc.confidence)
        o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
        o.integrate()
        rtb, rcb = o.optimize_adjoint(K, p["loss"], p["lam"], p["sigma_z"], p["step"], -1.0,
                                      yb.cpu().numpy().astype(np.float64))
        rec["tbar_max_err_rel"] = float(np.abs(tb.cpu().numpy() - rtb).max() /
                                        np.abs(rtb).max())
        rec["cbar_max_err_rel"] = float(np.abs(cb.cpu().numpy() - rcb).max() /
                                        np.abs(rcb).max())
    extra["dts_adjoint_1024x768"] = rec
    g.close()


TRAFFIC_JSON = os.path.join(ROOT, "profiles", "traffic_dts_colorization_1080p.json")


def measured_traffic(kernel_prefix: str):
    """HBM bytes per dispatch of a kernel from the committed counter summary
    (FETCH_SIZE + WRITE_SIZE scaled by the stream-kernel calibration), or
    None when there is no summary for it."""
    try:
        with open(TRAFFIC_JSON) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    for name, v in t.get("kernels", {}).items():
        if name.replace("dtsamd::", "").replace("void ", "").startswith(kernel_prefix):
            return round(v["bytes"], 0)
    return None


def batched_frames(S):
    """The 64-frame uneven-size depth super-resolution batch (seeded)."""
    from domain_transform_solver_amd import scenes
    specs = scenes.batch_specs(64, seed=4000)
    frames, P_tot, PK = [], 0, 0
    for i, (w, h, f) in enumerate(specs):
        sc = scenes.depth_scene(w, h, f, seed=4000 + i)
        p = scenes.depth_sr_params(f)
        frames.append(S.BatchFrameSpec(
            sc.rgba, sc.target, sc.confidence,
            S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
            S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"]),
            p["color_space"]))
        P_tot += w * h
        PK += w * h * p["num_iter"]
    return frames, P_tot, PK


def pmc_workload(S, which="dts", batch_tuning=(4, 1, 3)):
    """Workload for the counter runs: three eager 1080p colorization solves
    (DTS, or HFBS with which="hfbs"), three eager 1080p 8x depth
    super-resolution solves (which="depth": one launch per iteration), three
    forward + checkpointed forward + adjoint runs of the 1024x768 differentiable
    solve (which="adjoint"), or five graph replays of the 64-frame batch at a
    fixed tuning (which="batched"),
    then the calibration stream kernel over 1920 x 1080 floats (3
    dispatches)."""
    import torch
    from domain_transform_solver_amd import _lib, scenes
    lib = _lib.load()
    W, H = 1920, 1080
    if which == "batched":
        # Fixed tuning (no autotuner trials in the trace): 5 graph replays.
        frames, _, _ = batched_frames(S)
        b = S.BatchSolver(len(frames), sum(f.target.size for f in frames))
        b.SetFrames(frames)
        b.SetTuning(batch_tuning[0], bool(batch_tuning[1]))
        if len(batch_tuning) > 2:
            b.SetFused(batch_tuning[2])
        b.SetGraphMode(True)
        for _ in range(5):
            b.Solve()
        b.Synchronize()
        b.close()
    elif which == "depth":
        sc = scenes.depth_scene(W, H, 8, seed=1000)
        p = scenes.depth_sr_params(8)
        opts = S.DtsOptions(S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
                            S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"],
                                                   p["sigma_z"], p["step"]),
                            S.DomainFilterParams(), p["color_space"], (W, H))
        g = S.DomainTransformSolver(W, H)
        g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
        g.SetGraphMode(False)
        for _ in range(3):
            g.Solve(opts, False)
        g.Synchronize()
    elif which == "adjoint":
        Wa, Ha, K, C = 1024, 768, 20, 5
        sc = scenes.depth_scene(Wa, Ha, 8, seed=4100)
        p = scenes.depth_sr_params(8)
        fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
        op = S.DomainOptimizeParams(K, p["loss"], p["lam"], p["sigma_z"], p["step"])
        g = S.DomainTransformSolver(Wa, Ha)
        g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
        g.ComputeColorSpaceDifferential(fp)
        g.IntegrateColorDifferentials()
        yb = torch.randn(Ha, Wa, device="cuda")
        tb = torch.empty_like(yb)
        cb = torch.empty_like(yb)
        torch.cuda.synchronize()
        for _ in range(3):
            g.Optimize(op)
            g.OptimizeCheckpointed(op, -1.0, C)
            g.OptimizeAdjointDevice(yb.data_ptr(), tb.data_ptr(), cb.data_ptr())
        g.Synchronize()
    elif which == "hfbs":
        sc = scenes.colorization_scene(W, H, seed=2000)
        gp = S.HFBSGridParams(sigma_x=8, sigma_y=8, sigma_l=8, lambda_=8.0, alpha=1.0,
                              beta=0.9, bistoch_iter_max=20, optim_iter_max=20)
        s = S.HFBS(W, H, 240, 135, 32)
        s.SetImageDim(W, H)
        s.SetFilterParams(gp)
        s.InitFrame(0, sc.rgba, sc.target_u, sc.confidence)
        s.SetGraphMode(False)
        for _ in range(3):
            s.Optimize()
        s.Synchronize()
    else:
        sc = scenes.colorization_scene(W, H, seed=2000)
        opts = _opts_colorization(S, 20)
        g = S.DomainTransformSolver(W, H)
        g.InitFrame(opts.color_space, sc.rgba, sc.target_u, sc.confidence)
        g.SetGraphMode(False)
        for _ in range(3):
            g.Solve(opts, False)
        g.Synchronize()
    n = W * H
    bufs = [torch.rand(n, device="cuda") for _ in range(5)]
    torch.cuda.synchronize()
    ms = ctypes.c_float()
    _lib.check(lib.dts_debug_stream(*[ctypes.c_void_p(b.data_ptr()) for b in bufs], n, 3,
                                    ctypes.byref(ms)))
    print(f"pmc workload done: stream kernel over {n} floats", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--pmc", action="store_true",
                    help="run the counter-collection workload only (no JSON line)")
    ap.add_argument("--pmc-workload", default="dts", choices=["dts", "depth", "hfbs", "batched", "adjoint"],
                    help="solver of the --pmc workload")
    ap.add_argument("--batch-tuning", default="4,1,3",
                    help="x_blocks_per_cu,y_register_tiles[,fused_max_chunks] of the batched "
                         "--pmc workload")
    ap.add_argument("--extras", default="hfbs,depth,stereo,adjoint,rf,batched",
                    help="comma list of extra workloads (hfbs, depth, stereo, adjoint, "
                         "batched) or none")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU; --gpus {args.gpus} is not supported "
              "(frames shard across GPUs without communication; see the design notes).",
              file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a GPU"
    torch.cuda.init()
    from domain_transform_solver_amd import solver as S
    if args.pmc:
        pmc_workload(S, args.pmc_workload,
                     tuple(int(v) for v in args.batch_tuning.split(",")))
        return

    extra = {}
    value, ms, roof, keep = bench_dts_colorization(S, args, extra)
    if not args.no_accuracy:
        accuracy_vs_oracle(*keep, extra)
    extras = set(x for x in args.extras.split(",") if x and x != "none")
    if "hfbs" in extras:
        bench_hfbs_colorization(S, args, extra)
    if "depth" in extras:
        bench_depth_sr(S, args, extra)
        bench_depth_sr(S, args, extra, 1920, 1080)
    if "stereo" in extras:
        bench_stereo(S, args, extra)
    if "adjoint" in extras:
        bench_adjoint(S, args, extra)
    if "rf" in extras:
        bench_rf(S, args, extra)
    if "batched" in extras:
        bench_batched(S, args, extra)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline_dts(keep[0], keep[1])
    line = {
        "metric": METRIC, "value": round(value, 1), "unit": "Mpix·iter/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic (seeded Voronoi colorization scene, 1% scribbles)",
        "config": {"workload": "colorization 1920x1080, DTS, Charbonnier, overwrite 0.85, "
                               "20 iterations (one chroma channel per step)",
                   "width": 1920, "height": 1080, "iterations": 20},
        "roofline": roof, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
