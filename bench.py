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
    """Algorithmic HBM bytes per pixel per iteration: X pass reads v, t, c,
    bounds and writes v' (20 B); Y pass reads v', bounds and writes v'' (12 B).
    The overwrite is fused into the X pass (no extra bytes)."""
    return 32 + (12 if stereo else 0)


X_PASS_BYTES = 20  # per pixel per launch


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
    # Per-kernel time of the horizontal pass, events on the handle's stream.
    gs.SetGraphMode(False)
    lib.dts_profile_passes(gs.handle, 1)
    nprof = 3
    for _ in range(nprof):
        solve()
    xms, xc, yms, yc = ctypes.c_double(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    lib.dts_profile_read(gs.handle, ctypes.byref(xms), ctypes.byref(xc), ctypes.byref(yms),
                         ctypes.byref(yc))
    lib.dts_profile_passes(gs.handle, 0)
    x_avg = xms.value / max(xc.value, 1)
    y_avg = yms.value / max(yc.value, 1)
    achieved = X_PASS_BYTES * P / (x_avg * 1e-3) / 1e9
    out = gs.Download(S.OPTIMIZED_QUANTITY)
    extra["dts_colorization_1080p"] = {
        "ms_per_solve_host": ms_host, "ms_per_solve_events": ms_ev,
        "mpix_iter_s": value, "alg_bytes_per_solve": bytes_solve,
        "achieved_GBs_solve": bytes_solve / (ms_host * 1e-3) / 1e9,
        "pct_hbm_solve": 100.0 * bytes_solve / (ms_host * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "x_pass_ms": x_avg, "y_pass_ms": y_avg,
        "y_pass_GBs": 12 * P / (y_avg * 1e-3) / 1e9,
        "graph_mode": not args.no_graph,
    }
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "kernel": "k_line_pass (horizontal pass, Charbonnier + overwrite)",
            "alg_bytes_per_launch": X_PASS_BYTES * P}
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
    until ~budget_s of CPU time (at most 30 solves)."""
    import oracle
    W, H = 1920, 1080
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    oracle.set_threads(threads)
    o = oracle.OracleDTS(W, H)
    f, op = opts.filter_options, opts.optim_options
    o.init_frame(opts.color_space, sc.rgba, sc.target_u, sc.confidence)
    n, t_tot = 0, 0.0
    while n < 30 and t_tot < budget_s:
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
    frame, all on one stream, each a graph replay)."""
    from domain_transform_solver_amd import _lib, scenes
    lib = _lib.load()
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
    steps = max(1, min(args.steps, 20))
    b = S.BatchSolver(len(frames), P_tot)
    b.SetFrames(frames)
    b.SetGraphMode(not args.no_graph)
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
    B = P_tot * 12 + PK * dts_bytes_per_iter(False)
    extra["dts_batched_64_frames"] = {
        "frames": len(frames), "total_pixels": P_tot, "steps": steps,
        "ms_per_batch_host": ms_b, "ms_per_batch_events": ev.value / steps,
        "mpix_iter_s": PK / (ms_b * 1e-3) / 1e6,
        "alg_bytes_per_batch": B, "achieved_GBs": B / (ms_b * 1e-3) / 1e9,
        "pct_hbm": 100.0 * B / (ms_b * 1e-3) / 1e9 / HBM_PEAK_GBS,
        "ms_64_plain_solves": ms_p, "speedup_vs_plain": ms_p / ms_b,
        "max_abs_diff_batched_vs_plain": diff,
    }
    for g in reversed(hs):  # hs[0] owns the shared stream
        g.close()
    b.close()


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--extras", default="hfbs,batched",
                    help="comma list of extra workloads (hfbs, batched) or none")
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

    extra = {}
    value, ms, roof, keep = bench_dts_colorization(S, args, extra)
    if not args.no_accuracy:
        accuracy_vs_oracle(*keep, extra)
    extras = set(x for x in args.extras.split(",") if x and x != "none")
    if "hfbs" in extras:
        bench_hfbs_colorization(S, args, extra)
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
