"""Writes the golden fixtures of this directory from the FP64 oracle.

Each fixture is an .npz holding the inputs of one small solve (guide, target,
confidence, parameters) and the oracle's outputs.  The reference ships no
golden vectors for this path (its tests assert nothing), so these pin the
oracle against regressions and give the GPU tests fixed expected values.

    python tests/golden/make_fixtures.py
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle  # noqa: E402
from domain_transform_solver_amd import scenes  # noqa: E402


def _dts(name, rgba, t, c, cs, fp, op, overwrite=-1.0, adjoint=False, filt=False):
    h, w = t.shape
    o = oracle.OracleDTS(w, h)
    o.init_frame(cs, rgba, t, c)
    o.compute_differential(*fp)
    o.integrate()
    out = {"rgba": rgba, "target": t, "confidence": c, "color_space": cs,
           "filter_params": np.array(fp, np.float32), "optimize_params": np.array(op, np.float32),
           "overwrite": np.float32(overwrite),
           "ctH": o.ctH_fx(), "ctV": o.ctV_fx()}
    lo, up = o.bounds(op[3], False)
    out["bounds_h"] = np.stack([lo, up])
    lo, up = o.bounds(op[3], True)
    out["bounds_v"] = np.stack([lo, up])
    K, loss, lam, sz, step = int(op[0]), int(op[1]), op[2], op[3], op[4]
    o.optimize(K, loss, lam, sz, step, overwrite)
    out["result"] = o.result()
    if adjoint:
        yb = np.random.default_rng(7).standard_normal((h, w))
        tb, cb = o.optimize_adjoint(K, loss, lam, sz, step, overwrite, yb)
        out.update(ybar=yb, tbar=tb, cbar=cb)
    if filt:
        x = np.random.default_rng(8).standard_normal((h, w)).astype(np.float32)
        out["filter_in"] = x
        out["filter_out"] = o.filter(fp[0], fp[1], fp[2], x, 3)
    np.savez_compressed(os.path.join(HERE, name), **out)


def main():
    oracle.set_threads(1)
    sc = scenes.depth_scene(48, 36, 4, seed=71)
    p = scenes.depth_sr_params(4)
    _dts("dts_depth_48x36.npz", sc.rgba, sc.target, sc.confidence, 1,
         (p["sigma_x"], p["sigma_y"], p["sigma_r"]), (10, 1, 0.99, p["sigma_z"], 0.99),
         adjoint=True)
    sc = scenes.depth_scene(40, 30, 4, seed=72)
    _dts("dts_wide_windows_40x30.npz", sc.rgba, sc.target, sc.confidence, 0,
         (8.0, 8.0, 1.0), (6, 1, 0.99, 9.0, 0.99), adjoint=True, filt=True)
    sc = scenes.colorization_scene(48, 32, seed=73, regions=5, scribble_frac=0.05)
    c = (sc.confidence * 0.002).astype(np.float32)
    _dts("dts_colorization_48x32.npz", sc.rgba, sc.target_u, c, 0,
         (64.0, 64.0, 0.25), (8, 2, 0.99, 30.0, 0.99), overwrite=0.0015)
    sc = scenes.colorization_scene(48, 32, seed=74, regions=5, scribble_frac=0.05)
    prm = oracle.hfbs_params(8, 8, 8, optim_iter_max=10, bistoch_iter_max=10)
    ref, _, _ = oracle.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, prm)
    np.savez_compressed(os.path.join(HERE, "hfbs_colorization_48x32.npz"), rgba=sc.rgba,
                        target=sc.target_u, confidence=sc.confidence,
                        params=np.array([8, 8, 8, 10, 10], np.float32), result=ref)


if __name__ == "__main__":
    main()
