"""The C++ class API as a drop-in: tests/cxx/dropin_solve.exe is compiled
against include/ only (DtsOptions / HfbsOptions + EdgeAwareOptimize, the
calls the reference CLIs make) and linked to libdts_amd.so.  Its output is
checked against the FP64 oracle."""
import os
import subprocess

import numpy as np
import pytest

from domain_transform_solver_amd import scenes

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cxx", "dropin_solve.exe")


def _run(tmp_path, algo, rgba, t, c, extra):
    h, w = t.shape
    paths = [tmp_path / n for n in ("g.u8", "t.f32", "c.f32", "o.f32")]
    np.ascontiguousarray(rgba, np.uint8).tofile(paths[0])
    np.ascontiguousarray(t, np.float32).tofile(paths[1])
    np.ascontiguousarray(c, np.float32).tofile(paths[2])
    cmd = [EXE, algo, str(w), str(h), *map(str, paths), *map(str, extra)]
    print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("ms=")
    return np.fromfile(paths[3], np.float32).reshape(h, w)


@pytest.fixture(scope="module")
def exe(gpu_lib):
    if not os.path.exists(EXE):
        pytest.fail("dropin_solve.exe is not built (run __graft_entry__.build())")
    return EXE


def test_dropin_dts_depth_superres(exe, oracle_mod, tmp_path):
    sc = scenes.depth_scene(320, 240, 8, seed=1000)
    p = scenes.depth_sr_params(8)
    out = _run(tmp_path, "dts", sc.rgba, sc.target, sc.confidence,
               [p["sigma_x"], p["sigma_y"], p["sigma_r"], p["sigma_z"], 10, 1])
    o = oracle_mod.OracleDTS(320, 240)
    o.init_frame(1, sc.rgba, sc.target, sc.confidence)
    o.compute_differential(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    o.integrate()
    o.optimize(10, 1, 0.99, p["sigma_z"], 0.99)
    err = np.abs(out - o.result()).max()
    assert err <= 1e-4 * float(sc.target.max() - sc.target.min()), err


def test_dropin_hfbs_colorization(exe, oracle_mod, tmp_path):
    sc = scenes.colorization_scene(256, 160, seed=11, regions=6, scribble_frac=0.03)
    out = _run(tmp_path, "hfbs", sc.rgba, sc.target_u, sc.confidence, [8, 8, 8, 0, 20, 0])
    prm = oracle_mod.hfbs_params(8, 8, 8, optim_iter_max=20)
    ref, _, _ = oracle_mod.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, prm)
    assert np.abs(out - ref).max() <= 1e-4 * 0.8


def test_dropin_rf_filter_and_adjoint(exe, oracle_mod, tmp_path):
    """DomainTransformFilter::RFFilter / RFFilterAdjoint (the RF-DT extension)
    through the C++ class API."""
    rng = np.random.default_rng(21)
    h, w = 70, 90
    rgba = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    x = rng.normal(size=(h, w)).astype(np.float32)
    yb = rng.normal(size=(h, w)).astype(np.float32)
    out = _run(tmp_path, "rf", rgba, x, yb, [12, 10, 0.4, 0, 3, 0])
    xb = np.fromfile(str(tmp_path / "o.f32") + ".xbar", np.float32).reshape(h, w)
    gb = np.fromfile(str(tmp_path / "o.f32") + ".gbar", np.float32).reshape(h, w, 4)
    g = oracle_mod.convert_color(rgba, 0).astype(np.float64)
    ref = oracle_mod.rf_filter(g, 12.0, 10.0, 0.4, 3, x)
    rxb, rgb = oracle_mod.rf_filter_adjoint(g, 12.0, 10.0, 0.4, 3, x, yb)
    assert np.abs(out - ref).max() <= 1e-5 * np.abs(ref).max()
    assert np.abs(xb - rxb).max() <= 1e-5 * np.abs(rxb).max()
    assert np.abs(gb - rgb).max() <= 1e-5 * np.abs(rgb).max()


def test_reference_style_program_runs(gpu_lib):
    """tests/cxx/reference_usage.exe: the reference's two-argument
    EdgeAwareSolver subclasses for DTS and HFBS run EdgeAwareOptimize on a
    ProblemImageData through the drop-in headers."""
    exe = os.path.join(ROOT, "tests", "cxx", "reference_usage.exe")
    if not os.path.exists(exe):
        pytest.fail("reference_usage.exe is not built (run __graft_entry__.build())")
    print(exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert r.stdout.startswith("dts ") and "hfbs" in r.stdout
