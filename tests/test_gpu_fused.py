"""One launch per Optimize iteration (dts_set_fused): the horizontal half of
the rows a vertical tile needs is recomputed inside the tile.  It runs for
Optimize without the target overwrite when the horizontal window span fits 4
chunks of 64 columns; the tests check which form ran (launch counts from the
pass profiler), parity with the FP64 oracle, agreement with the two-pass form
to float rounding and bit-identical graph replays."""
import ctypes

import numpy as np
import pytest

from domain_transform_solver_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-4          # vs the FP64 oracle, x the target's value range
TWO_PASS_TOL = 5e-5  # fused vs two-pass, x range: fp32 prefix sums from other columns


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


def _guide(rng, w, h, levels):
    """Random colours of `levels` levels per channel; levels = 0: flat in
    alternate 96x64 blocks and random in the others, so that some tiles'
    windows reach the caps and others stay within a few pixels."""
    g = np.zeros((h, w, 4), np.uint8)
    if levels == 0:
        g[..., :3] = 128
        blocks = (np.arange(w)[None, :] // 96 + np.arange(h)[:, None] // 64) % 2 == 1
        noisy = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
        g[..., :3][blocks] = noisy[blocks]
        return g
    g[..., :3] = rng.integers(0, levels, (h, w, 3)) * (255 // (levels - 1))
    return g


def _expect_fused(w, sigma_z, loss, K, overwrite):
    """The host's dispatch rule restated: window caps from floor(sigma_z)
    clipped to the row, chunks = max(2, ceil((63 + lo + hi) / 64)) <= 4 (with
    or without the overwrite)."""
    if loss == 0 or K <= 0:
        return False
    E = w if sigma_z >= w else int(np.floor(sigma_z))
    lo, hi = min(E + 1, w), min(E, w - 1)
    return max(2, -(-(63 + lo + hi) // 64)) <= 4


def _optimize(S, lib, g, op, overwrite, fused):
    g.SetFused(fused)
    lib.dts_profile_passes(g.handle, 1)
    g.Optimize(op, overwrite)
    ms, n = ctypes.c_double(), ctypes.c_int()
    lib.dts_profile_read_fused(g.handle, ctypes.byref(ms), ctypes.byref(n))
    xms, xc, yms, yc = ctypes.c_double(), ctypes.c_int(), ctypes.c_double(), ctypes.c_int()
    lib.dts_profile_read(g.handle, ctypes.byref(xms), ctypes.byref(xc), ctypes.byref(yms),
                         ctypes.byref(yc))
    lib.dts_profile_passes(g.handle, 0)
    return g.Download(S.OPTIMIZED_QUANTITY), n.value, xc.value, yc.value


CASES = [
    # w, h, sigma_z, loss, K, levels, overwrite
    (181, 133, 9.5, 1, 12, 6, -1.0),      # 2 chunks, L2
    (181, 133, 9.5, 2, 12, 6, -1.0),      # 2 chunks, Charbonnier
    (300, 200, 50.0, 1, 5, 2, -1.0),      # 3 chunks
    (300, 200, 90.0, 1, 4, 2, -1.0),      # 4 chunks
    (300, 200, 90.0, 2, 3, 2, -1.0),      # 4 chunks, Charbonnier
    (300, 200, 120.0, 1, 3, 2, -1.0),     # 5 chunks: two passes
    (300, 200, 50.0, 1, 5, 0, -1.0),      # 3 chunks; windows at the caps in flat blocks only
    (300, 200, 50.0, 2, 4, 0, 0.85),      # the same with the overwrite, Charbonnier
    (64, 300, 200.0, 1, 3, 2, -1.0),      # caps clipped by the row: 3 chunks, tall windows
    (181, 133, 9.5, 1, 6, 6, 0.85),       # overwrite
    (181, 133, 9.5, 2, 5, 6, 0.85),       # overwrite, Charbonnier
    (300, 200, 90.0, 1, 3, 2, 0.5),       # overwrite, 4 chunks
    (1, 40, 6.0, 1, 2, 6, -1.0),
    (40, 1, 6.0, 1, 1, 6, -1.0),
    (2, 2, 6.0, 1, 3, 6, -1.0),
    (7, 300, 6.0, 1, 2, 6, -1.0),
    (300, 7, 6.0, 2, 2, 6, -1.0),
    (65, 33, 6.0, 1, 2, 6, -1.0),
    (63, 31, 6.0, 1, 1, 6, -1.0),
    (1920, 1080, 35.555557, 1, 10, 6, -1.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}x{c[1]}-sz{c[2]}-l{c[3]}-K{c[4]}-ow{c[6]}")
def test_fused_matches_oracle_and_two_pass(S, oracle_mod, case):
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    w, h, sz, loss, K, levels, ow = case
    rng = np.random.default_rng(w * 7 + h + K)
    rgba = _guide(rng, w, h, levels)
    t = rng.normal(size=(h, w)).astype(np.float32)
    c = rng.uniform(0, 1, (h, w)).astype(np.float32)
    if loss == 2:
        # Charbonnier's step stays stable for small c / n (see the random
        # parity case of the two-pass solver).
        c = np.where(c > 0.85, c, c * 0.002).astype(np.float32) if ow > 0 else \
            (c * 0.002).astype(np.float32)
    fp = S.DomainFilterParams(16.0, 12.0, 0.5)
    op = S.DomainOptimizeParams(K, loss, 0.99, float(np.float32(sz)), 0.99)
    g = S.DomainTransformSolver(w, h)
    g.InitFrame(1, rgba, t, c)
    g.ComputeColorSpaceDifferential(fp)
    g.IntegrateColorDifferentials()
    fv, nf, xc, yc = _optimize(S, lib, g, op, ow, True)
    tv, nf2, xc2, yc2 = _optimize(S, lib, g, op, ow, False)
    expect = _expect_fused(w, op.sigma_z, loss, K, ow)
    print(f"{case}: fused launches {nf} (x {xc}, y {yc}); expected fused={expect}", flush=True)
    assert nf == (K if expect else 0)
    assert (xc, yc) == ((0, 0) if expect else (K, K))
    assert (nf2, xc2, yc2) == (0, K, K)
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(1, rgba, t, c)
    o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    o.integrate()
    o.optimize(K, loss, 0.99, op.sigma_z, 0.99, ow)
    rng_t = float(np.ptp(t)) or 1.0
    e_or = float(np.abs(fv.astype(np.float64) - o.result()).max()) / rng_t
    e_tp = float(np.abs(fv - tv).max()) / rng_t
    print(f"  vs oracle {e_or:.2e}, vs two-pass {e_tp:.2e} (x range)")
    assert e_or <= TOL
    if expect:
        assert e_tp <= TWO_PASS_TOL
    else:
        np.testing.assert_array_equal(fv, tv)
    g.close()


def test_fused_depth_superres_1080p_rmse(S, oracle_mod):
    """8x depth super-resolution at 1920x1080 (10 iterations, L2, no
    overwrite): the fused form runs and keeps the oracle's task error."""
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    W, H, f = 1920, 1080, 8
    sc = scenes.depth_scene(W, H, f, seed=1000)
    p = scenes.depth_sr_params(f)
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"])
    g = S.DomainTransformSolver(W, H)
    g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    g.ComputeColorSpaceDifferential(fp)
    g.IntegrateColorDifferentials()
    out, nf, _, _ = _optimize(S, lib, g, op, -1.0, True)
    assert nf == p["num_iter"]
    o = oracle_mod.OracleDTS(W, H)
    o.init_frame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    o.integrate()
    o.optimize(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"])
    rng_t = float(np.ptp(sc.target))
    err = float(np.abs(out.astype(np.float64) - o.result()).max()) / rng_t
    rg, ro = scenes.rmse(out, sc.gt), scenes.rmse(o.result(), sc.gt)
    print(f"1080p depth SR fused: {err:.2e} x range; RMSE gpu {rg:.6f} oracle {ro:.6f}")
    assert err <= TOL
    assert abs(rg - ro) <= 0.005 * ro
    g.close()


def test_fused_graph_replays_match_eager(S):
    """Solve through a captured graph with the fused iterations: every
    replay equals the eager result bit for bit, with eager work between."""
    sc = scenes.depth_scene(320, 240, 8, seed=1000)
    p = scenes.depth_sr_params(8)
    opts = S.DtsOptions(
        S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
        S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"]),
        S.DomainFilterParams(), p["color_space"], (320, 240))
    g = S.DomainTransformSolver(320, 240)
    g.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    g.Solve(opts, False)
    eager = g.Download(S.OPTIMIZED_QUANTITY)
    g.SetFused(False)
    g.Solve(opts, False)
    two_pass = g.Download(S.OPTIMIZED_QUANTITY)
    g.SetFused(True)
    g.SetGraphMode(True)
    other = S.DomainTransformSolver(320, 240)  # eager work between the replays
    other.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    for i in range(4):
        g.Solve(opts, False)
        np.testing.assert_array_equal(g.Download(S.OPTIMIZED_QUANTITY), eager)
        other.Solve(opts, False)
        np.testing.assert_array_equal(other.Download(S.OPTIMIZED_QUANTITY), eager)
        g.Synchronize()
    assert float(np.abs(eager - two_pass).max()) <= TWO_PASS_TOL * float(np.ptp(sc.target))
    other.close()
    g.close()


@pytest.mark.parametrize("dims", [(256, 128), (200, 77)])
def test_fused_stereo_matches_oracle_and_two_pass(S, oracle_mod, dims):
    """The stereo solve with one launch per iteration (the photometric term
    in the recomputed horizontal half).  Where the fp32 and FP64 v straddle
    an integer the term gathers different right-view pixels, so agreement is
    bounded by the fraction of pixels off and the task metric, as for the
    two-pass form; a 200-wide frame leaves lanes past the row end in its
    last tile."""
    from domain_transform_solver_amd import _lib
    lib = _lib.load()
    w, h = dims
    sc = scenes.stereo_scene(w, h, seed=3001, regions=8, max_disp=24.0)
    p = scenes.stereo_params()
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(6, p["loss"], p["lam"], p["sigma_z"], p["step"],
                                photo_consis_lambda=p["photo_lambda"])
    g = S.DomainTransformStereo(w, h)
    g.InitFrame(0, sc.rgba, sc.target, sc.confidence)
    g.InitRightFrame(0.0, 0, sc.rgba_right)
    g.ComputeColorSpaceDifferential(fp)
    g.IntegrateColorDifferentials()
    g.ProcessRightFrame()
    fv, nf, _, _ = _optimize(S, lib, g, op, -1.0, True)
    tv, nf2, xc, yc = _optimize(S, lib, g, op, -1.0, False)
    assert nf == 6 and (nf2, xc, yc) == (0, 6, 6)
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(0, sc.rgba, sc.target, sc.confidence)
    o.init_right_frame(0.0, 0, sc.rgba_right)
    o.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    o.integrate()
    o.process_right_frame()
    o.optimize_stereo(6, p["loss"], p["lam"], p["sigma_z"], p["step"], p["photo_lambda"])
    ref = o.result()
    rng_t = float(np.ptp(sc.target))
    frac = float(np.mean(np.abs(fv - ref) > TOL * rng_t))
    frac2 = float(np.mean(np.abs(fv - tv) > TOL * rng_t))
    print(f"stereo {dims}: fused vs oracle {frac:.5f}, vs two-pass {frac2:.5f} of pixels off")
    assert frac < 0.01 and frac2 < 0.01
    bg, bo = scenes.bad_pixel_rate(fv, sc.gt), scenes.bad_pixel_rate(ref, sc.gt)
    assert abs(bg - bo) <= 0.005 * max(bo, 1e-3) + 1e-3
    g.close()
