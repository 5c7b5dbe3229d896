"""GPU parity of the DTS path (HIP library through the C ABI) against the
FP64 oracle.

Bars:
  * colour conversion, 32.32 ct and window bounds: bit-exact;
  * solver / filter / confidence values: max |gpu - oracle| <= 1e-4 x the
    value range of the target (fp32 storage and fp32 update arithmetic
    against FP64);
  * task error (depth RMSE) within 0.5% of the oracle's.
"""
import math

import numpy as np
import pytest

from domain_transform_solver_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


def _range(t):
    t = np.asarray(t, np.float64)
    r = float(t.max() - t.min())
    return r if r > 0 else 1.0


def _assert_close(gpu, ref, value_range, tol=TOL, what=""):
    err = float(np.max(np.abs(np.asarray(gpu, np.float64) - ref)))
    print(f"[{what}] max-abs {err:.3e}  / range {value_range:.3e} = {err / value_range:.3e}")
    assert err <= tol * value_range, f"{what}: {err} > {tol} * {value_range}"


def _random_guide(rng, w, h, levels=6):
    g = np.zeros((h, w, 4), np.uint8)
    g[..., :3] = rng.integers(0, levels, (h, w, 3)) * (255 // (levels - 1))
    g[..., 3] = rng.integers(0, 2, (h, w))
    return g


def _pair(S, oracle_mod, w, h, stereo=False):
    gs = (S.DomainTransformStereo if stereo else S.DomainTransformSolver)(w, h)
    orc = oracle_mod.OracleDTS(w, h)
    return gs, orc


@pytest.mark.parametrize("cs", [0, 1, 2])
def test_color_conversion_exact(S, oracle_mod, cs):
    rng = np.random.default_rng(100 + cs)
    w, h = 97, 41
    g = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(cs, g, t)
    orc.init_frame(cs, g, t)
    assert np.array_equal(gs.Download(S.COLOR_IMAGE), orc.color())


@pytest.mark.parametrize("dims", [(320, 240), (129, 77), (1, 33), (45, 1), (1, 1), (1100, 9),
                                  (2100, 5), (6, 2200), (257, 1025), (2048, 3), (1920, 1080),
                                  (2048, 4), (8, 1028)])
def test_ct_bit_exact(S, oracle_mod, dims):
    _check_ct(S, oracle_mod, dims, (64.0, 48.0, 0.1))


@pytest.mark.parametrize("dims", [(64, 32), (36, 20), (8, 1028)])
def test_ct_bit_exact_large_steps(S, oracle_mod, dims):
    # sigma_s / sigma_r = 1e6: steps up to 2e6, beyond the range where the
    # scan converts them with the 2^52 addition.
    _check_ct(S, oracle_mod, dims, (100.0, 100.0, 1e-4))


def _check_ct(S, oracle_mod, dims, params):
    w, h = dims
    rng = np.random.default_rng(w * 7 + h)
    g = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(1, g, t)
    orc.init_frame(1, g, t)
    fp = S.DomainFilterParams(*params)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(*params)
    orc.integrate()
    assert np.array_equal(gs.debug_ct(False), orc.ctH_fx())
    assert np.array_equal(gs.debug_ct(True), orc.ctV_fx())
    # Download(DIFFERENTIAL) is the fp32 rounding of the FP64 dH.
    np.testing.assert_array_equal(gs.Download(S.DIFFERENTIAL), orc.dHdx().astype(np.float32))
    ct = gs.Download(S.INTEGRAL)
    np.testing.assert_allclose(ct, orc.ctH_fx() / 2.0 ** 32, rtol=1e-6)


@pytest.mark.parametrize("formed", [False, True], ids=["from_guide", "from_ct"])
@pytest.mark.parametrize("dims", [(203, 117), (256, 120)])
@pytest.mark.parametrize("sigma", [0.0, 0.75, 2.0, 7.3, 35.555557, 300.0, float("nan"), 1e12])
def test_bounds_bit_exact(S, oracle_mod, sigma, dims, formed):
    """Bounds after a (lazy) integration come from the guide and the step
    tables; once the ct planes are formed (here by reading them) they come
    from the planes.  Both equal the oracle's bit for bit, on ragged dims
    and on dims whose lines take 16-byte guide loads."""
    rng = np.random.default_rng(5)
    w, h = dims
    g = _random_guide(rng, w, h, levels=3)
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(0, g, t)
    orc.init_frame(0, g, t)
    fp = S.DomainFilterParams(9.0, 9.0, 1.0)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(9.0, 9.0, 1.0)
    orc.integrate()
    if formed:
        gs.debug_ct(False)
    for vertical in (False, True):
        lo, up = gs.debug_bounds(sigma, vertical)
        olo, oup = orc.bounds(sigma, vertical)
        assert np.array_equal(lo, olo), f"lower bounds differ (vertical={vertical})"
        assert np.array_equal(up, oup), f"upper bounds differ (vertical={vertical})"


@pytest.mark.parametrize("dims", [(203, 117), (256, 120)])
@pytest.mark.parametrize("sigma", [15.5, 16.0, 16.99, 17.0, 24.0])
def test_bounds_near_range_staging(S, oracle_mod, sigma, dims):
    """Radii at and just past the 16-position near probes, on a guide flat in
    alternate 96x64 blocks (unit steps: windows as wide as the radius) and
    noisy in the others (short windows).  A bounds tile stages only the near
    range of its lines when the radius is 16 or more, and a wavefront whose
    windows reach past the probes restages its lines over the whole range:
    the blocks give tiles of either kind and tiles that mix both, in both
    directions."""
    rng = np.random.default_rng(11)
    w, h = dims
    g = np.full((h, w, 4), 128, np.uint8)
    noisy = rng.integers(0, 256, (h, w, 4), dtype=np.uint8)
    blocks = (np.arange(w)[None, :] // 96 + np.arange(h)[:, None] // 64) % 2 == 1
    g[blocks] = noisy[blocks]
    g[..., 3] = 255
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(0, g, t)
    orc.init_frame(0, g, t)
    gs.ComputeColorSpaceDifferential(S.DomainFilterParams(9.0, 9.0, 1.0))
    gs.IntegrateColorDifferentials()
    orc.compute_differential(9.0, 9.0, 1.0)
    orc.integrate()
    for vertical in (False, True):
        lo, up = gs.debug_bounds(sigma, vertical)
        olo, oup = orc.bounds(sigma, vertical)
        assert np.array_equal(lo, olo), f"lower bounds differ (vertical={vertical})"
        assert np.array_equal(up, oup), f"upper bounds differ (vertical={vertical})"
        pos = np.arange(h)[:, None] if vertical else np.arange(w)[None, :]
        ext = np.maximum(pos - lo, up - pos)
        assert ext.max() >= min(int(sigma), 16) and ext.min() <= 2


@pytest.mark.parametrize("dims", [(520, 300), (517, 301)])
def test_bounds_from_guide_two_chunk_stages(S, oracle_mod, dims):
    """Radius 100: every staged line of a bounds tile spans more than 256
    positions, so the guide staging scans two chunks per line with a carry;
    natural-image-like guide (smooth ramps, noise, edges) with wide SSDs."""
    rng = np.random.default_rng(9)
    w, h = dims
    yy, xx = np.mgrid[0:h, 0:w]
    g = np.zeros((h, w, 4), np.uint8)
    base = (xx * 0.3 + yy * 0.2) % 256
    for k in range(3):
        g[..., k] = np.clip(base + rng.normal(0, 3, (h, w)) + 120 * ((xx + 37 * k) % 97 < 9),
                            0, 255).astype(np.uint8)
    g[..., 3] = 255
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(0, g, t)
    orc.init_frame(0, g, t)
    gs.ComputeColorSpaceDifferential(S.DomainFilterParams(2.0, 2.0, 40.0))
    gs.IntegrateColorDifferentials()
    orc.compute_differential(2.0, 2.0, 40.0)
    orc.integrate()
    for vertical in (False, True):
        lo, up = gs.debug_bounds(100.0, vertical)
        olo, oup = orc.bounds(100.0, vertical)
        assert np.array_equal(lo, olo) and np.array_equal(up, oup), f"vertical={vertical}"
        pos = np.arange(h)[:, None] if vertical else np.arange(w)[None, :]
        assert int((pos - lo).max()) > 60  # windows reach well into the second chunk


def test_lazy_integration_survives_guide_change(S, oracle_mod):
    """Integrate, then a new guide without integrating again: the Optimize
    uses the ct of the guide that was integrated (as the reference, whose ct
    buffer is only rewritten by an integration), then a re-integration uses
    the new guide."""
    rng = np.random.default_rng(12)
    w, h = 160, 96
    g1 = _random_guide(rng, w, h, levels=4)
    g2 = _random_guide(rng, w, h, levels=2)
    t = rng.uniform(0.5, 5.0, (h, w)).astype(np.float32)
    c = (rng.uniform(size=(h, w)) < 0.3).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    fp = S.DomainFilterParams(6.0, 6.0, 0.4)
    op = S.DomainOptimizeParams(4, S.L2, 0.99, 9.0, 0.99)
    gs.InitFrame(0, g1, t, c)
    orc.init_frame(0, g1, t, c)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(6.0, 6.0, 0.4)
    orc.integrate()
    gs.InitFrame(0, g2, t, c)
    orc.init_frame(0, g2, t, c)
    gs.Optimize(op)
    orc.optimize(4, S.L2, 0.99, 9.0, 0.99)
    _assert_close(gs.Download(S.OPTIMIZED_QUANTITY), orc.result(), _range(t),
                  what="old ct after a guide change")
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(6.0, 6.0, 0.4)
    orc.integrate()
    gs.Optimize(op)
    orc.optimize(4, S.L2, 0.99, 9.0, 0.99)
    _assert_close(gs.Download(S.OPTIMIZED_QUANTITY), orc.result(), _range(t),
                  what="re-integrated")


def test_graph_solves_switching_step_tables(S, oracle_mod):
    """Graph-mode solves alternating between two range sigmas (two step
    tables for one handle's table slot) and eager solves in between: every
    result equals the eager solve of the same options."""
    sc = scenes.depth_scene(192, 128, 8, seed=1010)
    p = scenes.depth_sr_params(8)

    def opts(sr):
        return S.DtsOptions(S.DomainFilterParams(p["sigma_x"], p["sigma_y"], sr),
                            S.DomainOptimizeParams(6, S.L2, 0.99, p["sigma_z"], 0.99),
                            S.DomainFilterParams(), p["color_space"], (192, 128))

    gs = S.DomainTransformSolver(192, 128)
    gs.InitFrame(p["color_space"], sc.rgba, sc.target, sc.confidence)
    ref = {}
    for sr in (0.1, 0.05):
        gs.Solve(opts(sr), False)
        ref[sr] = gs.Download(S.OPTIMIZED_QUANTITY)
    assert not np.array_equal(ref[0.1], ref[0.05])
    for mode in (True, False, True, True, False, True):
        for sr in (0.1, 0.05, 0.1):
            gs.SetGraphMode(mode)
            gs.Solve(opts(sr), False)
            np.testing.assert_array_equal(gs.Download(S.OPTIMIZED_QUANTITY), ref[sr])
    gs.SetGraphMode(False)


def test_bounds_global_search_paths(S, oracle_mod):
    """Windows wider than the LDS stage (radius 200 on 700x450) take the
    global-memory search in both directions."""
    rng = np.random.default_rng(6)
    w, h = 700, 450
    g = _random_guide(rng, w, h, levels=3)
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(0, g, t)
    orc.init_frame(0, g, t)
    fp = S.DomainFilterParams(2.0, 2.0, 1.0)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(2.0, 2.0, 1.0)
    orc.integrate()
    for vertical in (False, True):
        lo, up = gs.debug_bounds(200.0, vertical)
        olo, oup = orc.bounds(200.0, vertical)
        assert np.array_equal(lo, olo) and np.array_equal(up, oup), f"vertical={vertical}"


def test_wide_bounds_elements(S, oracle_mod):
    """Extents above 254 do not fit the 2-byte bounds elements: on 320x300 with
    radius 280 the LDS-tiled search writes 4-byte elements, which the bounds
    download and both passes of Optimize read (against the oracle)."""
    rng = np.random.default_rng(61)
    w, h = 320, 300
    g = np.full((h, w, 4), 90, np.uint8)  # ct steps of exactly 1 away from two soft edges
    g[:, 160:] = 140
    g[150:] = 40
    gs, orc = _pair(S, oracle_mod, w, h)
    t = rng.uniform(0.5, 5.0, (h, w)).astype(np.float32)
    c = (rng.uniform(size=(h, w)) < 0.2).astype(np.float32)
    gs.InitFrame(0, g, t, c)
    orc.init_frame(0, g, t, c)
    fp = S.DomainFilterParams(0.5, 0.5, 1.0)  # ct steps close to 1: windows near +-280
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    orc.compute_differential(0.5, 0.5, 1.0)
    orc.integrate()
    for vertical in (False, True):
        lo, up = gs.debug_bounds(280.0, vertical)
        olo, oup = orc.bounds(280.0, vertical)
        assert np.array_equal(lo, olo) and np.array_equal(up, oup), f"vertical={vertical}"
        pos = np.arange(h)[:, None] if vertical else np.arange(w)[None, :]
        assert max(int((pos - lo).max()), int((up - pos).max())) > 255  # beyond 8-bit offsets
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    gs.Optimize(S.DomainOptimizeParams(4, S.L2, 0.99, 280.0, 0.99))
    orc.optimize(4, S.L2, 0.99, 280.0, 0.99)
    err = float(np.abs(gs.Download(S.OPTIMIZED_QUANTITY) - orc.result()).max())
    print(f"radius 280 on 320x300: max-abs {err:.2e}")
    assert err <= 1e-4 * 4.5


def test_optimize_wide_windows_lds_tiles(S, oracle_mod):
    """Radius 100: vertical tiles too tall for the register-staged form use
    the LDS-loop tile; the horizontal segments use 8 chunks."""
    rng = np.random.default_rng(8)
    w, h = 600, 400
    g = _random_guide(rng, w, h, levels=2)
    t = rng.normal(size=(h, w)).astype(np.float32)
    c = rng.uniform(0, 1, (h, w)).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    fp = S.DomainFilterParams(2.0, 2.0, 1.0)
    op = S.DomainOptimizeParams(3, S.L2, 0.99, 100.0, 0.99)
    gv, ov = _run_optimize(gs, orc, g, t, c, 0, fp, op)
    _assert_close(gv, ov, _range(t), what="radius 100")


@pytest.mark.parametrize("loss,overwrite", [(1, -1.0), (2, 0.85)])
def test_wide_window_fallback(S, oracle_mod, loss, overwrite):
    """Radius 300 on 1100x300: horizontal windows beyond the 8-chunk segments
    and vertical ones beyond the LDS tile take the double-precision prefix
    path in both directions; eager and graph-replayed solves match the oracle."""
    rng = np.random.default_rng(70 + loss)
    w, h = 1100, 300
    g = _random_guide(rng, w, h, levels=2)
    t = rng.uniform(0.5, 5.0, (h, w)).astype(np.float32)
    c = (rng.uniform(size=(h, w)) < 0.3).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    fp = S.DomainFilterParams(1.0, 1.0, 1.0)
    op = S.DomainOptimizeParams(3, loss, 0.99, 300.0, 0.99)
    gv, ov = _run_optimize(gs, orc, g, t, c, 0, fp, op, overwrite)
    _assert_close(gv, ov, _range(t), tol=1e-5, what=f"radius 300 loss {loss}")
    opts = S.DtsOptions(fp, op, S.DomainFilterParams(), 0, (w, h), overwrite)
    gs.SetGraphMode(True)
    for _ in range(2):
        gs.InitFrame(0, g, t, c)
        gs.Solve(opts, False)
        np.testing.assert_array_equal(gs.Download(S.OPTIMIZED_QUANTITY), gv)
    gs.SetGraphMode(False)


def test_wide_window_fallback_filter_and_stereo(S, oracle_mod):
    """Filter with sigma 400 (first iteration beyond both fast paths) and a
    stereo solve with radius 200 on 700-pixel rows, against the oracle."""
    rng = np.random.default_rng(77)
    w, h = 900, 700
    g = _random_guide(rng, w, h, levels=3)
    x = rng.uniform(0.0, 5.0, (h, w)).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    gs.InitFrame(0, g, np.zeros((h, w), np.float32))
    orc.init_frame(0, g, np.zeros((h, w), np.float32))
    fp = S.DomainFilterParams(400.0, 400.0, 100.0)
    _assert_close(gs.Filter(fp, x, 3), orc.filter(400.0, 400.0, 100.0, x, 3), 5.0,
                  what="filter sigma 400")
    sc = scenes.stereo_scene(700, 96, seed=3002, regions=8, max_disp=24.0)
    p = scenes.stereo_params()
    gs, orc = _pair(S, oracle_mod, 700, 96, stereo=True)
    fp = S.DomainFilterParams(2.0, 2.0, 1.0)
    op = S.DomainOptimizeParams(3, p["loss"], p["lam"], 200.0, p["step"],
                                photo_consis_lambda=p["photo_lambda"])
    gs.InitFrame(0, sc.rgba, sc.target, sc.confidence)
    gs.InitRightFrame(0.0, 0, sc.rgba_right)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    gs.ProcessRightFrame()
    gs.Optimize(op)
    orc.init_frame(0, sc.rgba, sc.target, sc.confidence)
    orc.init_right_frame(0.0, 0, sc.rgba_right)
    orc.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    orc.integrate()
    orc.process_right_frame()
    orc.optimize_stereo(3, p["loss"], p["lam"], 200.0, p["step"], p["photo_lambda"])
    d = np.abs(gs.Download(S.OPTIMIZED_QUANTITY) - orc.result())
    frac = float(np.mean(d > TOL * _range(sc.target)))
    print(f"wide stereo: {frac:.5f} of pixels beyond tolerance, max {d.max():.3e}")
    assert frac < 0.01


def _run_optimize(gs, orc, sc_rgba, t, c, cs, fp, op, overwrite=-1.0):
    w, h = t.shape[1], t.shape[0]
    gs.InitFrame(cs, sc_rgba, t, c)
    orc.init_frame(cs, sc_rgba, t, c)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    gs.Optimize(op, overwrite)
    orc.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    orc.integrate()
    orc.optimize(op.num_iterations, op.loss, op.lambda_, op.sigma_z, op.step_size, overwrite)
    return gs.Download(5), orc.result()  # OPTIMIZED_QUANTITY


def test_depth_superres_c0_parity(S, oracle_mod):
    """320x240 8x depth super-resolution, 10 iterations (the reference preset)."""
    sc = scenes.depth_scene(320, 240, 8, seed=1000)
    p = scenes.depth_sr_params(8)
    gs, orc = _pair(S, oracle_mod, 320, 240)
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(p["num_iter"], p["loss"], p["lam"], p["sigma_z"], p["step"])
    g, o = _run_optimize(gs, orc, sc.rgba, sc.target, sc.confidence, p["color_space"], fp, op)
    _assert_close(g, o, _range(sc.target), what="c0 depth SR")
    rg, ro = scenes.rmse(g, sc.gt), scenes.rmse(o, sc.gt)
    print(f"RMSE gpu {rg:.6f} oracle {ro:.6f}")
    assert abs(rg - ro) <= 0.005 * ro


@pytest.mark.parametrize("loss", [1, 2])
@pytest.mark.parametrize("overwrite", [-1.0, 0.85])
def test_optimize_random_parity(S, oracle_mod, loss, overwrite):
    """Random guide/target/confidence.  The Charbonnier step is unstable
    near v = t once 0.5*step*(c/n)/sqrt(1e-6) > 2 (a limit cycle that the
    FP64 oracle itself cannot reproduce under a 1e-7 input perturbation, see
    test_oracle.py::test_charbonnier_step_conditioning), so its random case
    keeps c/n small; the overwrite case pins c > 0.85 as colorization does."""
    rng = np.random.default_rng(loss * 10 + int(overwrite > 0))
    w, h = 181, 133
    g = _random_guide(rng, w, h)
    t = rng.normal(size=(h, w)).astype(np.float32)
    c = rng.uniform(0, 1, (h, w)).astype(np.float32)
    if loss == 2:
        c = np.where(c > 0.85, c, c * 0.002).astype(np.float32) if overwrite > 0 else \
            (c * 0.002).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    fp = S.DomainFilterParams(16.0, 12.0, 0.5)
    op = S.DomainOptimizeParams(12, loss, 0.99, 9.5, 0.99)
    gv, ov = _run_optimize(gs, orc, g, t, c, 1, fp, op, overwrite)
    _assert_close(gv, ov, _range(t), what=f"optimize loss={loss} ow={overwrite}")


def test_optimize_l1_and_zero_iterations_return_target(S, oracle_mod):
    rng = np.random.default_rng(3)
    w, h = 64, 48
    g = _random_guide(rng, w, h)
    t = rng.normal(size=(h, w)).astype(np.float32)
    gs = S.DomainTransformSolver(w, h)
    gs.InitFrame(0, g, t)
    gs.ComputeColorSpaceDifferential(S.DomainFilterParams(8, 8, 0.2))
    gs.IntegrateColorDifferentials()
    gs.Optimize(S.DomainOptimizeParams(5, S.L1, 0.99, 20.0, 0.99))
    np.testing.assert_array_equal(gs.Download(S.OPTIMIZED_QUANTITY), t)
    gs.Optimize(S.DomainOptimizeParams(0, S.L2, 0.99, 20.0, 0.99))
    np.testing.assert_array_equal(gs.Download(S.OPTIMIZED_QUANTITY), t)


@pytest.mark.parametrize("dims", [(1, 40), (40, 1), (2, 2), (300, 7), (7, 300)])
def test_optimize_degenerate_shapes(S, oracle_mod, dims):
    w, h = dims
    rng = np.random.default_rng(w + 100 * h)
    g = _random_guide(rng, w, h)
    t = rng.normal(size=(h, w)).astype(np.float32)
    c = rng.uniform(0, 1, (h, w)).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    fp = S.DomainFilterParams(8.0, 8.0, 0.4)
    op = S.DomainOptimizeParams(4, S.L2, 0.9, 6.0, 0.95)
    gv, ov = _run_optimize(gs, orc, g, t, c, 0, fp, op)
    _assert_close(gv, ov, _range(t), what=f"shape {dims}")


def test_set_image_dim_reuse(S, oracle_mod):
    """One handle, frames of different sizes; invalid dims keep the size."""
    gs = S.DomainTransformSolver(300, 200)
    assert not gs.SetImageDim(301, 10)
    assert gs.width == 300
    for (w, h, seed) in [(300, 200, 1), (120, 90, 2), (257, 31, 3)]:
        assert gs.SetImageDim(w, h)
        sc = scenes.depth_scene(w, h, 4, seed=seed)
        orc = oracle_mod.OracleDTS(w, h)
        p = scenes.depth_sr_params(4)
        fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
        op = S.DomainOptimizeParams(6, 1, 0.99, p["sigma_z"], 0.99)
        gv, ov = _run_optimize(gs, orc, sc.rgba, sc.target, sc.confidence, 1, fp, op)
        _assert_close(gv, ov, _range(sc.target), what=f"reuse {w}x{h}")


def test_filter_parity(S, oracle_mod):
    rng = np.random.default_rng(17)
    w, h = 211, 150
    g = _random_guide(rng, w, h)
    x = rng.normal(size=(h, w)).astype(np.float32)
    gs, orc = _pair(S, oracle_mod, w, h)
    t = np.zeros((h, w), np.float32)
    gs.InitFrame(1, g, t)
    orc.init_frame(1, g, t)
    fp = S.DomainFilterParams(20.0, 14.0, 0.6)
    out = gs.Filter(fp, x, 3)
    ref = orc.filter(20.0, 14.0, 0.6, x.astype(np.float64), 3)
    _assert_close(out, ref, _range(x), what="filter")
    # Filter leaves its result as OPTIMIZED_QUANTITY and rebuilds ct.
    np.testing.assert_array_equal(gs.Download(S.OPTIMIZED_QUANTITY), out)
    assert np.array_equal(gs.debug_ct(False), orc.ctH_fx())


def test_confidence_parity(S, oracle_mod):
    sc = scenes.depth_scene(256, 192, 8, seed=31)
    gs, orc = _pair(S, oracle_mod, 256, 192)
    gs.InitFrame(1, sc.rgba, sc.target)
    orc.init_frame(1, sc.rgba, sc.target)
    cp = S.DomainFilterParams(8.0, 8.0, 0.2, [16.0])
    gs.ComputeConfidence(cp, 20)
    orc.compute_confidence(8.0, 8.0, 0.2, 16.0, 20)
    _assert_close(gs.Download(S.CONFIDENCE), orc.confidence(), 1.0, what="confidence")


def test_solve_with_computed_confidence(S, oracle_mod):
    """Solve<DtsOptions> with no confidence: differential, integrate,
    ComputeConfidence (which rebuilds ct with its own params), Optimize."""
    sc = scenes.depth_scene(200, 160, 8, seed=41)
    p = scenes.depth_sr_params(8)
    opts = S.DtsOptions(
        S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
        S.DomainOptimizeParams(8, 1, 0.99, p["sigma_z"], 0.99),
        S.DomainFilterParams(8.0, 8.0, 0.2, [16.0]), 1, (200, 160))
    gs, orc = _pair(S, oracle_mod, 200, 160)
    g, ms = S.edge_aware_optimize(sc.rgba, sc.target, None, opts, gs)
    orc.init_frame(1, sc.rgba, sc.target)
    orc.compute_differential(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    orc.integrate()
    orc.compute_confidence(8.0, 8.0, 0.2, 16.0, 0)
    orc.optimize(8, 1, 0.99, p["sigma_z"], 0.99)
    _assert_close(g, orc.result(), _range(sc.target), what="solve+confidence")
    assert ms > 0


def test_stereo_parity(S, oracle_mod):
    sc = scenes.stereo_scene(256, 128, seed=3001, regions=8, max_disp=24.0)
    p = scenes.stereo_params()
    gs, orc = _pair(S, oracle_mod, 256, 128, stereo=True)
    fp = S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"])
    op = S.DomainOptimizeParams(6, p["loss"], p["lam"], p["sigma_z"], p["step"],
                                photo_consis_lambda=p["photo_lambda"])
    gs.InitFrame(0, sc.rgba, sc.target, sc.confidence)
    gs.InitRightFrame(0.0, 0, sc.rgba_right)
    gs.ComputeColorSpaceDifferential(fp)
    gs.IntegrateColorDifferentials()
    gs.ProcessRightFrame()
    gs.Optimize(op)
    orc.init_frame(0, sc.rgba, sc.target, sc.confidence)
    orc.init_right_frame(0.0, 0, sc.rgba_right)
    orc.compute_differential(fp.sigma_x, fp.sigma_y, fp.sigma_r)
    orc.integrate()
    orc.process_right_frame()
    orc.optimize_stereo(6, p["loss"], p["lam"], p["sigma_z"], p["step"], p["photo_lambda"])
    g, o = gs.Download(S.OPTIMIZED_QUANTITY), orc.result()
    # The photometric term reads the right view at int(doff + x - v): where
    # the fp32 and FP64 v straddle an integer the two gather different
    # pixels, so pixel agreement is checked on the fraction of pixels and
    # the task metric, not the max.
    d = np.abs(g - o)
    frac = float(np.mean(d > TOL * _range(sc.target)))
    print(f"stereo: {frac:.5f} of pixels beyond tolerance, max {d.max():.3e}")
    assert frac < 0.01
    bg, bo = scenes.bad_pixel_rate(g, sc.gt), scenes.bad_pixel_rate(o, sc.gt)
    assert abs(bg - bo) <= 0.005 * max(bo, 1e-3) + 1e-3


def test_graph_mode_replays_match_eager(S, oracle_mod):
    sc = scenes.depth_scene(320, 240, 8, seed=1000)
    p = scenes.depth_sr_params(8)
    opts = S.DtsOptions(
        S.DomainFilterParams(p["sigma_x"], p["sigma_y"], p["sigma_r"]),
        S.DomainOptimizeParams(10, 1, 0.99, p["sigma_z"], 0.99),
        S.DomainFilterParams(8.0, 8.0, 0.2, [16.0]), 1, (320, 240))
    gs = S.DomainTransformSolver(320, 240)
    gs.InitFrame(1, sc.rgba, sc.target, sc.confidence)
    gs.Solve(opts, False)
    eager = gs.Download(S.OPTIMIZED_QUANTITY)
    gs.SetGraphMode(True)
    for _ in range(4):
        gs.Solve(opts, False)
        rep = gs.Download(S.OPTIMIZED_QUANTITY)
        np.testing.assert_array_equal(rep, eager)
        gs.Synchronize()


def test_invalid_arguments_raise(S):
    from domain_transform_solver_amd._lib import DtsError
    gs = S.DomainTransformSolver(32, 32)
    g = np.zeros((32, 32, 4), np.uint8)
    gs.InitFrame(0, g, np.zeros((32, 32), np.float32))
    gs.ComputeColorSpaceDifferential(S.DomainFilterParams(8, 8, 0.2))
    gs.IntegrateColorDifferentials()
    with pytest.raises(DtsError):
        gs.Optimize(S.DomainOptimizeParams(3, 7, 0.99, 5.0, 0.99))
    with pytest.raises(DtsError):
        gs.Optimize(S.DomainOptimizeParams(3, 1, 0.99, -1.0, 0.99))
    with pytest.raises(DtsError):
        gs.InitFrame(9, g, np.zeros((32, 32), np.float32))
    gs2 = S.DomainTransformSolver(8, 8)
    assert gs2._lib.dts_process_right_frame(gs2.handle) == 4  # DTS_ERR_STATE
