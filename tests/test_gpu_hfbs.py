"""GPU parity of the HFBS path (C ABI) against the FP64 oracle.

Grid quantities that are sums of integers (splat counts) are compared
exactly; the splat sums and the solver values in fp32 against FP64 to
1e-4 of the value range (the splat uses fp32 LDS atomics whose order is
not fixed, so the grids are compared with a tolerance, not bitwise).
"""
import numpy as np
import pytest

from domain_transform_solver_amd import scenes

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


def _params(S, oracle_mod, **kw):
    gp = S.HFBSGridParams(**kw)
    op = oracle_mod.hfbs_params(gp.sigma_x, gp.sigma_y, gp.sigma_l, gp.lambda_, gp.alpha,
                                gp.beta, gp.bistoch_iter_max, gp.optim_iter_max,
                                gp.neg_exp_beta)
    return gp, op


def _run_gpu(S, sc_rgba, t, c, gp, cs=0, max_grid=None):
    h, w = t.shape
    mg = max_grid or (int(np.ceil(w / gp.sigma_x)) + 1, int(np.ceil(h / gp.sigma_y)) + 1, 256)
    s = S.HFBS(w, h, *mg)
    s.SetImageDim(w, h)
    s.SetFilterParams(gp)
    s.ClearAll()
    s.InitFrame(cs, sc_rgba, t, c)
    s.Optimize()
    return s, s.Download(S.OPTIMIZED_QUANTITY)


def _check(g, o, rng, what):
    err = float(np.max(np.abs(g.astype(np.float64) - o)))
    print(f"[{what}] max-abs {err:.3e} range {rng:.3e} -> {err / rng:.3e}")
    assert err <= TOL * rng, what


def test_hfbs_grids_parity(S, oracle_mod):
    sc = scenes.colorization_scene(203, 117, seed=77, regions=9, scribble_frac=0.05)
    gp, op = _params(S, oracle_mod, sigma_x=8, sigma_y=8, sigma_l=8, optim_iter_max=10)
    s, out = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp)
    gr = oracle_mod.hfbs_grids(sc.rgba, sc.target_u, sc.confidence, op)
    assert s.grid_dims() == tuple(reversed(gr["b"].shape))
    np.testing.assert_array_equal(s.debug_grid(2), gr["splat_sum"].astype(np.float32))
    _check(s.debug_grid(0), gr["b"], max(1e-6, np.abs(gr["b"]).max()), "b")
    _check(s.debug_grid(3), gr["splat_norm"], gr["splat_norm"].max(), "splat_norm")
    _check(s.debug_grid(4), gr["inv_diag_a"], gr["inv_diag_a"].max(), "inv_diag_A")
    _check(s.debug_grid(1), gr["splat_c"], gr["splat_c"].max(), "blur(c)")
    _check(s.debug_grid(5), gr["z_final"], 0.8, "z")


@pytest.mark.parametrize("dims,sig", [((320, 240), (8, 8, 8)), ((257, 131), (5.5, 7, 4)),
                                      ((64, 1), (8, 8, 16)), ((1, 50), (4, 4, 32))])
def test_hfbs_colorization_parity(S, oracle_mod, dims, sig):
    w, h = dims
    sc = scenes.colorization_scene(w, h, seed=w + h, regions=7, scribble_frac=0.04)
    gp, op = _params(S, oracle_mod, sigma_x=sig[0], sigma_y=sig[1], sigma_l=sig[2],
                     optim_iter_max=20)
    s, out = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp)
    ref, dims_o, fell = oracle_mod.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, op)
    assert not fell and not s.fell_back
    _check(out, ref, 0.8, f"hfbs {dims} {sig}")


def test_hfbs_ycbcr_depth_parity(S, oracle_mod):
    sc = scenes.depth_scene(320, 240, 8, seed=5)
    gp, op = _params(S, oracle_mod, sigma_x=16, sigma_y=16, sigma_l=16, optim_iter_max=30)
    s, out = _run_gpu(S, sc.rgba, sc.target, sc.confidence, gp, cs=1)
    g8 = oracle_mod.convert_color(sc.rgba, oracle_mod.YCBCR)
    ref, _, fell = oracle_mod.hfbs_optimize(g8, sc.target, sc.confidence, op)
    _check(out, ref, float(sc.target.max() - sc.target.min()), "hfbs depth")


def test_hfbs_guard_falls_back(S, oracle_mod):
    sc = scenes.colorization_scene(64, 32, seed=1, regions=3)
    gp, op = _params(S, oracle_mod, sigma_x=8, sigma_y=8, sigma_l=8)
    s, out = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp, max_grid=(4, 4, 32))
    assert s.fell_back
    np.testing.assert_array_equal(out, sc.target_u)


def test_hfbs_graph_replays(S, oracle_mod):
    sc = scenes.colorization_scene(320, 240, seed=3, regions=6, scribble_frac=0.02)
    gp, _ = _params(S, oracle_mod, sigma_x=8, sigma_y=8, sigma_l=8, optim_iter_max=20)
    s, eager = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp)
    s.SetGraphMode(True)
    for _ in range(4):
        s.Optimize()
        rep = s.Download(S.OPTIMIZED_QUANTITY)
        # LDS float atomics in the splat may add in another order.
        np.testing.assert_allclose(rep, eager, rtol=0, atol=1e-5)
        s.Synchronize()


@pytest.mark.parametrize("kb,k", [(0, 5), (1, 5), (3, 0), (2, 3)])
def test_hfbs_iteration_counts(S, oracle_mod, kb, k):
    """No, one and an odd number of bistochastization steps (the inverse
    diagonal comes from the last step, or from the starting n = 1 when there
    is none) and no heavy-ball step (the initial solution is sliced)."""
    sc = scenes.colorization_scene(97, 61, seed=kb * 10 + k, regions=5, scribble_frac=0.05)
    gp, op = _params(S, oracle_mod, sigma_x=8, sigma_y=8, sigma_l=8, bistoch_iter_max=kb,
                     optim_iter_max=k)
    s, out = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp)
    gr = oracle_mod.hfbs_grids(sc.rgba, sc.target_u, sc.confidence, op)
    _check(s.debug_grid(3), gr["splat_norm"], gr["splat_norm"].max(), f"n Kb={kb}")
    _check(s.debug_grid(4), gr["inv_diag_a"], gr["inv_diag_a"].max(), f"invA Kb={kb}")
    ref, _, fell = oracle_mod.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, op)
    assert not fell
    _check(out, ref, 0.8, f"hfbs Kb={kb} K={k}")


@pytest.mark.parametrize("beta", [0.9, 0.0])
def test_hfbs_long_lines_general_sweeps(S, oracle_mod, beta):
    """Grid lines beyond the register sweeps: 300 voxels along y (sigma_y = 1)
    take the per-line sweep kernel; with beta = 0 every axis does, and the
    initial solution comes from its own kernel."""
    sc = scenes.colorization_scene(40, 300, seed=11, regions=4, scribble_frac=0.05)
    gp, op = _params(S, oracle_mod, sigma_x=8, sigma_y=1, sigma_l=1, optim_iter_max=5,
                     neg_exp_beta=beta)
    s, out = _run_gpu(S, sc.rgba, sc.target_u, sc.confidence, gp, max_grid=(6, 301, 256))
    assert not s.fell_back
    assert s.grid_dims() == (5, 300, 255)
    gr = oracle_mod.hfbs_grids(sc.rgba, sc.target_u, sc.confidence, op)
    _check(s.debug_grid(1), gr["splat_c"], gr["splat_c"].max(), "blur(c)")
    ref, _, fell = oracle_mod.hfbs_optimize(sc.rgba, sc.target_u, sc.confidence, op)
    assert not fell
    _check(out, ref, 0.8, f"hfbs long lines beta={beta}")
