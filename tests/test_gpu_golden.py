"""The GPU path against the committed golden fixtures: ct and window bounds
bit-exact, solver results within 1e-4 of the value range, adjoints within
2e-4 of their largest entry."""
import glob
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DTS = sorted(glob.glob(os.path.join(GOLDEN, "dts_*.npz")))


@pytest.fixture(scope="module")
def S(gpu_lib):
    from domain_transform_solver_amd import solver
    return solver


@pytest.mark.parametrize("path", DTS, ids=[os.path.basename(p) for p in DTS])
def test_gpu_matches_dts_fixture(S, path):
    d = np.load(path)
    t = d["target"]
    h, w = t.shape
    fp, op = d["filter_params"], d["optimize_params"]
    print(os.path.basename(path), flush=True)
    g = S.DomainTransformSolver(w, h)
    g.InitFrame(int(d["color_space"]), d["rgba"], t, d["confidence"])
    fparams = S.DomainFilterParams(*map(float, fp))
    g.ComputeColorSpaceDifferential(fparams)
    g.IntegrateColorDifferentials()
    np.testing.assert_array_equal(g.debug_ct(False), d["ctH"])
    np.testing.assert_array_equal(g.debug_ct(True), d["ctV"])
    for vert, key in ((False, "bounds_h"), (True, "bounds_v")):
        lo, up = g.debug_bounds(float(op[3]), vert)
        np.testing.assert_array_equal(np.stack([lo, up]), d[key])
    rng = float(t.max() - t.min()) or 1.0
    oparams = S.DomainOptimizeParams(int(op[0]), int(op[1]), float(op[2]), float(op[3]),
                                     float(op[4]))
    if "tbar" in d.files:
        g.OptimizeCheckpointed(oparams, float(d["overwrite"]), 0)
        tb, cb = g.OptimizeAdjoint(d["ybar"].astype(np.float32))
        assert np.abs(tb - d["tbar"]).max() <= 2e-4 * np.abs(d["tbar"]).max()
        assert np.abs(cb - d["cbar"]).max() <= 2e-4 * np.abs(d["cbar"]).max()
    else:
        g.Optimize(oparams, float(d["overwrite"]))
    assert np.abs(g.Download(S.OPTIMIZED_QUANTITY) - d["result"]).max() <= 1e-4 * rng
    if "filter_out" in d.files:
        out = g.Filter(fparams, d["filter_in"], 3)
        frng = float(d["filter_in"].max() - d["filter_in"].min())
        assert np.abs(out - d["filter_out"]).max() <= 1e-4 * frng


def test_gpu_matches_hfbs_fixture(S):
    d = np.load(os.path.join(GOLDEN, "hfbs_colorization_48x32.npz"))
    sx, sy, sl, kb, k = map(float, d["params"])
    h, w = d["target"].shape
    s = S.HFBS(w, h, 64, 64, 64)
    s.SetImageDim(w, h)
    s.SetFilterParams(S.HFBSGridParams(sigma_x=sx, sigma_y=sy, sigma_l=sl,
                                       bistoch_iter_max=int(kb), optim_iter_max=int(k)))
    s.InitFrame(0, d["rgba"], d["target"], d["confidence"])
    s.Optimize()
    out = s.Download(S.OPTIMIZED_QUANTITY)
    assert np.abs(out - d["result"]).max() <= 1e-4 * 0.8
