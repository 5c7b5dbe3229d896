"""The FP64 oracle reproduces the committed golden fixtures
(tests/golden/make_fixtures.py): integer work exactly, floating point to
1e-12 relative."""
import glob
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DTS = sorted(glob.glob(os.path.join(GOLDEN, "dts_*.npz")))


def _close(a, b, rel=1e-12):
    scale = max(float(np.abs(b).max()), 1e-300)
    assert float(np.abs(np.asarray(a, np.float64) - b).max()) <= rel * scale


@pytest.mark.parametrize("path", DTS, ids=[os.path.basename(p) for p in DTS])
def test_oracle_reproduces_dts_fixture(oracle_mod, path):
    d = np.load(path)
    t = d["target"]
    h, w = t.shape
    fp, op = d["filter_params"], d["optimize_params"]
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(int(d["color_space"]), d["rgba"], t, d["confidence"])
    o.compute_differential(*map(float, fp))
    o.integrate()
    np.testing.assert_array_equal(o.ctH_fx(), d["ctH"])
    np.testing.assert_array_equal(o.ctV_fx(), d["ctV"])
    for vert, key in ((False, "bounds_h"), (True, "bounds_v")):
        lo, up = o.bounds(float(op[3]), vert)
        np.testing.assert_array_equal(np.stack([lo, up]), d[key])
    K, loss = int(op[0]), int(op[1])
    o.optimize(K, loss, float(op[2]), float(op[3]), float(op[4]), float(d["overwrite"]))
    _close(o.result(), d["result"])
    if "tbar" in d.files:
        tb, cb = o.optimize_adjoint(K, loss, float(op[2]), float(op[3]), float(op[4]),
                                    float(d["overwrite"]), d["ybar"])
        _close(tb, d["tbar"], 1e-10)
        _close(cb, d["cbar"], 1e-10)
    if "filter_out" in d.files:
        _close(o.filter(*map(float, fp), d["filter_in"], 3), d["filter_out"])


def test_oracle_reproduces_hfbs_fixture(oracle_mod):
    d = np.load(os.path.join(GOLDEN, "hfbs_colorization_48x32.npz"))
    sx, sy, sl, kb, k = map(float, d["params"])
    prm = oracle_mod.hfbs_params(sx, sy, sl, optim_iter_max=int(k), bistoch_iter_max=int(kb))
    ref, _, _ = oracle_mod.hfbs_optimize(d["rgba"], d["target"], d["confidence"], prm)
    _close(ref, d["result"])
