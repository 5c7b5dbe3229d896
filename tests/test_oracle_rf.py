"""Pins of the FP64 RF-DT oracle (oracle/rf_oracle.c; the recursive-filter domain
transform is not in the reference): closed-form impulse responses, the variance
schedule, affine invariance, edge stopping, the exact transpose of the input adjoint
(dot-product test) and the guide gradient against central finite differences."""
import numpy as np
import pytest


def _guide(rng, h, w, levels=256):
    return rng.integers(0, levels, (h, w, 4)).astype(np.float64)


def test_line_impulse_response_closed_form(oracle_mod):
    # Causal then anticausal with constant w on a long line: the impulse
    # response is (1 - w) / (1 + w) * w^|k| away from the ends.
    L, c, w = 401, 200, 0.7
    x = np.zeros(L)
    x[c] = 1.0
    z = oracle_mod.rf_line(x, np.full(L, w))
    k = np.arange(-30, 31)
    np.testing.assert_allclose(z[c + k], (1 - w) / (1 + w) * w ** np.abs(k), rtol=1e-12)
    assert abs(z.sum() - 1.0) < 1e-12


def test_sigma_schedule_sums_variances(oracle_mod):
    for n in (1, 2, 3, 5):
        s = [oracle_mod.rf_sigma_i(10.0, i, n) for i in range(n)]
        assert abs(sum(v * v for v in s) - 100.0) < 1e-9
        assert all(s[i] == pytest.approx(2 * s[i + 1]) for i in range(n - 1))


def test_flat_guide_is_separable_exponential(oracle_mod):
    # Constant guide: d = 1, so one iteration is the separable closed form
    # h(kx) h(ky) with w = exp(-sqrt(2) / sigma) per direction.
    H, W, cy, cx = 161, 181, 80, 90
    g = np.full((H, W, 4), 77.0)
    x = np.zeros((H, W))
    x[cy, cx] = 1.0
    sx, sy = 4.0, 3.0
    out = oracle_mod.rf_filter(g, sx, sy, 0.5, 1, x)
    wx, wy = np.exp(-np.sqrt(2) / sx), np.exp(-np.sqrt(2) / sy)
    kx, ky = np.arange(-20, 21), np.arange(-15, 16)
    ref = np.outer((1 - wy) / (1 + wy) * wy ** np.abs(ky), (1 - wx) / (1 + wx) * wx ** np.abs(kx))
    np.testing.assert_allclose(out[cy + ky[0]:cy + ky[-1] + 1, cx + kx[0]:cx + kx[-1] + 1], ref,
                               rtol=1e-10, atol=1e-15)


def test_constant_input_is_preserved(oracle_mod):
    rng = np.random.default_rng(1)
    g = _guide(rng, 23, 31)
    out = oracle_mod.rf_filter(g, 12.0, 9.0, 0.3, 3, np.full((23, 31), 2.5))
    np.testing.assert_allclose(out, 2.5, rtol=0, atol=1e-12)


def test_strong_edge_stops_diffusion(oracle_mod):
    rng = np.random.default_rng(2)
    H, W = 20, 40
    g = np.zeros((H, W, 4))
    g[:, 20:, :3] = 255.0                       # vertical edge between columns 19 and 20
    x1 = rng.normal(size=(H, W))
    x2 = x1.copy()
    x2[:, 20:] += 10.0 * rng.normal(size=(H, 20))
    o1 = oracle_mod.rf_filter(g, 30.0, 30.0, 0.01, 3, x1)
    o2 = oracle_mod.rf_filter(g, 30.0, 30.0, 0.01, 3, x2)
    np.testing.assert_allclose(o1[:, :20], o2[:, :20], rtol=0, atol=1e-12)
    assert np.abs(o1[:, 20:] - o2[:, 20:]).max() > 1.0


@pytest.mark.parametrize("shape,n", [((13, 19), 3), ((1, 30), 2), ((25, 1), 1), ((9, 9), 4)])
def test_input_adjoint_is_exact_transpose(oracle_mod, shape, n):
    rng = np.random.default_rng(3)
    h, w = shape
    g = _guide(rng, h, w)
    x, yb = rng.normal(size=shape), rng.normal(size=shape)
    y = oracle_mod.rf_filter(g, 7.0, 5.0, 0.4, n, x)
    xb, _ = oracle_mod.rf_filter_adjoint(g, 7.0, 5.0, 0.4, n, x, yb)
    assert abs(np.vdot(yb, y) - np.vdot(xb, x)) <= 1e-12 * max(1.0, abs(np.vdot(yb, y)))


@pytest.mark.parametrize("sr", [0.4, 2.0])
def test_guide_gradient_matches_central_differences(oracle_mod, sr):
    rng = np.random.default_rng(4)
    h, w, n = 11, 14, 3
    g = _guide(rng, h, w, levels=64) * 4.0
    x, yb = rng.normal(size=(h, w)), rng.normal(size=(h, w))
    _, gb = oracle_mod.rf_filter_adjoint(g, 6.0, 5.0, sr, n, x, yb)
    assert np.abs(gb).max() > 1e-6
    hstep = 1e-4
    for k in range(16):
        v = rng.normal(size=g.shape)
        fp = np.vdot(yb, oracle_mod.rf_filter(g + hstep * v, 6.0, 5.0, sr, n, x))
        fm = np.vdot(yb, oracle_mod.rf_filter(g - hstep * v, 6.0, 5.0, sr, n, x))
        fd = (fp - fm) / (2 * hstep)
        an = float(np.vdot(gb, v))
        assert abs(fd - an) <= 1e-6 * max(1e-3, abs(an)), (k, fd, an)
