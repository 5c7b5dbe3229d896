"""Independent pins for the oracle's Charbonnier gradient, stereo photometric
term and confidence map, and for the HFBS bistochastization, inverse diagonal,
heavy-ball solve and slice.  CPU only.

Each expected value is built here from its definition with explicit windows,
explicit neighbour lists, a sparse direct solve or an interpolation identity,
not from the oracle's own arithmetic, and each case is chosen so that a
plausible slip (a sign, a swapped operand, a missing factor, an off-by-one
range check) moves the result by far more than the tolerance.
"""
import math

import numpy as np
import pytest

from test_oracle import _brute_bounds, _brute_pass, _random_guide

FX = 2.0 ** 32
K_INV255 = float(np.float32(0.00392156863))
K_INV255SQ = float(np.float32(0.0000153787))


# ----------------------------------------------------------------- Charbonnier
def test_charbonnier_three_iterations_explicit(oracle_mod):
    """src/domain_transform_solver.cuh:18-23 and :53-79: the Charbonnier data
    gradient (c/n) x / sqrt(x^2 + 1e-6), x = v - t, enters the X half with the
    factor 1/2.  From the second iteration on v != t, so the sign of x, the c/n
    weight and the epsilon all matter; c is kept small so the fixed step stays
    well conditioned (test_charbonnier_step_conditioning)."""
    rng = np.random.default_rng(23)
    w, h = 18, 13
    g = _random_guide(rng, w, h)
    t = rng.normal(size=(h, w)).astype(np.float32)
    c = rng.uniform(0, 0.3, (h, w)).astype(np.float32)
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.RGB, g, t, c)
    o.compute_differential(7.0, 7.0, 1.5)
    o.integrate()
    sigma, lam, step = 5.0, float(np.float32(0.7)), float(np.float32(0.9))
    o.optimize(3, oracle_mod.CHARBONNIER, lam, sigma, step)
    got = o.result()
    t = t.astype(np.float64)
    c = c.astype(np.float64)
    ctH, ctV = o.ctH_fx(), o.ctV_fx()
    n = np.empty((h, w))
    for y in range(h):
        lo, up = _brute_bounds(ctH[y], sigma)
        n[y] = up - lo
    v = t.copy()
    no_data = t.copy()
    for _ in range(3):
        B = _brute_pass(v, ctH, sigma, False)
        x = v - t
        gdat = (c / n) * x / np.sqrt(x * x + 1e-6)
        v = v - step * (lam * (v - B) + 0.5 * gdat)
        v = v - step * lam * (v - _brute_pass(v, ctV, sigma, True))
        B = _brute_pass(no_data, ctH, sigma, False)
        no_data = no_data - step * lam * (no_data - B)
        no_data = no_data - step * lam * (no_data - _brute_pass(no_data, ctV, sigma, True))
    np.testing.assert_allclose(got, v, rtol=0, atol=1e-10)
    # The data term is large enough here that dropping it would be seen.
    assert np.abs(no_data - v).max() > 1e-2


# --------------------------------------------------------------------- stereo
def _dH(row, x, q):
    """Backward differential of one RGBA row at x (pixel -1 is black)."""
    prev = row[x - 1].astype(np.int64) if x > 0 else np.zeros(4, np.int64)
    ssd = float(((row[x].astype(np.int64) - prev) ** 2).sum())
    return math.sqrt(ssd * K_INV255SQ * q * q + 1.0)


def _photo_grad(L, R, doff, v, y, x, q_lr, grad_alpha, photo_lambda):
    """src/domain_transform_loss_term.cuh:20-39 with the stereo solver's
    arguments (src/domain_transform_stereo.cu:77-87): gather the right view at
    xr = int(doff + x - v) when 0 <= doff + x - v < W."""
    W = L.shape[1]
    xi = doff + x - v
    if not (0.0 <= xi < W):
        return 0.0
    xr = int(xi)
    sad = float(np.abs(L[y, x].astype(np.int64) - R[y, xr].astype(np.int64)).sum()) * K_INV255
    dhr = _dH(R[y], xr, q_lr)
    return photo_lambda * 2.0 * sad * ((dhr * dhr - 1.0) * grad_alpha)


def _stereo_case():
    rng = np.random.default_rng(31)
    w, h = 12, 3
    L = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    R = rng.integers(0, 256, (h, w, 4)).astype(np.uint8)
    doff = 2.5
    t = (rng.integers(0, 17, (h, w)) * 0.25).astype(np.float32)
    # Range edges of xi = doff + x - v: -0.5 (outside although int() gives 0),
    # exactly 0 (inside), 11.75 (inside, xr = W - 1), 12.0 = W (outside).
    t[0, 0] = 3.0
    t[0, 1] = 3.5
    t[1, 11] = 1.75
    t[1, 10] = 0.5
    t[2, 11] = 1.25
    c = rng.uniform(0.2, 1.0, (h, w)).astype(np.float32)
    return L, R, doff, t, c


def test_stereo_photometric_term_one_step(oracle_mod):
    """lambda = 0 and K = 1: the result is t - s/2 * g_photo(t) per pixel (the
    data term is 0 at v = t and the Y half does nothing).  g_photo uses the
    right view's own differential dH_R at xr with the factor
    (dH_R^2 - 1) (sigma_r / sigma_x)^2, the SAD over all four channels, and the
    range test on the real-valued xi."""
    L, R, doff, t, c = _stereo_case()
    h, w = t.shape
    sx, sy, sr = 8.0, 8.0, 0.5
    step, plam = float(np.float32(0.8)), float(np.float32(0.3))
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.RGB, L, t, c)
    o.init_right_frame(doff, oracle_mod.RGB, R)
    o.compute_differential(sx, sy, sr)
    o.integrate()
    o.process_right_frame()
    o.optimize_stereo(1, oracle_mod.L2, 0.0, 6.0, step, plam)
    q_lr = float(np.float32(sx) / np.float32(sr))
    qa = np.float32(sr) / np.float32(sx)
    grad_alpha = float(np.float32(qa * qa))
    exp = np.empty((h, w))
    inside = 0
    for y in range(h):
        for x in range(w):
            v = float(t[y, x])
            gp = _photo_grad(L, R, doff, v, y, x, q_lr, grad_alpha, plam)
            inside += 0.0 <= doff + x - v < w
            exp[y, x] = v - step * 0.5 * gp
    np.testing.assert_allclose(o.result(), exp, rtol=0, atol=1e-12)
    # The edge pixels of the range test: unchanged outside, moved inside.
    assert o.result()[0, 0] == float(t[0, 0])
    assert o.result()[1, 10] == float(t[1, 10])
    assert o.result()[2, 11] == float(t[2, 11])
    assert o.result()[0, 1] != float(t[0, 1])
    assert o.result()[1, 11] != float(t[1, 11])
    assert 0 < inside < h * w


def test_stereo_term_with_smoothing_two_iterations(oracle_mod):
    """lambda > 0, K = 2, L2: the photometric term is evaluated at the pass
    input v (not at the target), beside the box step with explicit windows."""
    L, R, doff, t, c = _stereo_case()
    h, w = t.shape
    sx, sy, sr = 8.0, 8.0, 0.5
    sigma, lam = 3.0, float(np.float32(0.6))
    step, plam = float(np.float32(0.8)), float(np.float32(0.3))
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.RGB, L, t, c)
    o.init_right_frame(doff, oracle_mod.RGB, R)
    o.compute_differential(sx, sy, sr)
    o.integrate()
    o.process_right_frame()
    o.optimize_stereo(2, oracle_mod.L2, lam, sigma, step, plam)
    ctH, ctV = o.ctH_fx(), o.ctV_fx()
    q_lr = float(np.float32(sx) / np.float32(sr))
    qa = np.float32(sr) / np.float32(sx)
    grad_alpha = float(np.float32(qa * qa))
    tt = t.astype(np.float64)
    cc = c.astype(np.float64)
    n = np.empty((h, w))
    for y in range(h):
        lo, up = _brute_bounds(ctH[y], sigma)
        n[y] = up - lo
    v = tt.copy()
    for _ in range(2):
        B = _brute_pass(v, ctH, sigma, False)
        g = 2 * (cc / n) * (v - tt)
        for y in range(h):
            for x in range(w):
                g[y, x] += _photo_grad(L, R, doff, v[y, x], y, x, q_lr, grad_alpha, plam)
        v = v - step * (lam * (v - B) + 0.5 * g)
        v = v - step * lam * (v - _brute_pass(v, ctV, sigma, True))
    np.testing.assert_allclose(o.result(), v, rtol=0, atol=1e-12)


# ----------------------------------------------------------------- confidence
def _flat_ct_fx(n, first_d):
    """32.32 ct of a line whose first step is first_d and every other step 1."""
    d0 = int(round(first_d * FX))
    return d0 + np.arange(n, dtype=np.int64) * int(FX)


def _brute_filter(v, ctH, ctV, sx, sy, num_iter):
    """src/domain_transform_filter.cu:124-188 with explicit windows: N
    iterations of a row pass then a column pass, the sigmas scaled (in fp32)
    by 1 and then 3 * 2^(N-i-1) / sqrt(4^N - 1)."""
    for it in range(num_iter):
        m = 1.0 if it == 0 else float(np.float32(3 * 2 ** (num_iter - (it + 1)) /
                                                 math.sqrt(4 ** num_iter - 1)))
        s1 = float(np.float32(np.float32(sx) * np.float32(m)))
        s2 = float(np.float32(np.float32(sy) * np.float32(m)))
        v = _brute_pass(v, ctH, s1, False)
        v = _brute_pass(v, ctV, s2, True)
    return v


def test_confidence_step_edge_flat_guide(oracle_mod):
    """src/domain_transform_solver.cu:206-246 and :21-44: conf =
    exp(-(F(t^2) - F(t)^2) * 0.5 / sigma0^2) with F the 3-iteration domain
    transform filter, 0 at x <= left_clear.  A flat guide makes ct a closed
    form (d = 1 except the first pixel of each line, whose neighbour outside
    the image is black), and a step-edge target has a known, non-zero local
    variance along the edge."""
    w, h = 30, 20
    val = (10, 20, 30, 0)
    g = np.zeros((h, w, 4), np.uint8)
    g[...] = val
    t = np.where(np.arange(w)[None, :] < 13, 1.0, 3.0) + np.where(
        np.arange(h)[:, None] < 9, 0.0, 0.5)
    t = t.astype(np.float32)
    sx, sy, sr, sigma0, left_clear = 6.0, 5.0, 0.3, 0.8, 4
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.RGB, g, t)
    o.compute_confidence(sx, sy, sr, sigma0, left_clear)
    conf = o.confidence()
    k = K_INV255SQ
    ssd = float(sum(c * c for c in val))
    qx = float(np.float32(sx) / np.float32(sr))
    qy = float(np.float32(sy) / np.float32(sr))
    ctH = np.broadcast_to(_flat_ct_fx(w, math.sqrt(ssd * k * qx * qx + 1)), (h, w))
    ctV = np.broadcast_to(_flat_ct_fx(h, math.sqrt(ssd * k * qy * qy + 1))[:, None], (h, w))
    tt = t.astype(np.float64)
    m1 = _brute_filter(tt, ctH, ctV, sx, sy, 3)
    m2 = _brute_filter(tt * tt, ctH, ctV, sx, sy, 3)
    var = m2 - m1 * m1
    scale = float(np.float32(0.5) / (np.float32(sigma0) * np.float32(sigma0)))
    exp = np.where(var < 0, 0.0, np.exp(-var * scale))
    exp[:, : left_clear + 1] = 0.0
    np.testing.assert_allclose(conf, exp, rtol=1e-12, atol=1e-13)
    # The edge is seen: a band of columns with clearly reduced confidence,
    # flat regions far from both edges at 1.
    assert conf[:, 10:16].min() < 0.6
    assert conf[2, left_clear + 1] > 0.0 and conf[2, left_clear] == 0.0


# ----------------------------------------------------------- HFBS: small grid
def _grid_coord(v, sigma, gmax):
    """Grid coordinate as the reference computes it (fp32 v * 1 / sigma),
    clamped into the grid."""
    gcoord = np.float32(v) * np.float32(1.0) / np.float32(sigma)
    return min(gcoord, np.float32(gmax - 1))


def _brute_splat(rgba, t, c, sx, sy, sl, gx, gy, gl):
    """Per-voxel pixel count S, confidence sum and confidence-weighted target sum."""
    h, w = t.shape
    S = np.zeros((gl, gy, gx))
    C = np.zeros((gl, gy, gx))
    B = np.zeros((gl, gy, gx))
    for y in range(h):
        for x in range(w):
            ix = int(math.floor(_grid_coord(x, sx, gx)))
            iy = int(math.floor(_grid_coord(y, sy, gy)))
            iz = int(math.floor(_grid_coord(int(rgba[y, x, 0]), sl, gl)))
            S[iz, iy, ix] += 1
            C[iz, iy, ix] += float(c[y, x])
            B[iz, iy, ix] += float(c[y, x]) * float(t[y, x])
    return S, C, B


def _neighbour_sum(a, weights=None):
    """Sum over the 6 face neighbours of every voxel (zero outside)."""
    gl, gy, gx = a.shape
    out = np.zeros_like(a)
    for z in range(gl):
        for y in range(gy):
            for x in range(gx):
                s = 0.0
                for dz, dy, dx in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0),
                                   (1, 0, 0)):
                    zz, yy, xx = z + dz, y + dy, x + dx
                    if 0 <= zz < gl and 0 <= yy < gy and 0 <= xx < gx:
                        s += a[zz, yy, xx] * (1.0 if weights is None else weights[zz, yy, xx])
                out[z, y, x] = s
    return out


def _tiny_hfbs_case():
    """16 x 16 pixels on a 2 x 2 x 2 grid (sigma 8, 8, 128): one voxel empty,
    one cell split between both luma bins, confidence zero on part of it."""
    w = h = 16
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., :3] = 50
    cols = np.arange(8)[None, :, None]
    rows = np.arange(8)[:, None, None]
    rgba[:8, 8:, :3] = np.where(cols % 3 == 0, 200, 50)  # cell (1,0): mixed
    rgba[8:, :8, :3] = 200  # cell (0,1): bin 1 only
    rgba[8:, 8:, :3] = np.where(rows % 2 == 0, 255, 90)  # cell (1,1): mixed
    rng = np.random.default_rng(5)
    t = rng.uniform(-0.5, 0.5, (h, w)).astype(np.float32)
    c = rng.uniform(0.1, 1.0, (h, w)).astype(np.float32)
    c[:8, 8:12] = 0.0
    return rgba, t, c


def test_hfbs_splat_bistochastization_inverse_diagonal_tiny_grid(oracle_mod):
    """src/hfbs.cu:71-120 (splat), :125-155 (Jacobi steps
    n' = sqrt(n (S + 1e-5) / (2 n + sum_6 n_nb + 1e-5)) from n = 1) and
    :158-195 (invA = 1 / max(1e-3, lambda (S - 2 n^2) + c_g)) on a 2x2x2 grid
    with hand-placed occupancy: the empty voxel takes the 1e-3 floor."""
    rgba, t, c = _tiny_hfbs_case()
    kb, lam = 5, 8.0
    p = oracle_mod.hfbs_params(8, 8, 128, lambda_=lam, bistoch_iter_max=kb, optim_iter_max=0)
    gr = oracle_mod.hfbs_grids(rgba, t, c, p)
    S, C, B = _brute_splat(rgba, t, c, 8.0, 8.0, 128.0, 2, 2, 2)
    assert S[1, 0, 0] == 0 and S.sum() == 256
    np.testing.assert_array_equal(gr["splat_sum"], S)
    np.testing.assert_allclose(gr["b"], B, rtol=1e-13, atol=1e-13)
    eps = float(np.float32(0.00001))
    n = np.ones_like(S)
    for _ in range(kb):
        n = np.sqrt(n * (S + eps) / (2 * n + _neighbour_sum(n) + eps))
    np.testing.assert_allclose(gr["splat_norm"], n, rtol=1e-13, atol=0)
    diag = lam * (S - 2 * n * n) + C
    exp = 1.0 / np.maximum(1e-3, diag)
    np.testing.assert_allclose(gr["inv_diag_a"], exp, rtol=1e-13, atol=0)
    assert gr["inv_diag_a"][1, 0, 0] == 1e3  # the empty voxel is floored
    assert (diag > 1e-3).sum() >= 5  # the others are not


def _hand_grids():
    """3 x 2 x 1 grid with hand values: b, n, invA, Y0 (gl, gy, gx)."""
    b = np.array([[[0.5, -0.2, 1.0], [0.1, 0.0, -0.4]]])
    n = np.array([[[0.8, 1.1, 0.6], [0.9, 0.7, 1.2]]])
    invA = np.array([[[0.25, 0.5, 0.2], [0.4, 1.0, 0.3]]])
    Y0 = np.array([[[3.0, -0.1, 0.2], [0.05, -2.5, 0.0]]])
    return b, n, invA, Y0


def test_hfbs_heavy_ball_two_hand_steps(oracle_mod):
    """src/hfbs.cu:200-239: q = (b + lambda n sum_6 n_nb Y_nb) invA, g = Y - q,
    Huber clamp of g to +-1 beyond 1.345, h = beta h + g_c, Y = Y - alpha h,
    two steps from h = 0 so the momentum enters.  Two voxels are clamped in
    the first step (one each sign)."""
    b, n, invA, Y0 = _hand_grids()
    lam, alpha, beta = 2.0, 0.5, float(np.float32(0.9))  # the parameters are fp32
    p = oracle_mod.hfbs_params(lambda_=lam, alpha=alpha, beta=beta, optim_iter_max=2)
    got = oracle_mod.hfbs_solve(b, n, invA, Y0, p)
    Y = Y0.copy()
    hm = np.zeros_like(Y)
    clamped = []
    for _ in range(2):
        q = (b + lam * n * _neighbour_sum(Y, n)) * invA
        grad = Y - q
        clamped.append(np.abs(grad) > 1.345)
        gc = np.where(np.abs(grad) > 1.345, np.sign(grad), grad)
        hm = beta * hm + gc
        Y = Y - alpha * hm
    assert clamped[0].sum() == 2 and (~clamped[0]).sum() == 4
    np.testing.assert_allclose(got, Y, rtol=0, atol=1e-14)
    # Without the clamp the result would differ at the clamped voxels.
    p1 = oracle_mod.hfbs_params(lambda_=lam, alpha=alpha, beta=beta, optim_iter_max=1)
    one = oracle_mod.hfbs_solve(b, n, invA, Y0, p1)
    np.testing.assert_allclose(one[clamped[0]], (Y0 - alpha * np.sign(Y0 - (
        b + lam * n * _neighbour_sum(Y0, n)) * invA))[clamped[0]], rtol=0, atol=1e-14)


def test_hfbs_heavy_ball_fixed_point_is_linear_solve(oracle_mod):
    """After many steps the heavy-ball iterate is the fixed point g = 0, i.e.
    the linear system (1/invA) Y - lambda n sum_6 n_nb Y_nb = b.  Solved here
    with a sparse direct solver on a 10 x 8 x 6 grid in which every voxel is
    occupied (so no diagonal is floored)."""
    from scipy.sparse import lil_matrix
    from scipy.sparse.linalg import spsolve

    w, h = 80, 64
    yy, xx = np.mgrid[0:h, 0:w]
    luma = (20 + 45 * ((xx + 3 * yy) % 6)).astype(np.uint8)
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., 0] = rgba[..., 1] = rgba[..., 2] = luma
    t = (0.1 * np.sin(xx / 7.0) * np.cos(yy / 5.0) + 0.03 * ((xx + 3 * yy) % 6)).astype(
        np.float32)
    rng = np.random.default_rng(9)
    c = rng.uniform(0.5, 1.0, (h, w)).astype(np.float32)
    lam = 8.0
    p = oracle_mod.hfbs_params(8, 8, 45, lambda_=lam, alpha=1.0, beta=0.9,
                               bistoch_iter_max=200, optim_iter_max=3000)
    gr = oracle_mod.hfbs_grids(rgba, t, c, p)
    S = gr["splat_sum"]
    assert S.shape == (6, 8, 10) and (S > 0).all()
    n, invA, b = gr["splat_norm"], gr["inv_diag_a"], gr["b"]
    diag = 1.0 / invA
    assert (diag > 1e-3 * 1.01).all()
    gl, gy, gx = S.shape
    N = S.size
    A = lil_matrix((N, N))
    idx = np.arange(N).reshape(S.shape)
    for z in range(gl):
        for y in range(gy):
            for x in range(gx):
                i = idx[z, y, x]
                A[i, i] = diag[z, y, x]
                for dz, dy, dx in ((0, 0, -1), (0, 0, 1), (0, -1, 0), (0, 1, 0), (-1, 0, 0),
                                   (1, 0, 0)):
                    zz, yv, xv = z + dz, y + dy, x + dx
                    if 0 <= zz < gl and 0 <= yv < gy and 0 <= xv < gx:
                        A[i, idx[zz, yv, xv]] -= lam * n[z, y, x] * n[zz, yv, xv]
    Y = spsolve(A.tocsr(), b.ravel()).reshape(S.shape)
    scale = np.abs(Y).max()
    assert scale > 0.01
    np.testing.assert_allclose(gr["z_final"], Y, rtol=0, atol=1e-9 * scale)
    # The initialisation is not already the solution.
    assert np.abs(gr["z_init"] - Y).max() > 1e-3 * scale


def test_hfbs_slice_reproduces_linear_grid(oracle_mod):
    """src/hfbs.cu:244-330 with a uniform splat count: trilinear interpolation
    of Y = a + b x + c y + d z at the pixels' fp32 grid coordinates (clamped
    into the grid) is exact."""
    rng = np.random.default_rng(17)
    w, h = 40, 24
    sx, sy, sl = 8.0, 8.0, 32.0
    gx, gy, gl = 5, 3, 8
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., 0] = rng.integers(0, 256, (h, w))
    rgba[0, :3, 0] = [0, 255, 224]  # luma at the bottom and clamped top of the grid
    a, bx, cy, dz = 0.3, 0.7, -1.1, 0.45
    zi, yi, xi = np.mgrid[0:gl, 0:gy, 0:gx].astype(np.float64)
    Y = a + bx * xi + cy * yi + dz * zi
    S = np.full((gl, gy, gx), 3.0)
    p = oracle_mod.hfbs_params(sx, sy, sl)
    got = oracle_mod.hfbs_slice(rgba, S, Y, p)
    exp = np.empty((h, w))
    for y in range(h):
        for x in range(w):
            fx = float(_grid_coord(x, sx, gx))
            fy = float(_grid_coord(y, sy, gy))
            fz = float(_grid_coord(int(rgba[y, x, 0]), sl, gl))
            exp[y, x] = a + bx * fx + cy * fy + dz * fz
    np.testing.assert_allclose(got, exp, rtol=0, atol=1e-13)


def test_hfbs_slice_zero_count_corner_drops_out(oracle_mod):
    """The slice is sum_8 w_k S_k Y_k / sum_8 w_k S_k over the 8 corners with
    trilinear weights w_k: a corner with S = 0 contributes to neither sum, so
    a huge Y there leaves the result untouched; corners outside the grid
    count as S = 0."""
    rng = np.random.default_rng(18)
    w, h = 40, 24
    sx, sy, sl = 8.0, 8.0, 32.0
    gx, gy, gl = 5, 3, 8
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., 0] = rng.integers(0, 256, (h, w))
    S = rng.integers(1, 9, (gl, gy, gx)).astype(np.float64)
    Y = rng.normal(size=(gl, gy, gx))
    S[3, 1, 2] = 0.0
    Y[3, 1, 2] = 1e6
    p = oracle_mod.hfbs_params(sx, sy, sl)
    got = oracle_mod.hfbs_slice(rgba, S, Y, p)

    def get(a, z, y, x):
        return a[z, y, x] if (0 <= z < gl and 0 <= y < gy and 0 <= x < gx) else 0.0

    exp = np.empty((h, w))
    for y in range(h):
        for x in range(w):
            fx = float(_grid_coord(x, sx, gx))
            fy = float(_grid_coord(y, sy, gy))
            fz = float(_grid_coord(int(rgba[y, x, 0]), sl, gl))
            x0, y0, z0 = math.floor(fx), math.floor(fy), math.floor(fz)
            ax, ay, az = fx - x0, fy - y0, fz - z0
            num = den = 0.0
            for kz in (0, 1):
                for ky in (0, 1):
                    for kx in (0, 1):
                        wgt = ((ax if kx else 1 - ax) * (ay if ky else 1 - ay) *
                               (az if kz else 1 - az))
                        s = get(S, z0 + kz, y0 + ky, x0 + kx)
                        num += wgt * s * get(Y, z0 + kz, y0 + ky, x0 + kx)
                        den += wgt * s
            exp[y, x] = num / den
    assert np.abs(got).max() < 10.0  # the 1e6 never leaks in
    np.testing.assert_allclose(got, exp, rtol=1e-12, atol=1e-12)
