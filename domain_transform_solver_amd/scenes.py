"""Seeded synthetic scenes for the depth super-resolution, colorization and
stereo workloads.

The reference's CLIs load real images (src/depth_super_res_dataset.h:38-85,
src/benchmark_colorization_cli.cpp:218-283, src/benchmark_stereo_cli.cpp:120-200);
these generators build inputs of the same kind and shape from a seed so that
every run sees identical data.  Regions are Voronoi cells of uniformly placed
sites (the lowest site index wins ties).  Guides are RGBA uint8 with alpha 0,
as the reference's image loader leaves it (src/domain_transform_image_io.h:15-21).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

SIZES = [(320, 240), (640, 480), (1280, 720), (1920, 1080)]
FACTORS = [4, 8, 16]


def voronoi_labels(w: int, h: int, sites: np.ndarray) -> np.ndarray:
    """Label of the nearest site for every pixel (float64 distances)."""
    labels = np.empty((h, w), np.int32)
    xs = np.arange(w, dtype=np.float64)
    sx = sites[:, 0][None, :]
    sy = sites[:, 1][None, :]
    dx2 = (xs[:, None] - sx) ** 2  # (w, n)
    rows = max(1, (1 << 20) // max(1, w * sites.shape[0]))
    for y0 in range(0, h, rows):
        ys = np.arange(y0, min(h, y0 + rows), dtype=np.float64)
        d = dx2[None, :, :] + ((ys[:, None] - sy) ** 2)[:, None, :]
        labels[y0:y0 + len(ys)] = np.argmin(d, axis=2)
    return labels


def _sites(rng, n, w, h):
    return np.stack([rng.uniform(0, w, n), rng.uniform(0, h, n)], axis=1)


@dataclass
class DepthScene:
    rgba: np.ndarray        # (H, W, 4) uint8, RGB guide
    target: np.ndarray      # (H, W) float32, nearest-upsampled noisy low-res depth
    confidence: np.ndarray  # (H, W) float32, 1 at low-res sample centres
    gt: np.ndarray          # (H, W) float32, ground-truth depth in metres
    factor: int
    seed: int


def depth_scene(w: int = 320, h: int = 240, factor: int = 8, seed: int = 1000,
                regions: int = 12) -> DepthScene:
    """Voronoi-plane depth scene: each region has a random colour (+N(0,4^2)
    noise) and a random plane depth in [0.5, 5.0] m.  The input is the
    factor x factor block mean plus N(0, 0.01^2), nearest-upsampled;
    confidence is 1 at the block centres."""
    rng = np.random.default_rng(seed)
    sites = _sites(rng, regions, w, h)
    lab = voronoi_labels(w, h, sites)
    colors = rng.integers(0, 256, size=(regions, 3)).astype(np.float64)
    d0 = rng.uniform(1.0, 4.5, regions)
    gxs = rng.uniform(-1.5, 1.5, regions) / w
    gys = rng.uniform(-1.5, 1.5, regions) / h
    xs = np.arange(w)[None, :]
    ys = np.arange(h)[:, None]
    gt = d0[lab] + gxs[lab] * (xs - sites[lab, 0]) + gys[lab] * (ys - sites[lab, 1])
    gt = np.clip(gt, 0.5, 5.0)
    rgb = colors[lab] + rng.normal(0.0, 4.0, size=(h, w, 3))
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., :3] = np.clip(np.rint(rgb), 0, 255).astype(np.uint8)
    lw, lh = w // factor, h // factor
    blocks = gt[: lh * factor, : lw * factor].reshape(lh, factor, lw, factor).mean(axis=(1, 3))
    low = blocks + rng.normal(0.0, 0.01, size=blocks.shape)
    target = np.empty((h, w), np.float64)
    yi = np.minimum(np.arange(h) // factor, lh - 1)
    xi = np.minimum(np.arange(w) // factor, lw - 1)
    target[:] = low[yi[:, None], xi[None, :]]
    conf = np.zeros((h, w), np.float32)
    cy = np.arange(lh) * factor + factor // 2
    cx = np.arange(lw) * factor + factor // 2
    conf[np.ix_(cy[cy < h], cx[cx < w])] = 1.0
    return DepthScene(rgba, target.astype(np.float32), conf, gt.astype(np.float32),
                      factor, seed)


@dataclass
class ColorizationScene:
    rgba: np.ndarray        # (H, W, 4) uint8, YYY0 gray guide
    target_u: np.ndarray    # (H, W) float32 scribbled U (0 elsewhere)
    target_v: np.ndarray
    confidence: np.ndarray  # (H, W) float32 1 at scribbles
    gt_u: np.ndarray        # (H, W) float32 region chroma
    gt_v: np.ndarray
    seed: int


def colorization_scene(w: int = 1920, h: int = 1080, seed: int = 2000,
                       regions: int = 40, scribble_frac: float = 0.01
                       ) -> ColorizationScene:
    """Gray Voronoi scene with a random linear luma ramp per region plus
    N(0, (2/255)^2) noise; 1% of pixels carry the region's chroma
    (U, V uniform in [-0.4, 0.4]) with confidence 1."""
    rng = np.random.default_rng(seed)
    sites = _sites(rng, regions, w, h)
    lab = voronoi_labels(w, h, sites)
    base = rng.uniform(0.15, 0.85, regions)
    ax = rng.uniform(-0.3, 0.3, regions) / w
    ay = rng.uniform(-0.3, 0.3, regions) / h
    xs = np.arange(w)[None, :]
    ys = np.arange(h)[:, None]
    luma = base[lab] + ax[lab] * (xs - sites[lab, 0]) + ay[lab] * (ys - sites[lab, 1])
    luma = luma + rng.normal(0.0, 2.0 / 255.0, size=(h, w))
    g8 = np.clip(np.rint(luma * 255.0), 0, 255).astype(np.uint8)
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., 0] = g8
    rgba[..., 1] = g8
    rgba[..., 2] = g8
    cu = rng.uniform(-0.4, 0.4, regions)
    cv = rng.uniform(-0.4, 0.4, regions)
    gt_u = cu[lab].astype(np.float32)
    gt_v = cv[lab].astype(np.float32)
    n = w * h
    k = max(1, int(round(scribble_frac * n)))
    idx = rng.choice(n, size=k, replace=False)
    conf = np.zeros(n, np.float32)
    conf[idx] = 1.0
    tu = np.zeros(n, np.float32)
    tv = np.zeros(n, np.float32)
    tu[idx] = gt_u.reshape(-1)[idx]
    tv[idx] = gt_v.reshape(-1)[idx]
    return ColorizationScene(rgba, tu.reshape(h, w), tv.reshape(h, w), conf.reshape(h, w),
                             gt_u, gt_v, seed)


@dataclass
class StereoScene:
    rgba: np.ndarray        # (H, W, 4) uint8 left RGB guide
    rgba_right: np.ndarray  # (H, W, 4) uint8 right view warped from the left
    target: np.ndarray      # (H, W) float32 noisy initial disparity
    confidence: np.ndarray  # (H, W) float32 0.1 on outliers else 1
    gt: np.ndarray          # (H, W) float32 ground-truth disparity
    seed: int


def stereo_scene(w: int = 1280, h: int = 720, seed: int = 3000, regions: int = 30,
                 max_disp: float = 128.0) -> StereoScene:
    """30 fronto-parallel textured planes with disparity uniform in
    [0, max_disp]; the initial disparity is GT + N(0,1) with 20% uniform
    outliers (confidence 0.1).  The right view samples the left texture of
    each plane at x + d (nearest), which is what a rectified pair of
    fronto-parallel planes shows away from occlusions."""
    rng = np.random.default_rng(seed)
    sites = _sites(rng, regions, w, h)
    lab = voronoi_labels(w, h, sites)
    disp = rng.uniform(0.0, max_disp, regions)
    gt = disp[lab]
    base = rng.integers(40, 216, size=(regions, 3)).astype(np.float64)
    # Per-plane texture: a fixed random pattern per region in plane coordinates.
    tex_w = w + int(max_disp) + 2
    pattern = rng.normal(0.0, 25.0, size=(regions, 3, h, tex_w)).astype(np.float32) \
        if regions * 3 * h * tex_w <= 64_000_000 else None
    if pattern is None:
        # Large scenes: a separable texture keeps memory small.
        px = rng.normal(0.0, 18.0, size=(regions, 3, tex_w)).astype(np.float32)
        py = rng.normal(0.0, 18.0, size=(regions, 3, h)).astype(np.float32)

        def tex(r, c, y, x):
            return px[r, c, x] + py[r, c, y]
    else:
        def tex(r, c, y, x):
            return pattern[r, c, y, x]
    ys, xs = np.mgrid[0:h, 0:w]
    left = np.empty((h, w, 3), np.float64)
    for c in range(3):
        left[..., c] = base[lab, c] + tex(lab, c, ys, xs)
    # Right view: pixel (xr, y) shows the plane of the left pixel xr + d.
    right = np.empty((h, w, 3), np.float64)
    xl = np.clip(np.rint(xs + gt).astype(np.int64), 0, w - 1)
    lab_r = lab[ys, xl]
    xr_src = np.clip(np.rint(xs + disp[lab_r]).astype(np.int64), 0, tex_w - 1)
    for c in range(3):
        right[..., c] = base[lab_r, c] + tex(lab_r, c, ys, xr_src)
    rgba = np.zeros((h, w, 4), np.uint8)
    rgba[..., :3] = np.clip(np.rint(left), 0, 255).astype(np.uint8)
    rgba_r = np.zeros((h, w, 4), np.uint8)
    rgba_r[..., :3] = np.clip(np.rint(right), 0, 255).astype(np.uint8)
    target = gt + rng.normal(0.0, 1.0, size=(h, w))
    out = rng.random((h, w)) < 0.2
    target[out] = rng.uniform(0.0, max_disp, size=int(out.sum()))
    conf = np.where(out, 0.1, 1.0).astype(np.float32)
    return StereoScene(rgba, rgba_r, target.astype(np.float32), conf,
                       gt.astype(np.float32), seed)


def batch_specs(n: int = 64, seed: int = 4000):
    """Frame sizes and upsampling factors for the batched workload: sizes
    uniform over SIZES, factors uniform over FACTORS."""
    rng = np.random.default_rng(seed)
    si = rng.integers(0, len(SIZES), n)
    fi = rng.integers(0, len(FACTORS), n)
    return [(SIZES[a][0], SIZES[a][1], FACTORS[b]) for a, b in zip(si, fi)]


# --------------------------------------------------------------- task metrics
def rmse(a, b) -> float:
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return float(np.sqrt(np.mean(d * d)))


def psnr(a, b, peak: float = 1.0) -> float:
    """PSNR of two chroma planes for a signal of the given peak-to-peak range."""
    m = float(np.mean((np.asarray(a, np.float64) - np.asarray(b, np.float64)) ** 2))
    return float("inf") if m == 0 else 10.0 * np.log10(peak * peak / m)


def bad_pixel_rate(d, gt, thresh: float = 1.0) -> float:
    return float(np.mean(np.abs(np.asarray(d, np.float64) - np.asarray(gt, np.float64)) > thresh))


# ------------------------------------------------- solver presets (reference)
def depth_sr_params(factor: int):
    """DtsSuperResSolver presets (src/benchmark_depth_super_res_cli.cpp:45-85)
    with the solve length of the depth workloads (10 iterations)."""
    s = 8.0 * factor
    return dict(color_space=1, sigma_x=s, sigma_y=s, sigma_r=0.1,
                num_iter=10, loss=1, lam=0.99, step=0.99,
                sigma_z=float(np.float32(8 * factor / (1 + 0.2 * 4))), overwrite=-1.0)


def colorization_params():
    """DtsColorizationSolver presets (src/benchmark_colorization_cli.cpp:298-359)
    with 20 iterations."""
    return dict(color_space=0, sigma_x=64.0, sigma_y=64.0, sigma_r=0.25,
                num_iter=20, loss=2, lam=0.99, step=0.99, sigma_z=30.0, overwrite=0.85)


def stereo_params():
    """DtsStereoSolver presets (src/benchmark_stereo_cli.cpp:242-261) with 15
    iterations and the photometric weight 0.01."""
    return dict(color_space=0, sigma_x=8.0, sigma_y=8.0, sigma_r=0.1,
                num_iter=15, loss=2, lam=0.99, step=0.99, sigma_z=30.0,
                overwrite=-1.0, photo_lambda=0.01)
