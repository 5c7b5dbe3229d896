"""The depth super-resolution program (domain_transform_solver_amd/bin/depth_super_res_cli,
the drop-in for src/benchmark_depth_super_res_cli.cpp) on crops of the reference's own
dataset (tests/golden/depth_superres: 160x128 windows of art, books and moebius, four
upsampling levels each), checked image by image against the FP64 oracle run with the
same presets, and its printed RMSE against the one computed from the oracle's images."""
import math
import os
import re
import shutil
import subprocess

import numpy as np
import pytest
from PIL import Image

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "domain_transform_solver_amd", "bin")
CROPS = os.path.join(ROOT, "tests", "golden", "depth_superres")
SCENES = ["art_big", "books_big", "moebius_big"]
SCALE = 1.0 / (256.0 * 256.0 - 1)


def _load(top, input_name, s, k):
    rgba = np.array(Image.open(f"{top}/_input/{s}/{s}_color.png").convert("RGB"))
    rgba = np.concatenate([rgba, np.zeros(rgba.shape[:2] + (1,), np.uint8)], axis=2)
    tgt = np.array(Image.open(f"{top}/{input_name}/{s}/output_{k}_n.png")).astype(np.float32)
    conf = np.array(Image.open(f"{top}/weights/{s}/weight_{k}_n.png")).astype(np.float32)
    gt = np.array(Image.open(f"{top}/_gth/{s}.png")).astype(np.float64)
    tgt = tgt * np.float32(SCALE)
    conf = conf * np.float32(SCALE)
    conf = conf / conf.max()
    return np.ascontiguousarray(rgba), tgt, conf, gt


def _to_u8(v):
    return np.clip(np.rint(np.asarray(v, np.float64) * 255.0), 0, 255)


def _oracle_dts(oracle_mod, rgba, tgt, conf, up):
    h, w = tgt.shape
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.YCBCR, rgba, tgt, conf)
    o.compute_differential(8.0 * up, 8.0 * up, 0.1)
    o.integrate()
    o.optimize(10, oracle_mod.L2, 0.99, float(np.float32(8 * up / (1 + 0.2 * 4))), 0.99)
    return o.result()


def _oracle_hfbs(oracle_mod, rgba, tgt, conf, up):
    g8 = oracle_mod.convert_color(rgba, oracle_mod.YCBCR)
    out, _, fell = oracle_mod.hfbs_optimize(g8, (tgt - np.float32(0.5)) * np.float32(128), conf,
                                            oracle_mod.hfbs_params(), (1536, 1536, 64))
    assert not fell
    return out / 128.0 + 0.5


def _run_cli(tmp_path, algo):
    exe = os.path.join(BIN, "depth_super_res_cli")
    assert os.path.exists(exe), "depth_super_res_cli is not built (run build())"
    top = str(tmp_path / "data")
    shutil.copytree(CROPS, top)
    for s in SCENES:
        os.makedirs(f"{top}/{algo}/{s}")
    print(f"running {exe} --algorithm_name={algo}", flush=True)
    r = subprocess.run([exe, f"--path_to_data_top_dir={top}", f"--algorithm_name={algo}",
                        "--algorithm_input_name=bicubic"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    assert r.stdout.strip() == f"Using {algo} algorithm."
    m = re.search(r"RMSE ([0-9.eE+-]+), time ([0-9.eE+-]+) ms, num_pixels ([0-9.eE+-]+)",
                  r.stderr)
    assert m, r.stderr
    return top, float(m.group(1)), float(m.group(2)), float(m.group(3))


@pytest.mark.parametrize("algo", ["dts", "hfbs"])
def test_depth_super_res_cli(tmp_path, oracle_mod, gpu_lib, algo):
    top, rmse, ms, npix = _run_cli(tmp_path, algo)
    assert npix == 160 * 128 and ms > 0
    logs, logs_bp, worst = [], [], 0
    for s in SCENES:
        for k in range(1, 5):
            rgba, tgt, conf, gt = _load(top, "bicubic", s, k)
            run = _oracle_dts if algo == "dts" else _oracle_hfbs
            ref = _to_u8(run(oracle_mod, rgba, tgt, conf, 1 << k))
            got = np.array(Image.open(f"{top}/{algo}/{s}/output_{k}_n.png")).astype(np.float64)
            assert got.shape == ref.shape
            # FP32 against FP64 before rounding to 8 bits: at most one level apart.
            d = np.abs(got - ref)
            worst = max(worst, d.max())
            assert d.max() <= 1, (s, k, d.max())
            assert np.mean(d > 0) < 0.01, (s, k)
            logs.append(math.log(math.sqrt(np.mean((gt - ref) ** 2))))
            bp = np.array(Image.open(f"{top}/barronpoole/{s}/output_{k}_n.png")).astype(np.float64)
            logs_bp.append(math.log(math.sqrt(np.mean((gt - bp / 257.0) ** 2))))
    rmse_ref = math.exp(sum(logs) / len(logs))
    print(f"{algo}: RMSE printed {rmse:.4f}, from the oracle {rmse_ref:.4f}, "
          f"stored Barron-Poole results {math.exp(sum(logs_bp) / len(logs_bp)):.4f}; "
          f"worst level difference {worst:.0f}")
    assert abs(rmse - rmse_ref) <= 0.01 * rmse_ref


# ---- colorization (src/benchmark_colorization_cli.cpp) ----------------------------------

def _ycbcr_f32(bgr):
    """The program's float YCbCr of an 8-bit BGR image (same operation order)."""
    b, g, r = (bgr[..., i].astype(np.float32) for i in range(3))
    f = np.float32
    y = f(0.299) * r + f(0.587) * g + f(0.114) * b
    cb = f(128) - f(0.168736) * r - f(0.331264) * g + f(0.5) * b
    cr = f(128) + f(0.5) * r - f(0.418688) * g - f(0.081312) * b
    return y, cb, cr


def _colorization_images(seed, w, h):
    """Ground truth of flat-coloured regions, its gray image and a stroke image
    that carries the true colours on a few short horizontal strokes."""
    from domain_transform_solver_amd import scenes
    rng = np.random.default_rng(seed)
    sites = np.stack([rng.uniform(0, w, 9), rng.uniform(0, h, 9)], axis=1)
    lab = scenes.voronoi_labels(w, h, sites)
    cols = rng.integers(20, 236, (9, 3))
    gt = cols[lab].astype(np.uint8)                                   # BGR
    y = np.clip(np.rint(_ycbcr_f32(gt)[0]), 0, 255).astype(np.uint8)
    gray = np.repeat(y[..., None], 3, axis=2)
    strokes = gray.copy()
    for _ in range(24):
        yy, xx = int(rng.integers(0, h)), int(rng.integers(0, w - 12))
        strokes[yy, xx:xx + 12] = gt[yy, xx:xx + 12]
    return gt, gray, strokes


def _bmp(path, bgr):
    Image.fromarray(np.ascontiguousarray(bgr[..., ::-1])).save(path)


def _read_bgr(path):
    return np.array(Image.open(path).convert("RGB"))[..., ::-1]


def _ssim_scipy(a, b):
    """Mean SSIM per channel with scipy's Gaussian (sigma 1.5, radius 5,
    mirrored borders), the reference's formula."""
    from scipy.ndimage import gaussian_filter

    def blur(x):
        return gaussian_filter(x, 1.5, mode="mirror", truncate=5 / 1.5)
    c1, c2 = (0.01 * 255) ** 2, (0.03 * 255) ** 2
    out = []
    for c in range(3):
        x, y = a[..., c].astype(np.float64), b[..., c].astype(np.float64)
        m1, m2 = blur(x), blur(y)
        v1, v2, v12 = blur(x * x) - m1 * m1, blur(y * y) - m2 * m2, blur(x * y) - m1 * m2
        out.append(np.mean((2 * m1 * m2 + c1) * (2 * v12 + c2) /
                           ((m1 * m1 + m2 * m2 + c1) * (v1 + v2 + c2))))
    return out


def _ycbcr8(bgr):
    return np.stack([np.clip(np.trunc(v), 0, 255) for v in _ycbcr_f32(bgr)], axis=2)


@pytest.mark.parametrize("algo", ["dts", "hfbs"])
def test_colorization_cli(tmp_path, oracle_mod, gpu_lib, algo):
    exe = os.path.join(BIN, "colorization_cli")
    assert os.path.exists(exe), "colorization_cli is not built (run build())"
    top = str(tmp_path)
    names, shapes = ["IMG_3743", "IMG_3748"], [(192, 144), (131, 97)]
    data = {}
    for i, (n, (w, h)) in enumerate(zip(names, shapes)):
        gt, gray, strokes = _colorization_images(500 + i, w, h)
        _bmp(f"{top}/{n}.bmp", gray)
        _bmp(f"{top}/{n}_strokes.bmp", strokes)
        _bmp(f"{top}/{n}_gt.bmp", gt)
        data[n] = (gt, gray, strokes)
    print(f"running {exe} --algorithm_name={algo}", flush=True)
    r = subprocess.run([exe, f"--path_to_data_top_dir={top}", f"--algorithm_name={algo}"],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = r.stdout.strip().splitlines()
    assert lines[0] == f"Using {algo} algorithm." and len(lines) == 3
    m = re.search(r"RMSE 0, time ([0-9.eE+-]+) ms, num_pixels ([0-9.eE+-]+)", r.stderr)
    assert m and float(m.group(2)) == (192 * 144 + 131 * 97) / 2
    for n, line in zip(names, lines[1:]):
        gt, gray, strokes = data[n]
        h, w = gray.shape[:2]
        yyya = np.zeros((h, w, 4), np.uint8)
        yyya[..., :3] = gray[..., :1]
        _, cb_t, cr_t = _ycbcr_f32(strokes)
        cb8 = np.clip(np.rint(cb_t), 0, 255)
        cr8 = np.clip(np.rint(cr_t), 0, 255)
        conf = ((cb8 > 129) | (cb8 < 127) | (cr8 > 129) | (cr8 < 127)).astype(np.float32)
        assert conf.sum() > 50
        est = []
        for t in (cb_t, cr_t):
            if algo == "dts":
                o = oracle_mod.OracleDTS(w, h)
                o.init_frame(oracle_mod.RGB, yyya, t, conf)
                o.compute_differential(64.0, 64.0, 0.25)
                o.integrate()
                o.optimize(10, oracle_mod.CHARBONNIER, 0.99, 30.0, 0.99, 0.85)
                est.append(o.result())
            else:
                out, _, fell = oracle_mod.hfbs_optimize(
                    yyya, t - np.float32(128), (conf * 110000000164.0).astype(np.float32),
                    oracle_mod.hfbs_params(), (410, 410, 64))
                assert not fell
                est.append(out + 128)
        yf = gray[..., 0].astype(np.float64)
        rr = yf + 1.402 * (est[1] - 128)
        gg = yf - 0.344136 * (est[0] - 128) - 0.714136 * (est[1] - 128)
        bb = yf + 1.772 * (est[0] - 128)
        ref = np.clip(np.round(np.stack([bb, gg, rr], axis=2)), 0, 255)
        got = _read_bgr(f"{top}/{algo}_{n}.bmp").astype(np.float64)
        d = np.abs(got - ref)
        print(f"{n}: max level difference {d.max():.0f}, "
              f"{np.mean(d > 0):.5f} of samples differ")
        assert d.max() <= 1 and np.mean(d > 0) < 0.01
        # Printed SSIM / PSNR of this output against scipy's Gaussian.
        res8, gt8 = _ycbcr8(got.astype(np.uint8)), _ycbcr8(gt)
        ssim = _ssim_scipy(gt8, res8)
        rm = [np.sqrt(np.mean((gt8[..., c] - res8[..., c]) ** 2)) for c in (1, 2)]
        psnr = [20 * np.log10(255) - 10 * np.log10(v) for v in rm]
        mm = re.fullmatch(n + r", SSIM: \[(.*)\], PSNR: \[(.*)\]", line)
        assert mm, line
        s_pr = [float(v) for v in mm.group(1).split(",")]
        p_pr = [float(v) for v in mm.group(2).split(",")]
        np.testing.assert_allclose(s_pr[:3], ssim, rtol=1e-4, atol=1e-5)
        np.testing.assert_allclose(p_pr[1:3], psnr, rtol=1e-5)
        assert s_pr[3] == 0 and p_pr[0] == 0 and p_pr[3] == 0


# ---- stereo (src/benchmark_stereo_cli.cpp) on the crop of the ArtL pair ---------------

def test_stereo_cli_art(tmp_path, oracle_mod, gpu_lib):
    exe = os.path.join(BIN, "stereo_cli")
    assert os.path.exists(exe), "stereo_cli is not built (run build())"
    top = str(tmp_path / "stereo")
    shutil.copytree(os.path.join(ROOT, "tests", "golden", "stereo"), top)
    print(f"running {exe}", flush=True)
    r = subprocess.run([exe, f"--path_to_data_top_dir={top}"], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.strip() == "Using DTS algorithm."
    d = f"{top}/Art/"

    def rgba(p):
        a = np.array(Image.open(p).convert("RGB"))
        return np.ascontiguousarray(np.concatenate([a, np.zeros(a.shape[:2] + (1,), np.uint8)], 2))
    left, right = rgba(d + "im0.png"), rgba(d + "im1.png")
    h, w = left.shape[:2]
    t = np.fromfile(d + "im0.png_disp.bin", dtype="<f4").reshape(h, w)
    with open(d + "disp0GT.pfm", "rb") as f:
        for _ in range(3):
            f.readline()
        gt = np.frombuffer(f.read(), "<f4").reshape(h, w)[::-1].astype(np.float64)
    gt = np.where(np.isinf(gt), 0.0, gt)
    o = oracle_mod.OracleDTS(w, h)
    o.init_frame(oracle_mod.RGB, left, t, None)
    o.init_right_frame(100.0, oracle_mod.RGB, right)      # calib.txt: doffs=100
    o.compute_confidence(8.0, 8.0, 0.2, 16.0, 20)
    o.compute_differential(8.0, 8.0, 0.1)
    o.integrate()
    o.process_right_frame()
    o.optimize_stereo(1, oracle_mod.CHARBONNIER, 0.99, 30.0, 0.99, 0.01)
    ref = o.result()
    rmse_ref = float(np.sqrt(np.mean((gt - ref) ** 2)))
    m = re.search(r"RMSE ([0-9.eE+-]+), time ([0-9.eE+-]+) ms, num_pixels ([0-9.eE+-]+)",
                  r.stderr)
    assert m and float(m.group(3)) == w * h
    print(f"RMSE printed {float(m.group(1)):.5f}, from the oracle {rmse_ref:.5f}, "
          f"of the initialisation {np.sqrt(np.mean((gt - t) ** 2)):.5f}")
    assert abs(float(m.group(1)) - rmse_ref) <= 1e-3 * rmse_ref
    mm = re.search(r"disparity_estimate min:([0-9.eE+-]+) max:([0-9.eE+-]+)", r.stderr)
    assert mm and abs(float(mm.group(2)) - ref.max()) <= 1e-3 * ref.max()
    mg = re.search(r"image_data.disparity_gt min:([0-9.eE+-]+) max:([0-9.eE+-]+)", r.stderr)
    assert mg and float(mg.group(1)) == gt.min() and abs(float(mg.group(2)) - gt.max()) < 1e-3
    got = np.array(Image.open(f"{top}/dts_Art.png")).astype(np.float64)
    diff = np.abs(got - np.clip(np.rint(ref), 0, 255))
    # The photometric term gathers at int(doff + x - v): where fp32 and FP64
    # straddle an integer the two read different pixels.
    print(f"{np.mean(diff > 0):.5f} of pixels differ, max {diff.max():.0f}")
    assert np.mean(diff > 1) < 0.005 and np.mean(diff > 0) < 0.02
