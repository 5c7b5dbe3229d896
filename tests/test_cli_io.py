"""The command-line programs' image codec (domain_transform_solver_amd/cli/image_io.cpp)
against PIL: PNG decoding of every colour type and bit depth the reference's data uses
(8-bit RGB colour, 16-bit gray depth and weights, 8-bit gray ground truth) and a few it
does not (palette, gray + alpha, RGBA), PNG writing read back by PIL, and PFM rows stored
bottom first (src/pfm_image.cpp:19-72)."""
import glob
import os
import subprocess

import numpy as np
import pytest
from PIL import Image

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cxx", "png_dump.exe")
CROPS = os.path.join(HERE, "golden", "depth_superres")


@pytest.fixture(scope="module", autouse=True)
def harness():
    """Builds the codec harness (g++ only, no GPU) so these tests never skip."""
    from domain_transform_solver_amd import build
    assert build.build_codec_harness() == EXE
    return EXE


def dump(path, tmp_path):
    raw = str(tmp_path / "out.raw")
    r = subprocess.run([EXE, "read", path, raw], capture_output=True, text=True, check=True)
    w, h, c, d = map(int, r.stdout.split())
    s = np.fromfile(raw, dtype=np.uint16).reshape(h, w, c)
    return s, d


def pil_array(path):
    im = Image.open(path)
    a = np.array(im)
    if im.mode == "P":
        a = np.array(im.convert("RGBA" if "transparency" in im.info else "RGB"))
    if a.ndim == 2:
        a = a[:, :, None]
    return a.astype(np.uint16)


def test_reads_dataset_crops(tmp_path):
    files = sorted(glob.glob(os.path.join(CROPS, "**", "*.png"), recursive=True))
    assert len(files) == 42
    for f in files:
        s, d = dump(f, tmp_path)
        ref = pil_array(f)
        assert s.shape == ref.shape, f
        np.testing.assert_array_equal(s, ref, err_msg=f)
        assert d == (16 if Image.open(f).mode.startswith("I") else 8)


@pytest.mark.parametrize("mode", ["L", "LA", "RGB", "RGBA", "P", "P4", "I;16", "1"])
def test_reads_colour_types(tmp_path, mode):
    rng = np.random.default_rng(7)
    h, w = 37, 53                                   # odd sizes: ragged rows
    if mode == "I;16":
        im = Image.fromarray(rng.integers(0, 65536, (h, w), dtype=np.uint16))
    elif mode == "1":
        im = Image.fromarray(rng.integers(0, 2, (h, w), dtype=np.uint8) * 255).convert("1")
    elif mode == "P":
        im = Image.fromarray(rng.integers(0, 256, (h, w, 3), dtype=np.uint8)).quantize(97)
    elif mode == "P4":                              # 16 colours: 4-bit packed indices
        im = Image.fromarray(rng.integers(0, 256, (h, w, 3), dtype=np.uint8)).quantize(16)
    else:
        c = {"L": 1, "LA": 2, "RGB": 3, "RGBA": 4}[mode]
        a = rng.integers(0, 256, (h, w, c), dtype=np.uint8)
        # smooth ramps too, so that PIL's filter choice varies between rows
        a[: h // 2] = (np.arange(w)[None, :, None] * 3 + np.arange(c)).astype(np.uint8)
        im = Image.fromarray(a[:, :, 0] if c == 1 else a, mode)
    p = str(tmp_path / "in.png")
    im.save(p)
    s, _ = dump(p, tmp_path)
    ref = pil_array(p)
    if mode == "1":
        ref = ref // 255 * 255 if ref.max() > 1 else ref * 255
        s = s * 255 if s.max() <= 1 else s
    np.testing.assert_array_equal(s, ref)


@pytest.mark.parametrize("bits", [8, 16])
def test_writes_gray(tmp_path, bits):
    rng = np.random.default_rng(bits)
    h, w = 29, 41
    v = rng.integers(0, 1 << bits, (h, w), dtype=np.uint16)
    raw, png = str(tmp_path / "in.raw"), str(tmp_path / "o.png")
    v.tofile(raw)
    subprocess.run([EXE, f"write{bits}", str(w), str(h), raw, png], check=True)
    got = np.array(Image.open(png)).astype(np.uint16)
    np.testing.assert_array_equal(got, v)


@pytest.mark.parametrize("scale", ["-1.0", "1.0"])
def test_reads_pfm_bottom_first(tmp_path, scale):
    rng = np.random.default_rng(3)
    h, w = 9, 14
    top_first = rng.standard_normal((h, w)).astype(np.float32)
    p = str(tmp_path / "d.pfm")
    dt = "<f4" if scale.startswith("-") else ">f4"
    with open(p, "wb") as f:
        f.write(f"Pf\n{w} {h}\n{scale}\n".encode())
        f.write(top_first[::-1].astype(dt).tobytes())
    raw = str(tmp_path / "o.raw")
    r = subprocess.run([EXE, "pfm", p, raw], capture_output=True, text=True, check=True)
    assert list(map(int, r.stdout.split())) == [w, h]
    np.testing.assert_array_equal(np.fromfile(raw, dtype=np.float32).reshape(h, w), top_first)


def _bgr(path, tmp_path):
    raw = str(tmp_path / "bgr.raw")
    r = subprocess.run([EXE, "bgr", path, raw], capture_output=True, text=True, check=True)
    w, h = map(int, r.stdout.split())
    return np.fromfile(raw, dtype=np.uint8).reshape(h, w, 3)


@pytest.mark.parametrize("mode", ["RGB", "RGBA", "L", "P", "P4", "1"])
def test_reads_bmp(tmp_path, mode):
    """BMP as cv::imread(IMREAD_COLOR) returns it: 8-bit BGR, top row first;
    odd widths exercise the 4-byte row padding."""
    rng = np.random.default_rng(11)
    h, w = 23, 37
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    im = Image.fromarray(a)
    if mode in ("P", "P4"):
        im = im.quantize(200 if mode == "P" else 16)
    elif mode != "RGB":
        im = im.convert(mode)
    p = str(tmp_path / "in.bmp")
    im.save(p)
    ref = np.array(im.convert("RGB"))[..., ::-1]
    np.testing.assert_array_equal(_bgr(p, tmp_path), ref)


def test_reads_top_down_bmp(tmp_path):
    rng = np.random.default_rng(12)
    h, w = 5, 6
    a = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    stride = (w * 3 + 3) // 4 * 4
    rows = b"".join(a[y, :, ::-1].tobytes() + b"\0" * (stride - 3 * w) for y in range(h))
    hdr = (b"BM" + (54 + len(rows)).to_bytes(4, "little") + b"\0\0\0\0" +
           (54).to_bytes(4, "little") + (40).to_bytes(4, "little") +
           w.to_bytes(4, "little") + (-h).to_bytes(4, "little", signed=True) +
           (1).to_bytes(2, "little") + (24).to_bytes(2, "little") + b"\0" * 24)
    p = str(tmp_path / "td.bmp")
    with open(p, "wb") as f:
        f.write(hdr + rows)
    np.testing.assert_array_equal(_bgr(p, tmp_path), a[..., ::-1])


def test_writes_bmp_and_reads_png_as_bgr(tmp_path):
    rng = np.random.default_rng(13)
    h, w = 19, 31
    bgr = rng.integers(0, 256, (h, w, 3), dtype=np.uint8)
    raw, p = str(tmp_path / "in.raw"), str(tmp_path / "o.bmp")
    bgr.tofile(raw)
    subprocess.run([EXE, "wbmp", str(w), str(h), raw, p], check=True)
    np.testing.assert_array_equal(np.array(Image.open(p))[..., ::-1], bgr)
    png = str(tmp_path / "o.png")
    Image.fromarray(bgr[..., ::-1].copy()).save(png)
    np.testing.assert_array_equal(_bgr(png, tmp_path), bgr)
    g16 = str(tmp_path / "g16.png")
    v = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    Image.fromarray(v).save(g16)
    np.testing.assert_array_equal(_bgr(g16, tmp_path), np.repeat((v >> 8)[..., None], 3, 2))


def _bmp_bytes(w, h, bpp, rows, palette=b"", comp=0, masks=b""):
    off = 54 + len(masks) + len(palette)
    ncol = len(palette) // 4
    return (b"BM" + (off + len(rows)).to_bytes(4, "little") + b"\0\0\0\0" +
            off.to_bytes(4, "little") + (40).to_bytes(4, "little") + w.to_bytes(4, "little") +
            h.to_bytes(4, "little") + (1).to_bytes(2, "little") + bpp.to_bytes(2, "little") +
            comp.to_bytes(4, "little") + b"\0" * 12 + ncol.to_bytes(4, "little") + b"\0" * 4 +
            masks + palette + rows)


def test_reads_4bit_and_16bit_bmp(tmp_path):
    rng = np.random.default_rng(14)
    h, w = 4, 7
    pal = rng.integers(0, 256, (16, 3), dtype=np.uint8)
    idx = rng.integers(0, 16, (h, w))
    stride = ((w * 4 + 31) // 32) * 4
    rows = b""
    for y in range(h - 1, -1, -1):                     # bottom row first
        r = bytearray(stride)
        for x in range(w):
            r[x // 2] |= int(idx[y, x]) << (4 if x % 2 == 0 else 0)
        rows += bytes(r)
    palette = b"".join(bytes([p[2], p[1], p[0], 0]) for p in pal)   # stored B G R 0
    p4 = str(tmp_path / "p4.bmp")
    with open(p4, "wb") as f:
        f.write(_bmp_bytes(w, h, 4, rows, palette))
    np.testing.assert_array_equal(_bgr(p4, tmp_path), pal[idx][..., ::-1])
    # 16-bit 5-6-5 with bit fields
    v = rng.integers(0, 65536, (h, w), dtype=np.uint16)
    stride = ((w * 16 + 31) // 32) * 4
    rows = b"".join(v[y].astype("<u2").tobytes() + b"\0" * (stride - 2 * w)
                    for y in range(h - 1, -1, -1))
    masks = b"".join(m.to_bytes(4, "little") for m in (0xF800, 0x07E0, 0x001F))
    p16 = str(tmp_path / "p16.bmp")
    with open(p16, "wb") as f:
        f.write(_bmp_bytes(w, h, 16, rows, comp=3, masks=masks))
    r5, g6, b5 = (v >> 11) & 31, (v >> 5) & 63, v & 31
    ref = np.stack([(b5 * 255 + 15) // 31, (g6 * 255 + 31) // 63, (r5 * 255 + 15) // 31], 2)
    np.testing.assert_array_equal(_bgr(p16, tmp_path), ref.astype(np.uint8))
