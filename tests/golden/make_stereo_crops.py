"""A crop of the reference's stereo test data (tests/testdata/stereo/Art in
akashbapat/domain_transform_solver, the Middlebury ArtL pair) for the stereo
program's real-data test: a 256x192 window of both views, of the float32
disparity initialisation (im0.png_disp.bin, rows top first) and of the ground
truth (disp0GT.pfm, rows bottom first, little-endian), plus calib.txt, in the
dataset's layout and formats.  Both views are cut at the same columns, so
the stereo gather at doff + x - v keeps its meaning inside the crop.

    python tests/golden/make_stereo_crops.py /root/reference/tests/testdata/stereo
"""
import os
import shutil
import sys

import numpy as np
from PIL import Image

X0, Y0, W, H = 200, 160, 256, 192


def read_pfm(path):
    with open(path, "rb") as f:
        assert f.readline().strip() == b"Pf"
        w, h = map(int, f.readline().split())
        scale = float(f.readline())
        data = np.frombuffer(f.read(4 * w * h), dtype="<f4" if scale < 0 else ">f4")
    return data.reshape(h, w)[::-1]            # top row first


def main(top):
    src = os.path.join(top, "Art")
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "stereo", "Art")
    os.makedirs(dst, exist_ok=True)
    for name in ("im0.png", "im1.png"):
        im = Image.open(os.path.join(src, name))
        full_w, full_h = im.size
        im.crop((X0, Y0, X0 + W, Y0 + H)).save(os.path.join(dst, name))
    init = np.fromfile(os.path.join(src, "im0.png_disp.bin"), dtype="<f4",
                       count=full_w * full_h).reshape(full_h, full_w)
    init[Y0:Y0 + H, X0:X0 + W].astype("<f4").tofile(os.path.join(dst, "im0.png_disp.bin"))
    gt = read_pfm(os.path.join(src, "disp0GT.pfm"))[Y0:Y0 + H, X0:X0 + W]
    with open(os.path.join(dst, "disp0GT.pfm"), "wb") as f:
        f.write(f"Pf\n{W} {H}\n-1\n".encode())
        f.write(np.ascontiguousarray(gt[::-1]).astype("<f4").tobytes())
    shutil.copyfile(os.path.join(src, "calib.txt"), os.path.join(dst, "calib.txt"))


if __name__ == "__main__":
    main(sys.argv[1])
