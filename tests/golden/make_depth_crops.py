"""Crops of the reference's depth super-resolution test data
(tests/testdata/depth_superres in akashbapat/domain_transform_solver) for the
command-line program's real-data test: for each scene (art_big, books_big,
moebius_big) a 160x128 window of the colour image, the ground truth, and the
bicubic inputs and confidence weights of the four upsampling levels, in the
dataset's directory layout and PNG formats (8-bit RGB / gray, 16-bit gray).

    python tests/golden/make_depth_crops.py /root/reference/tests/testdata/depth_superres
"""
import os
import sys

from PIL import Image

X0, Y0, W, H = 608, 480, 160, 128
SCENES = ["art_big", "books_big", "moebius_big"]


def crop(src, dst):
    os.makedirs(os.path.dirname(dst), exist_ok=True)
    im = Image.open(src)
    im.crop((X0, Y0, X0 + W, Y0 + H)).save(dst)


def main(top):
    here = os.path.join(os.path.dirname(os.path.abspath(__file__)), "depth_superres")
    for s in SCENES:
        crop(f"{top}/_input/{s}/{s}_color.png", f"{here}/_input/{s}/{s}_color.png")
        crop(f"{top}/_gth/{s}.png", f"{here}/_gth/{s}.png")
        for k in range(1, 5):
            crop(f"{top}/bicubic/{s}/output_{k}_n.png", f"{here}/bicubic/{s}/output_{k}_n.png")
            crop(f"{top}/weights/{s}/weight_{k}_n.png", f"{here}/weights/{s}/weight_{k}_n.png")
            crop(f"{top}/barronpoole/{s}/output_{k}_n.png",
                 f"{here}/barronpoole/{s}/output_{k}_n.png")


if __name__ == "__main__":
    main(sys.argv[1])
