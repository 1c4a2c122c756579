"""Writes piv_challenge_a.npz: the PIV Challenge (2001) case A pair that the
reference's real-data test loads (test/test_reference.jl:8-26, images
test/reference_images/A/A001_1.tif and A001_2.tif, 8-bit, 1024 rows x 1280
columns). The pixels are stored unchanged as uint8; the test divides by 255,
as the reference's Float64.(TiffImages.load(...)) of an 8-bit image does.

Run from the repository root with the reference checked out at
/root/reference: python3 tests/golden/make_piv_challenge_a.py
"""
import pathlib

import numpy as np
from PIL import Image

SRC = pathlib.Path("/root/reference/test/reference_images/A")
OUT = pathlib.Path(__file__).with_name("piv_challenge_a.npz")


def main():
    frames = []
    for name in ("A001_1.tif", "A001_2.tif"):
        im = Image.open(SRC / name)
        assert im.mode == "L", im.mode
        frames.append(np.asarray(im, dtype=np.uint8))
    assert frames[0].shape == frames[1].shape
    np.savez_compressed(OUT, A=frames[0], B=frames[1])
    print(OUT, frames[0].shape, frames[0].dtype)


if __name__ == "__main__":
    main()
