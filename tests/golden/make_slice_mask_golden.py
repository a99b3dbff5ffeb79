"""Writes the golden slice mask from the reference's committed evaluation run.

The reference tree holds the output of its own evaluation tool for the bumpy
sphere with the zebra k = 8 boundary (figures/zebra_k8_m2.json, command line
in that file: procedural_bumpy_sphere, --eval-mode slice --slice-view xy
--slice-plane 0.0 --slice-width 512 --slice-height 512) and the slice's
interior mask as a P6 image (figures/zebra_k8_m2_slice_mask.ppm: rows top to
bottom = v descending, inside 235, outside 20).  Only the mask bits are
committed, as slice_mask_zebra_k8_m2.npz beside this script:
  bits           np.packbits of the inside flags, row-major, rows as in the image
  width, height  512, 512
  inside_pixels  the count the reference's JSON reports (162215)
Run where the reference tree is present: python tests/golden/make_slice_mask_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REF = "/root/reference/figures"


def read_p6(path):
    with open(path, "rb") as f:
        data = f.read()
    parts = data.split(maxsplit=4)
    assert parts[0] == b"P6"
    w, h, maxv = int(parts[1]), int(parts[2]), int(parts[3])
    assert maxv == 255
    pix = np.frombuffer(parts[4], np.uint8)
    assert pix.size == w * h * 3
    return w, h, pix.reshape(h, w, 3)


def main():
    w, h, img = read_p6(os.path.join(REF, "zebra_k8_m2_slice_mask.ppm"))
    assert set(np.unique(img[:, :, 0]).tolist()) <= {20, 235}
    inside = img[:, :, 0] == 235
    doc = json.load(open(os.path.join(REF, "zebra_k8_m2.json")))
    n = int(doc["slice_eval"]["inside_pixels"])
    assert int(inside.sum()) == n
    np.savez_compressed(os.path.join(HERE, "slice_mask_zebra_k8_m2.npz"), bits=np.packbits(inside.ravel()),
                        width=w, height=h, inside_pixels=n)
    print("wrote slice_mask_zebra_k8_m2.npz:", w, h, n)


if __name__ == "__main__":
    main()
