"""Writes tests/golden/jpeg_crops.npz: crops of the reference's own test pair
(IWPP MP/Imagens_de_Teste/50-percent-{marker,mask}.jpg, the only complete pair
shipped with the reference) and their reconstructions, computed here by a
brute-force fixed-point iteration (scipy grey_dilation), independently of the
oracle.

The reference reads these files with cv::imread(..., IMREAD_GRAYSCALE)
(IWPP MP/src/main.cu:26-27), clips marker to mask (main.cu:55) and
reconstructs with 4-connectivity.  OpenCV is not installed here, so the JPEGs
are decoded with PIL (mode "L"); the SHA-256 of each full decode is stored so
that a different decoder shows up.  Run from the repository root:
    python tests/golden/make_golden.py
"""
import hashlib
import os
import sys

import numpy as np
from PIL import Image
from scipy import ndimage

REF = "/root/reference/IWPP MP/Imagens_de_Teste"
HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "jpeg_crops.npz")
# A 2048 x 2048 window of the decoded pair (the full 4096^2 pair compresses to
# 8.5 MB; this window to about 2 MB) and the SHA-256 of its reconstruction by
# the oracle, for the regression over every queue variant and tile size (the
# counterpart of IWPP MP/test_script.sh, without the external verifier).
PAIR = os.path.join(HERE, "pair50.npz")
PAIR_WINDOW = (1024, 1024, 2048, 2048)  # row, col, height, width

# (row, col, height, width) crops, one of them ragged.
CROPS = [(0, 0, 160, 160), (1024, 2048, 160, 160), (3936, 3936, 160, 160),
         (2000, 517, 97, 131), (700, 3100, 128, 192)]


def brute_force(marker, mask, conn):
    fp = np.ones((3, 3), bool) if conn == 8 else ndimage.generate_binary_structure(2, 1)
    J = np.minimum(marker, mask)
    while True:
        N = np.minimum(ndimage.grey_dilation(J, footprint=fp, mode="constant", cval=0), mask)
        if np.array_equal(N, J):
            return J
        J = N


def main():
    marker = np.asarray(Image.open(os.path.join(REF, "50-percent-marker.jpg")).convert("L"))
    mask = np.asarray(Image.open(os.path.join(REF, "50-percent-mask.jpg")).convert("L"))
    out = {
        "marker_sha256": hashlib.sha256(marker.tobytes()).hexdigest(),
        "mask_sha256": hashlib.sha256(mask.tobytes()).hexdigest(),
        "shape": np.array(marker.shape),
        "frac_marker_above_mask": np.array((marker > mask).mean()),
    }
    for k, (r, c, h, w) in enumerate(CROPS):
        jm = np.ascontiguousarray(marker[r:r + h, c:c + w])
        im = np.ascontiguousarray(mask[r:r + h, c:c + w])
        out[f"crop{k}_marker"] = jm
        out[f"crop{k}_mask"] = im
        for conn in (4, 8):
            out[f"crop{k}_out{conn}"] = brute_force(jm, im, conn)
    np.savez_compressed(OUT, **out)
    print("wrote", OUT, {k: v.shape for k, v in out.items() if hasattr(v, "shape")})

    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    import oracle
    r, c, h, w = PAIR_WINDOW
    pm = np.ascontiguousarray(marker[r:r + h, c:c + w])
    pk = np.ascontiguousarray(mask[r:r + h, c:c + w])
    pair = {"marker": pm, "mask": pk, "window": np.array(PAIR_WINDOW)}
    for conn in (4, 8):
        res = oracle.reconstruct(pm, pk, conn)
        pair[f"out{conn}_sha256"] = np.array(hashlib.sha256(res.tobytes()).hexdigest())
    np.savez_compressed(PAIR, **pair)
    print("wrote", PAIR, {k: str(v) if v.ndim == 0 else v.shape for k, v in pair.items()})


if __name__ == "__main__":
    sys.exit(main())
