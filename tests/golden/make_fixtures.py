"""Writes the committed golden fixtures from the reference's own test data.

Runs only where the reference checkout is present (it reads
/root/reference/tests); the fixtures it writes are plain data (image pixels
and the reference's gold result tables) and travel with the repository.

Sources (DICe, NTESS, BSD-3 licensed test data; see NOTICE in this folder):
  tests/component/images/*.tif   uncompressed 8-bit TIFF, decoded with PIL
  tests/component/images/*.rawi  12-byte header (uint32 w, h, bytes) + row-major data
  tests/regression/parallel_subsets/gold/DICe_solution_0.txt
  tests/regression/single_speckled_shift/gold/*.txt
"""
import os
import shutil
import sys

import numpy as np

REF = "/root/reference/tests"
HERE = os.path.dirname(os.path.abspath(__file__))


def read_tif(name):
    from PIL import Image

    im = Image.open(os.path.join(REF, "component", "images", name))
    assert im.mode == "L", (name, im.mode)
    return np.asarray(im, dtype=np.uint8).copy()


def read_rawi(name):
    """DICe_Rawi.cpp:73-100: uint32 w, uint32 h, uint32 bytes-per-value, row-major values."""
    raw = open(os.path.join(REF, "component", "images", name), "rb").read()
    w, h, nb = np.frombuffer(raw[:12], dtype="<u4")
    dt = {4: "<f4", 8: "<f8"}[int(nb)]
    return np.frombuffer(raw[12:], dtype=dt, count=int(w) * int(h)).reshape(int(h), int(w)).copy()


def main():
    if not os.path.isdir(REF):
        print("reference tests not present; nothing to do")
        return 1
    np.savez_compressed(os.path.join(HERE, "speckle_512.npz"),
                        ref=read_tif("refSpeckled.tif"), defm=read_tif("defSpeckled.tif"))
    np.savez_compressed(os.path.join(HERE, "subset_construction.npz"),
                        ref=read_tif("TestSubsetConstructionRef.tif"),
                        defm=read_tif("TestSubsetConstructionNormalRotDisp.tif"))
    np.savez_compressed(os.path.join(HERE, "synthetic_speckled.npz"),
                        **{"def%d" % k: read_tif("defSyntheticSpeckled%d.tif" % k) for k in range(6)})
    np.savez_compressed(os.path.join(HERE, "baboon.npz"),
                        baboon_f=read_rawi("baboon.rawi"), baboon_trans_f=read_rawi("baboon_trans.rawi"),
                        baboon_d=read_rawi("baboon_d.rawi"), baboon_trans_d=read_rawi("baboon_trans_d.rawi"))
    np.savez_compressed(os.path.join(HERE, "conformal_raster.npz"),
                        conformal_f=read_rawi("conformal.rawi"), conformal_d=read_rawi("conformal_d.rawi"))
    shutil.copy(os.path.join(REF, "regression", "parallel_subsets", "gold", "DICe_solution_0.txt"),
                os.path.join(HERE, "parallel_subsets_gold.txt"))
    gold_dir = os.path.join(REF, "regression", "single_speckled_shift", "gold")
    with open(os.path.join(HERE, "single_speckled_shift_gold.txt"), "w") as out:
        for k in range(7):
            lines = open(os.path.join(gold_dir, "single_speckled_shift_%d.txt" % k)).read().splitlines()
            rows = [ln for ln in lines if ln and not ln.startswith("***")]
            if k == 0:
                out.write(rows[0] + "\n")
            out.write(rows[1] + "\n")
    return 0


if __name__ == "__main__":
    sys.exit(main())
