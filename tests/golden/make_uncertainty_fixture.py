"""Fixture for the uncertainty fields: tests/regression/bitmap_star (translation-only Gauss-Newton, Keys-4,
31 x 31 subsets at step 25 over the DIC Challenge wave images, the uncertainty post-processor on), cut to
the 400 x 300 px top-left corner of its two 8-bit BMP images and the 117 gold rows whose subsets lie at
least 45 px inside the cut (x <= 330, y <= 230), so that their windows and gradient stencils never reach
it.  The deck's reference image is DIC_Challenge_Wave_Deformed_Noisy.bmp and its deformed image
DIC_Challenge_Wave_Reference_Noisy.bmp.  The gold's UNCERTAINTY is SIGMA / FIELD_1
(Uncertainty_Post_Processor, DICe_PostProcessor.cpp:1507-1537), so it pins FIELD_1 of
computeUncertaintyFields (DICe_Objective.cpp:193-299).

Run from the repository root: python tests/golden/make_uncertainty_fixture.py"""
import os

import numpy as np
from PIL import Image

SRC = "/root/reference/tests/regression/bitmap_star"
DST = os.path.join(os.path.dirname(os.path.abspath(__file__)), "uncertainty_wave.npz")


def main():
    ref = np.asarray(Image.open(os.path.join(SRC, "images", "DIC_Challenge_Wave_Deformed_Noisy.bmp")), np.uint8)
    de = np.asarray(Image.open(os.path.join(SRC, "images", "DIC_Challenge_Wave_Reference_Noisy.bmp")), np.uint8)
    lines = open(os.path.join(SRC, "gold", "DICe_solution_0.txt")).read().splitlines()
    header = [ln for ln in lines if ln.startswith("SUBSET_ID")][0].split(",")
    rows = np.array([[float(t) for t in ln.split(",")] for ln in lines if ln[:1].isdigit()])
    keep = (rows[:, 1] <= 330) & (rows[:, 2] <= 230)
    np.savez_compressed(DST, ref=ref[:300, :400], defm=de[:300, :400], gold=rows[keep],
                        columns=np.array(header))


if __name__ == "__main__":
    main()
