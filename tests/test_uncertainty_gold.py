"""The uncertainty fields against the reference's gold: tests/regression/bitmap_star (translation-only
Gauss-Newton, Keys-4, 31 x 31 subsets, the uncertainty post-processor on), 117 subsets of the corner cut of
tests/golden/make_uncertainty_fixture.py.  The gold's UNCERTAINTY is SIGMA / FIELD_1
(Uncertainty_Post_Processor, DICe_PostProcessor.cpp:1507-1537), so it pins FIELD_1, the cosine of
computeUncertaintyFields (DICe_Objective.cpp:193-299).

Each subset is solved by computeUpdateFast from the gold's own u, v (the deck's feature-matching initial
guess is out of scope; the solution does not depend on it), then sigma and gamma at the result.
Bar (five printed digits): u, v within 2e-5 px; sigma, gamma and UNCERTAINTY within 1e-4 relative; the
oracle in its double build, the GPU in float."""
import os

import numpy as np
import pytest

import oracle_py as o

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "uncertainty_wave.npz")


def load():
    z = np.load(FIX)
    cols = [str(c) for c in z["columns"]]
    return z["ref"], z["defm"], z["gold"], {c: i for i, c in enumerate(cols)}


def check(rows, c, solved):
    for r, (p, sigma, gamma, field1) in zip(rows, solved):
        assert abs(p[0] - r[c["DISPLACEMENT_X"]]) <= 2e-5 and abs(p[1] - r[c["DISPLACEMENT_Y"]]) <= 2e-5
        assert abs(sigma - r[c["SIGMA"]]) <= 1e-4 * r[c["SIGMA"]]
        assert abs(gamma - r[c["GAMMA"]]) <= 1e-4 * r[c["GAMMA"]]
        unc = np.float32(1.0 / np.float64(np.float32(field1)) * np.float64(np.float32(sigma)))
        assert abs(unc - r[c["UNCERTAINTY"]]) <= 1e-4 * r[c["UNCERTAINTY"]]


def test_oracle_uncertainty_fields_match_gold():
    ref, de, rows, c = load()
    p = o.default_params()
    p.optimization_method = 1  # GRADIENT_BASED
    solved = []
    for r in rows:
        x, y = r[c["COORDINATE_X"]], r[c["COORDINATE_Y"]]
        q, res = o.objective(p, ref, de, 31, x, y, 2, [r[c["DISPLACEMENT_X"]], r[c["DISPLACEMENT_Y"]], 0, 0, 0, 0])
        assert res[0] == 0
        _, rs = o.objective(p, ref, de, 31, x, y, 1, q)
        _, rg = o.objective(p, ref, de, 31, x, y, 0, q)
        solved.append((q, rs[0], rg[0], res[3]))
    check(rows, c, solved)


@pytest.mark.gpu
def test_gpu_uncertainty_fields_match_gold():
    """The same through the C ABI: dice_hip_objective for the solve, sigma and gamma, and the uncertainty
    post-processor (dice_hip_uncertainty) over the solved fields."""
    import dice_amd as dh
    ref, de, rows, c = load()
    cx, cy = rows[:, c["COORDINATE_X"]], rows[:, c["COORDINATE_Y"]]
    ctx = dh.Context(0)
    try:
        p = dh.default_params()
        p.optimization_method = dh.GRADIENT_BASED
        ctx.set_params(p)
        ctx.set_ref_image(ref.astype(np.float32))
        ctx.set_def_image(de.astype(np.float32))
        ctx.set_subsets(cx, cy, 31)
        solved = []
        f = dh.new_fields(cx, cy)
        for i, r in enumerate(rows):
            q, res = ctx.objective(i, dh.OBJ_UPDATE_FAST, [r[c["DISPLACEMENT_X"]], r[c["DISPLACEMENT_Y"]], 0, 0, 0, 0])
            assert res[0] == 0
            _, rs = ctx.objective(i, dh.OBJ_SIGMA, q)
            _, rg = ctx.objective(i, dh.OBJ_GAMMA, q)
            solved.append((q, rs[0], rg[0], res[3]))
            f[dh.F["SIGMA"], i], f[dh.F["FIELD_1"], i] = rs[0], res[3]
        check(rows, c, solved)
        unc = ctx.uncertainty(f)
        assert np.all(np.abs(unc[0] - rows[:, c["UNCERTAINTY"]]) <= 1e-4 * rows[:, c["UNCERTAINTY"]])
    finally:
        ctx.close()
