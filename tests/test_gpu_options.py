"""GPU parity of the driver and solver options, through the C ABI, against the CPU oracle (double build)
on the same seeded inputs: CONVOLUTION_5_POINT gradients (smooth_gradients_convolution_5_point), the
initial and final gamma thresholds (INITIALIZE_FAILED, FRAME_FAILED_DUE_TO_HIGH_GAMMA), the jump
tolerances (JUMP_TOLERANCE_EXCEEDED), the Levenberg-Marquardt diagonal, momentum, gamma normalised by
the active pixels (generic_correlation_routine, DICe_Schema.cpp:2025-2368; computeUpdateFast,
DICe_Objective.cpp:430-657), and the uncertainty fields FIELD_1..7 (computeUncertaintyFields,
DICe_Objective.cpp:193-299).  Bar as in tests/gpu_common.py."""
import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth
from gpu_common import compare, compare_spread, run_oracle, to_oracle

pytestmark = pytest.mark.gpu

MODES = [dh.MODE_BATCHED, dh.MODE_PLAIN]
MODE_NAMES = {dh.MODE_PLAIN: "plain", dh.MODE_BATCHED: "batched"}
S = 31


@pytest.fixture(scope="module")
def scene():
    ref, de = synth.config1(seed=7)
    cx, cy = dh.regular_grid(512, 512, S, 16)
    ctx = dh.Context(0)
    yield ctx, ref, de, cx, cy
    ctx.close()


def gn6():
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    return p


def run_both(scene, p, mode, fields=None, float_too=False):
    ctx, ref, de, cx, cy = scene
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(de)
    ctx.set_subsets(cx, cy, S)
    f0 = dh.new_fields(cx, cy) if fields is None else fields
    fg = ctx.correlate(f0.copy(), mode=mode)
    fo = run_oracle(p, ref, de, cx, cy, S, fields=f0)
    if float_too:
        return fg, fo, run_oracle(p, ref, de, cx, cy, S, fields=f0, precision="float")
    return fg, fo


def truth_fields(cx, cy, off=0.0):
    f = dh.new_fields(cx, cy)
    f[dh.F["DISP_X"]] = 0.5 + 0.01 * cx + off
    f[dh.F["DISP_Y"]] = -0.3 + 0.005 * cy - off
    return f


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_convolution_5_point_gradients(scene, mode):
    p = gn6()
    p.gradient_method = 1
    fg, fo, ff = run_both(scene, p, mode, truth_fields(scene[3], scene[4], 0.3), float_too=True)
    both = compare_spread(fg, fo, ff, label="gradient_method=1")
    assert both.mean() > 0.95
    # smoothing changes the solve: the unsmoothed run differs
    q = gn6()
    fu, _ = run_both(scene, q, mode, truth_fields(scene[3], scene[4], 0.3))
    assert np.abs(fu[dh.F["FIELD_7"]] - fg[dh.F["FIELD_7"]]).max() > 1.0


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_initial_gamma_threshold(scene, mode):
    ctx, ref, de, cx, cy = scene
    p = gn6()
    f0 = truth_fields(cx, cy, 0.4)
    q = to_oracle(p)
    g0 = np.array([o.objective(q, ref.astype(np.float64), de.astype(np.float64), S, cx[i], cy[i], 0,
                               [f0[dh.F["DISP_X"], i], f0[dh.F["DISP_Y"], i], 0, 0, 0, 0])[1][0]
                   for i in range(len(cx))])
    g_sorted = np.sort(g0)
    mid = len(g0) // 2
    p.initial_gamma_threshold = float(0.5 * (g_sorted[mid - 1] + g_sorted[mid]))
    fg, fo = run_both(scene, p, mode, f0)
    compare(fg, fo, label="initial_gamma_threshold")
    st = fg[dh.F["STATUS"]]
    assert 0.3 < np.mean(st == 5) < 0.7 and np.mean(st == 1) > 0.3
    assert np.all(fg[dh.F["SIGMA"]][st == 5] == -1.0)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_final_gamma_threshold(scene, mode):
    p = gn6()
    f0 = truth_fields(scene[3], scene[4], 0.2)
    fref, _ = run_both(scene, p, mode, f0)
    g_sorted = np.sort(fref[dh.F["GAMMA"]])
    mid = len(g_sorted) // 2
    p.final_gamma_threshold = float(0.5 * (g_sorted[mid - 1] + g_sorted[mid]))
    fg, fo = run_both(scene, p, mode, f0)
    compare(fg, fo, label="final_gamma_threshold")
    st = fg[dh.F["STATUS"]]
    assert 0.3 < np.mean(st == 23) < 0.7


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("opt", [dh.GRADIENT_BASED, dh.GRADIENT_BASED_THEN_SIMPLEX], ids=["gn", "gn_then_simplex"])
def test_jump_tolerances(scene, mode, opt):
    """From a zero guess, subsets whose solution moves more than disp_jump_tol (the imposed field reaches
    5.6 px) or turns by more than theta_jump_tol fail with JUMP_TOLERANCE_EXCEEDED.  With the simplex
    fallback the jumped subsets are solved again by Nelder-Mead, whose path forks on rounding: checked
    against the float/double spread of the reference algorithm."""
    ctx, ref, de, cx, cy = scene
    p = gn6()
    p.optimization_method = opt
    p.disp_jump_tol = 3.0
    fg, fo, ff = run_both(scene, p, mode, float_too=True)
    if opt == dh.GRADIENT_BASED:
        compare(fg, fo, label="disp_jump_tol")
    else:
        compare_spread(fg, fo, ff, label="disp_jump_tol", nelder_mead=True)
    st = fg[dh.F["STATUS"]]
    assert np.mean(st == 16) > 0.2 and np.mean(st == 1) > 0.2
    p.disp_jump_tol = 100.0
    p.theta_jump_tol = 2e-4
    fg, fo, ff = run_both(scene, p, mode, truth_fields(cx, cy, 0.2), float_too=True)
    if opt == dh.GRADIENT_BASED:
        compare(fg, fo, label="theta_jump_tol")
    else:
        compare_spread(fg, fo, ff, label="theta_jump_tol", nelder_mead=True)
    st = fg[dh.F["STATUS"]]
    assert np.mean(st == 16) > 0.2 and np.mean(st == 1) > 0.2


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("option", ["lm", "momentum", "normalize"])
def test_solver_options(scene, mode, option):
    p = gn6()
    if option == "lm":
        p.levenberg_marquardt_regularization_factor = 1e5
    elif option == "momentum":
        p.momentum_factor = 0.3
    else:
        p.normalize_gamma_with_active_pixels = 1
    fg, fo, ff = run_both(scene, p, mode, truth_fields(scene[3], scene[4], 0.4), float_too=True)
    if option == "momentum":
        # momentum makes the convergence linear: where the tolerance is met moves with rounding
        both = compare_spread(fg, fo, ff, label=option)
    else:
        agree, both = compare(fg, fo, label=option)
        same = both & (fg[dh.F["ITERATIONS"]] == fo[o.F["ITERATIONS"]])
        assert same.sum() >= 0.99 * both.sum()
    assert both.mean() > 0.9
    base, _ = run_both(scene, gn6(), mode, truth_fields(scene[3], scene[4], 0.4))
    if option == "normalize":
        # gamma divided by the active pixel count (Objective::gamma, DICe_Objective.cpp:66-72)
        ok = both & np.isin(base[dh.F["STATUS"]], dh.CONVERGED_STATUSES)
        np.testing.assert_allclose(fg[dh.F["GAMMA"]][ok] * fg[dh.F["ACTIVE_PIXELS"]][ok],
                                   base[dh.F["GAMMA"]][ok], rtol=1e-4)
    else:
        # the option changes the iteration path
        assert np.any(fg[dh.F["ITERATIONS"]] != base[dh.F["ITERATIONS"]])


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_uncertainty_fields_all_seven(scene, mode):
    """FIELD_1..7 on the subsets that took the same iterations on both sides.  FIELD_2 is exactly 1 without
    an image deformer.  FIELD_6 (and FIELD_4 = FIELD_6 / FIELD_1) weight each pixel by 1 / |grad|^2 and
    skip pixels with |grad|^2 < 1e-4, so a pixel whose float gradient sits within rounding of that cut
    moves them by its whole term; the bar is checked on the subsets without such a pixel."""
    ctx, ref, de, cx, cy = scene
    p = gn6()
    fg, fo = run_both(scene, p, mode, truth_fields(cx, cy, 0.3))
    agree, both = compare(fg, fo, label="uncertainty")
    same = both & (fg[dh.F["ITERATIONS"]] == fo[o.F["ITERATIONS"]])
    assert same.mean() > 0.9
    assert np.all(fg[dh.F["FIELD_2"]][same] == 1.0) and np.all(fo[o.F["FIELD_2"]][same] == 1.0)
    gx, gy = o.fd_gradients(ref.astype(np.float64))
    g2 = gx * gx + gy * gy
    h = S // 2
    near_cut = np.array([np.any(np.abs(g2[int(y) - h:int(y) + h + 1, int(x) - h:int(x) + h + 1] - 1e-4) < 1e-6)
                         for x, y in zip(cx, cy)])
    clean = same & ~near_cut
    assert clean.sum() > 0.8 * same.sum()
    for name in ("FIELD_1", "FIELD_3", "FIELD_4", "FIELD_5", "FIELD_6", "FIELD_7"):
        a, b = fg[dh.F[name]][clean], fo[o.F[name]][clean]
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-12)
        assert rel.max() <= 1e-4, (name, rel.max(), np.median(rel))
    np.testing.assert_allclose(fg[dh.F["FIELD_4"]][clean],
                               fg[dh.F["FIELD_6"]][clean] / fg[dh.F["FIELD_1"]][clean], rtol=1e-6)
