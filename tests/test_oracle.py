"""Pins the CPU oracle against the reference's own fixtures and closed forms.

Every check here uses data or known answers that the reference's tests hold
(component tests and regression golds), or a mathematical fact
that a plausible restatement mistake would break.  CPU only.
"""
import math
import os

import numpy as np
import pytest

import oracle_py as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    return np.load(os.path.join(GOLD, name))


def gold_table(name):
    rows = [ln.split() for ln in open(os.path.join(GOLD, name)) if ln[:1].isdigit()]
    return np.array([[float(x) for x in r] for r in rows])


def dice_diff_ok(value, gold, tol, floor):
    """tools/DICe_Diff.cpp:208-224: gold magnitudes below `floor` are skipped and the
    remaining values compare by ABSOLUTE difference (the quirk of its -t flag)."""
    value = np.asarray(value)
    gold = np.asarray(gold)
    mask = np.abs(gold) >= floor
    return bool(np.all(np.abs(value[mask] - gold[mask]) <= tol))


@pytest.mark.parametrize("precision", ["double", "float"])
def test_parallel_subsets_gold(precision):
    """tests/regression/parallel_subsets: 49 subsets, S=27, step 75, default params."""
    d = load("speckle_512.npz")
    g = gold_table("parallel_subsets_gold.txt")
    f = o.new_fields(g[:, 1], g[:, 2])
    o.correlate(o.default_params(), d["ref"], d["defm"], f, subset_size=27, precision=precision)
    # regression CMakeLists: DICe_Diff -t 1E-2 -f 1E-4 for this case
    for col, name in [(3, "DISP_X"), (4, "DISP_Y"), (8, "GAMMA"), (9, "ITERATIONS"), (10, "STATUS")]:
        assert dice_diff_ok(f[o.F[name]], g[:, col], 1e-2, 1e-4), name
    # and much tighter than the gold's own 5-significant-digit printout needs:
    assert np.abs(f[o.F["DISP_X"]] - g[:, 3]).max() < 5e-5
    assert np.abs(f[o.F["DISP_Y"]] - g[:, 4]).max() < 5e-5
    assert np.abs(f[o.F["GAMMA"]] - g[:, 8]).max() < 1e-6
    assert np.array_equal(f[o.F["ITERATIONS"]], g[:, 9])
    assert np.array_equal(f[o.F["STATUS"]], g[:, 10])


@pytest.mark.parametrize("precision", ["double", "float"])
def test_single_speckled_shift_gold(precision):
    """tests/regression/single_speckled_shift: u, v and SIGMA of 7 subsets (DICe_Diff -t 0.05 -f 5e-5)."""
    d = load("speckle_512.npz")
    g = gold_table("single_speckled_shift_gold.txt")
    f = o.new_fields(g[:, 1], g[:, 2])
    o.correlate(o.default_params(), d["ref"], d["defm"], f, subset_size=27, precision=precision)
    assert np.abs(f[o.F["DISP_X"]] - g[:, 3]).max() < 5e-5
    assert np.abs(f[o.F["DISP_Y"]] - g[:, 4]).max() < 5e-5
    assert np.abs(f[o.F["SIGMA"]] - g[:, 9]).max() < 2e-6
    assert np.all(f[o.F["STATUS"]] == 1)


@pytest.mark.parametrize("precision", ["double", "float"])
@pytest.mark.parametrize("opt", [0, 3, 1, 4])
@pytest.mark.parametrize("interp", [0, 2])
def test_correlate_params_known_answer(precision, opt, interp):
    """DICe_TestCorrelateParams.cpp:78-140: 6-param affine known answer within 1e-2."""
    d = load("subset_construction.npz")
    p = o.default_params()
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.robust_solver_tolerance = 1e-4
    p.optimization_method = opt
    p.interpolation_method = interp
    f = o.new_fields([30.0], [30.0])
    f[o.F["DISP_X"]] = 9.8 - 0.25
    f[o.F["DISP_Y"]] = -7.62 - 0.25
    f[o.F["ROT_Z"]] = 0.262 - 0.05
    o.correlate(p, d["ref"], d["defm"], f, subset_size=21, precision=precision)
    v = f[:, 0]
    assert abs(v[o.F["DISP_X"]] - 9.8) < 1e-2
    assert abs(v[o.F["DISP_Y"]] + 7.62) < 1e-2
    assert abs(v[o.F["ROT_Z"]] - 0.262) < 1e-2
    assert abs(v[o.F["NS_XX"]] - 0.121) < 1e-2
    assert abs(v[o.F["NS_YY"]] - 0.121) < 1e-2
    assert abs(v[o.F["SS_XY"]]) < 1e-2
    assert v[o.F["STATUS"]] == 1
    assert abs(v[o.F["SIGMA"]] + 1) >= 1e-2


@pytest.mark.parametrize("kind", ["d", "f"])
def test_baboon_transform_golden(kind):
    """DICe_TestImage.cpp:236-266: Keys-4 apply_transform golden, L2 diff < 0.1."""
    d = load("baboon.npz")
    img = d["baboon_" + kind]
    out = o.apply_transform(img, 100, 100, 25.0, -15.2, 22.8 * math.pi / 180.0)
    diff = math.sqrt(float(np.sum((out.astype(np.float64) - d["baboon_trans_" + kind]) ** 2)))
    assert diff < 0.1


def test_baboon_transform_catches_planted_mistakes():
    """The golden is sharp enough to catch a transposed map or a dropped Keys term."""
    d = load("baboon.npz")
    img = d["baboon_d"]
    wrong = o.apply_transform(img, 100, 100, 25.0, -15.2, -22.8 * math.pi / 180.0)  # sign flip
    diff = math.sqrt(float(np.sum((wrong - d["baboon_trans_d"]) ** 2)))
    assert diff > 100.0


def test_stripe_gamma_closed_form():
    """DICe_TestSubsetGamma.cpp:76-117: gamma = 0.4*k (+-0.05) for a 2k-px stripe shift."""
    w = h = 199
    ref = np.zeros((h, w), np.float32)
    for stripe in range(6):
        x0 = stripe * 40
        ref[:, x0:min(x0 + 20, w)] = 255
    for shift in range(11):
        de = np.zeros((h, w), np.float32)
        for stripe in range(6):
            x0 = stripe * 40 - shift * 2
            lo, hi = max(x0, 0), min(x0 + 20, w)
            if hi > lo:
                de[:, lo:hi] = 255
        gam = o.subset_gamma(ref, de, 100, 100, 99)
        assert abs(gam - shift * 0.4) <= 5e-2


def test_fd_gradients_sin_cos():
    """DICe_TestImage.cpp:166-232: FD gradients of 255cos(x/4pi)sin(y/4pi) within 1e-3."""
    w, h = 30, 20
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    k = 1.0 / (4 * math.pi)
    img = 255.0 * np.cos(x * k) * np.sin(y * k)
    gx_exact = -255.0 * k * np.sin(x * k) * np.sin(y * k)
    gy_exact = 255.0 * k * np.cos(x * k) * np.cos(y * k)
    for dt in (np.float64, np.float32):
        gx, gy = o.fd_gradients(img.astype(dt))
        assert np.abs(gx[2:-2, 2:-2] - gx_exact[2:-2, 2:-2]).max() < 1e-3
        assert np.abs(gy[2:-2, 2:-2] - gy_exact[2:-2, 2:-2]).max() < 1e-3
    # first-order one-sided differences within 2 px of the border
    gx, gy = o.fd_gradients(img)
    assert np.allclose(gx[:, 0], img[:, 1] - img[:, 0])
    assert np.allclose(gx[:, -1], img[:, -1] - img[:, -2])
    assert np.allclose(gy[0, :], img[1, :] - img[0, :])


def test_interpolants_reproduce_polynomials():
    """Bilinear is exact on a + bx + cy + dxy; Keys 4th order is exact on cubics (interior)."""
    rng = np.random.default_rng(3)
    h, w = 30, 40
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    bil = 3.0 + 0.5 * x - 0.25 * y + 0.1 * x * y
    px = rng.uniform(0, w - 1.6, 200)
    py = rng.uniform(0, h - 1.6, 200)
    exact = 3.0 + 0.5 * px - 0.25 * py + 0.1 * px * py
    assert np.abs(o.interp(0, bil, px, py) - exact).max() < 1e-9
    cub = 1.0 + 0.3 * x + 0.02 * x ** 2 - 0.001 * x ** 3 + 0.5 * y - 0.01 * y ** 2 + 0.0005 * y ** 3
    px = rng.uniform(3.0, w - 4.0, 200)
    py = rng.uniform(3.0, h - 4.0, 200)
    ce = 1.0 + 0.3 * px + 0.02 * px ** 2 - 0.001 * px ** 3 + 0.5 * py - 0.01 * py ** 2 + 0.0005 * py ** 3
    assert np.abs(o.interp(2, cub, px, py) - ce).max() < 1e-9
    assert np.abs(o.interp(3, cub, px, py) - ce).max() < 1e-9


def test_keys_border_fall_through():
    """DICe_Image.cpp:851-857: near the border keys_fourth_all = bilinear + Keys sum,
    while keys_fourth returns the bilinear value alone."""
    h, w = 20, 20
    img = np.arange(h * w, dtype=np.float64).reshape(h, w)
    x = np.array([2.25])
    y = np.array([10.5])
    single = o.interp(3, img, x, y)[0]
    bil = o.interp(0, img, x, y)[0]
    assert single == pytest.approx(bil)
    both = o.interp(2, img, x, y)[0]
    assert both == pytest.approx(bil + img[10, 2] + 0.25 + 0.5 * w, rel=1e-9)


def test_interp_shift_subpixel():
    """DICe_TestInterpShift.cpp:70-128: GN Keys recovers shifts 0..1 within 0.01 px."""
    w, h = 50, 40
    y, x = np.mgrid[0:h, 0:w].astype(np.float32)
    yv = np.abs(np.cos((y / np.float32(h)) * np.float32(5.0) * np.float32(math.pi)))
    ref = (np.abs(np.cos((x / np.float32(w)) * np.float32(5.0 * math.pi))) * yv * np.float32(255.0)).astype(np.float32)
    p = o.default_params()
    p.optimization_method = 1
    p.max_solver_iterations_fast = 100
    shift = np.float32(0.0)
    while shift <= 1.0:
        de = (np.abs(np.cos(((x - shift) / np.float32(w)) * np.float32(5.0 * math.pi))) * yv * np.float32(255.0)).astype(np.float32)
        f = o.new_fields([25.0], [20.0])
        o.correlate(p, ref, de, f, subset_size=21, precision="float")
        assert abs(f[o.F["DISP_X"], 0] - shift) <= 0.01
        shift = np.float32(shift + np.float32(0.1))


@pytest.mark.parametrize("precision", ["double", "float"])
def test_api_integer_shifts(precision):
    """DICe_TestAPI.cpp:70-150: dice_correlate with tracking defaults (SIMPLEX, Keys,
    u, v, theta); frame k is an integer shift u = v = k; fields carry over between frames.
    The test's 6-parameter list never reaches the static Schema (DICe_api.cpp:103-105
    calls set_params only on the first call or with update_params), so the tracking
    defaults are what both of its checks run with."""
    d = load("synthetic_speckled.npz")
    p = o.tracking_default_params()
    p.compute_def_gradients = 1  # dice_correlate's set_def_image(array) computes def gradients
    f = o.new_fields([47.0, 51.0, 51.0, 47.0], [47.0, 47.0, 51.0, 51.0])
    for k in range(6):
        o.correlate(p, d["def0"], d["def%d" % k], f, subset_size=21, first_frame=(k == 0), precision=precision)
        assert np.all(np.abs(f[o.F["DISP_X"]] - k) <= 1e-4)
        assert np.all(np.abs(f[o.F["DISP_Y"]] - k) <= 1e-4)
        assert np.all(np.abs(f[o.F["ROT_Z"]]) <= 1e-4)
        assert np.all(np.abs(f[o.F["GAMMA"]]) <= 1e-4)
        assert np.all(f[o.F["STATUS"]] == 1)


def test_condition_number_closed_form():
    """CONDITION_NUMBER (DICe_Objective.cpp:558-564): for a translation-only solve that ends at an
    integer shift every pixel is kept and theta = 0, so the u,v block of H is
    sum [gx^2, gx gy; gx gy, gy^2] over the subset with the reference gradients, and the stored value is
    ||H||_F * ||H^-1||_F = ||H||_F^2 / |det H|.  Computed here from the pinned FD gradients."""
    d = load("speckle_512.npz")
    ref = d["ref"].astype(np.float32)
    defm = np.roll(ref, (-1, 2), axis=(0, 1))  # defm[y, x] = ref[y + 1, x - 2]: u = 2, v = -1
    p = o.default_params()
    p.optimization_method = 1
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 0
    p.compute_uncertainty_fields = 0
    cx = np.array([100.0, 250.0, 400.0, 180.0])
    cy = np.array([120.0, 300.0, 200.0, 420.0])
    f = o.correlate(p, ref, defm, o.new_fields(cx, cy), subset_size=31, precision="double")
    assert np.all(f[o.F["STATUS"]] == 1)
    assert np.allclose(f[o.F["DISP_X"]], 2.0, atol=1e-3) and np.allclose(f[o.F["DISP_Y"]], -1.0, atol=1e-3)
    gx, gy = o.fd_gradients(ref.astype(np.float64))
    for i in range(len(cx)):
        sl = (slice(int(cy[i]) - 15, int(cy[i]) + 16), slice(int(cx[i]) - 15, int(cx[i]) + 16))
        a, b = gx[sl].ravel(), gy[sl].ravel()
        h00, h01, h11 = (a * a).sum(), (a * b).sum(), (b * b).sum()
        cond = (h00 ** 2 + 2 * h01 ** 2 + h11 ** 2) / abs(h00 * h11 - h01 ** 2)
        assert abs(f[o.F["CONDITION"]][i] - cond) <= 1e-5 * cond, (i, f[o.F["CONDITION"]][i], cond)


def test_lu_inverse_against_numpy_and_singular():
    rng = np.random.default_rng(7)
    for n in (2, 3, 4, 6):
        a = rng.normal(size=(n, n))
        a = a @ a.T + n * np.eye(n)
        inv, info = o.lu_inverse(a)
        assert info == 0
        assert np.allclose(inv, np.linalg.inv(a), rtol=1e-12, atol=1e-12)
        assert np.allclose(inv @ a, np.eye(n), atol=1e-12)
    # a zero matrix: INFO = 1 and the matrix is left as its (zero) factors, GETRI semantics
    z, info = o.lu_inverse(np.zeros((6, 6)))
    assert info == 1 and np.all(z == 0)


def test_flat_subset_statuses():
    """Flat reference subsets.  All-zero image: gradients are exactly 0, H = 0, det = 0 so
    cond = -1 (no singular flag), GETRI leaves the zero factors, the update is 0 and GN
    'converges' at iteration 0; sigma is -1 -> status 24 (DICe_Schema.cpp:2288-2295).
    Constant 100 image: the float FD stencil (c1, c2 = 1/12, -8/12) leaves ~2e-6 gradients,
    so sigma = 0 passes and the -1 ZNSSD gamma fails instead -> status 23."""
    p = o.default_params()
    p.optimization_method = 1
    for value, status in [(0.0, 24), (100.0, 23)]:
        ref = np.full((64, 64), value, np.float32)
        f = o.new_fields([32.0], [32.0])
        o.correlate(p, ref, ref.copy(), f, subset_size=21, precision="float")
        assert f[o.F["STATUS"], 0] == status
        assert f[o.F["ITERATIONS"], 0] == 0
        assert f[o.F["CONDITION"], 0] == -1.0


def test_out_of_view_and_failed_init():
    """FOV check on the first frame (status 12) and sigma = -1 initial field (status 5)."""
    d = load("speckle_512.npz")
    f = o.new_fields([5.0, 200.0], [200.0, 200.0])
    f[o.F["SIGMA"], 1] = -1.0
    o.correlate(o.default_params(), d["ref"], d["defm"], f, subset_size=27, precision="float")
    assert f[o.F["STATUS"], 0] == 12 and f[o.F["ITERATIONS"], 0] == -1
    assert f[o.F["STATUS"], 1] == 5 and f[o.F["GAMMA"], 1] == -1


def test_max_iterations_then_simplex_fallback():
    """max_solver_iterations_fast = 0 forces MAX_ITERATIONS_REACHED; GRADIENT_BASED records
    status 11, GRADIENT_BASED_THEN_SIMPLEX falls back to the simplex and succeeds."""
    d = load("speckle_512.npz")
    p = o.default_params()
    p.max_solver_iterations_fast = 0
    p.optimization_method = 1
    f = o.new_fields([200.0], [200.0])
    o.correlate(p, d["ref"], d["defm"], f, subset_size=27, precision="float")
    assert f[o.F["STATUS"], 0] == 11
    p.optimization_method = 4
    f = o.new_fields([200.0], [200.0])
    o.correlate(p, d["ref"], d["defm"], f, subset_size=27, precision="float")
    assert f[o.F["STATUS"], 0] == 1
    assert abs(f[o.F["DISP_X"], 0] - 0.4) < 0.1


@pytest.mark.parametrize("precision,key", [("float", "conformal_f"), ("double", "conformal_d")])
def test_polygon_raster_against_conformal_golden(precision, key):
    """DICe_TestSubset.cpp:248-309: the conformal subset (two polygons) written over its bounding box;
    non-subset pixels are 0 in the golden.  The reference's float build (conformal.rawi) and double build
    (conformal_d.rawi) differ by one pixel on a polygon edge, where the winding angle sums to pi up to
    rounding: the float and double rasters each reproduce their own golden exactly."""
    g = load("conformal_raster.npz")[key]
    xs, ys = [], []
    for vx, vy in [([940, 951, 964, 980, 959], [422, 399, 410, 413, 432]),
                   ([982, 1000, 991, 978], [426, 428, 456, 445])]:
        px, py = o.polygon_pixels(vx, vy, precision=precision)
        xs.append(px)
        ys.append(py)
    px = np.concatenate(xs)
    py = np.concatenate(ys)
    pts = sorted(set(zip(py.tolist(), px.tolist())))
    min_x = min(p[1] for p in pts)
    min_y = min(p[0] for p in pts)
    w = max(p[1] for p in pts) - min_x + 1
    h = max(p[0] for p in pts) - min_y + 1
    assert (h, w) == g.shape
    mask = np.zeros((h, w), bool)
    for yy, xx in pts:
        mask[yy - min_y, xx - min_x] = True
    # the speckle image has no zero intensity inside the subset: the non-zero pixels are the subset
    assert np.array_equal(mask, g != 0)