"""Pins of the oracle functions that the reference's own fixtures leave open, by what the mathematics
fixes (CPU only):

* Simplex::minimize (DICe_Simplex.cpp:60-277): evaluation sequences traced by hand on a quadratic bowl
  and a double well -- reflect, expand, contract, shrink, the extra evaluation at the end of every
  iteration, the stop rules (rtol, rtol difference, iteration cap, gamma(p0) below the threshold) and the
  reference's result, the LAST evaluated point.
* smooth_gradients_convolution_5_point (DICe_Image.cpp:1032-1062): the binomial 5 x 5 mask from an
  impulse, constant and linear fields unchanged inside, the 2-pixel border band untouched.
* computeUncertaintyFields FIELD_1..7 (DICe_Objective.cpp:193-299) without an image deformer, on the
  reference's integer-shift pairs (defSyntheticSpeckled0..5, u = v = k): FIELD_2 = 1, FIELD_3 =
  |u| sqrt(N_g), FIELD_5 = sqrt(sum (u.g)^2 / |g|^2), FIELD_1 = FIELD_5 / (|u| sqrt(N)), FIELD_6 =
  sqrt(sum (I_ref - I_def)^2 / |g|^2) (the interpolated deformed image at x + u is the reference pixel),
  FIELD_4 = FIELD_6 / FIELD_1, FIELD_7 = (1 + sum |g|^2) / N, with g the pinned FD gradients, N the
  subset's pixels and N_g those with |g|^2 >= 1e-4."""
import os

import numpy as np
import pytest

import oracle_py as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- Nelder-Mead

def test_simplex_reflect_expand_rtol_stop_last_point():
    """f(x) = x^2 from x0 = 3, delta 1.  Iteration 0: reflect to 2 (accepted, below the best vertex 3) so
    expand to 1 (accepted); the end-of-iteration evaluation at 1.  Iteration 1: reflect to -1 (f = 1,
    accepted, ties the best) so expand to -3 (f = 9, rejected); end evaluation at -3.  Iteration 2: both
    vertices at f = 1, rtol = 0 stops.  The result is the last evaluated point, -3."""
    r = o.nelder_mead(0, [3.0], [1.0], coef=[1.0, 0.0])
    expect = [3, 4, 2, 1, 1, -1, -3, -3]
    assert r["evals"] == 8
    np.testing.assert_array_equal(r["trace"][:, 0], expect)
    np.testing.assert_array_equal(r["trace"][:, 1], np.square(expect))
    assert r["status"] == 0 and r["iterations"] == 2 and r["stop"] == 2
    assert r["p"][0] == -3.0


def test_simplex_rtol_difference_stop():
    """f = x^2 + y^2, vertices (1,1), (2,1), (1,2).  Iteration 0: worst (2,1) reflects through the other
    two to (0,2), f = 4 < 5 accepted, not below the best (2) and not above the next worst (5): no
    expansion or contraction; end evaluation at (0,2).  Iteration 1: values (2, 4, 5) give the same rtol
    2*3/7 as before, so the rtol difference (0) stops it."""
    r = o.nelder_mead(0, [1.0, 1.0], [1.0, 1.0], coef=[1, 1, 0, 0])
    expect = [(1, 1, 2), (2, 1, 5), (1, 2, 5), (0, 2, 4), (0, 2, 4)]
    np.testing.assert_array_equal(r["trace"], np.array(expect, float))
    assert r["iterations"] == 1 and r["stop"] == 4 and r["status"] == 0
    np.testing.assert_array_equal(r["p"], [0.0, 2.0])


def test_simplex_contraction_sequence():
    """f = x^2 + y^2, vertices (0.5,0), (2.5,0), (0.5,2).  It 0: reflect (-1.5,2) ties the worst (6.25) and
    is not below the next worst (4.25), so contract to 0.5 centroid + 0.5 worst = (1.5,0.5), f = 2.5,
    accepted.  It 1: reflect (1.5,-1.5), f = 4.5, rejected; contract (0.75,1.125), f = 1.828125.  It 2:
    reflect (-0.25,0.625), f = 0.453125, accepted, neither below the best (0.25) nor above the next worst."""
    r = o.nelder_mead(0, [0.5, 0.0], [2.0, 2.0], coef=[1, 1, 0, 0])
    pts = [(0.5, 0), (2.5, 0), (0.5, 2), (-1.5, 2), (1.5, 0.5), (1.5, 0.5), (1.5, -1.5), (0.75, 1.125),
           (0.75, 1.125), (-0.25, 0.625), (-0.25, 0.625)]
    np.testing.assert_array_equal(r["trace"][:11, :2], np.array(pts, float))
    np.testing.assert_array_equal(r["trace"][:11, 2], [x * x + y * y for x, y in pts])
    # the same start with a cap of 3 iterations: iteration 2 is the cap (max_iterations - 1)
    c = o.nelder_mead(0, [0.5, 0.0], [2.0, 2.0], coef=[1, 1, 0, 0], max_iterations=3)
    assert c["status"] == 11 and c["iterations"] == 2 and c["evals"] == 9 and c["stop"] == 16
    np.testing.assert_array_equal(c["p"], [0.75, 1.125])


def test_simplex_shrink_sequence():
    """f = (x^2 - 1)^2 from -1.25, delta 2.4 (vertices -1.25 and 1.15).  It 0: reflect 3.55 is worse than
    the best; contract to the midpoint -0.05 (f = 0.995) is worse than the worst (0.316): shrink the worst
    vertex halfway to the best, 0.5 (-1.25 + 1.15) = -0.05, then the end evaluation there.  It 1: reflect
    2.35 (rejected), contract 0.55 (f = 0.4865, accepted).  It 2: reflect 1.75, contract 0.85."""
    r = o.nelder_mead(1, [-1.25], [2.4])
    x = np.array([-1.25, 1.15, 3.55, -0.05, -0.05, -0.05, 2.35, 0.55, 0.55, 1.75, 0.85, 0.85])
    np.testing.assert_allclose(r["trace"][:12, 0], x, rtol=0, atol=1e-14)
    np.testing.assert_allclose(r["trace"][:12, 1], (x * x - 1) ** 2, rtol=1e-12, atol=0)


def test_simplex_early_exit_below_threshold():
    """gamma(p0) in [0, threshold): success at iteration 0 after one evaluation, p unchanged."""
    r = o.nelder_mead(0, [0.25, -0.5], [1.0, 1.0], coef=[1, 1, 0.25, -0.5])
    assert r["status"] == 0 and r["iterations"] == 0 and r["evals"] == 1 and r["stop"] == 1
    np.testing.assert_array_equal(r["p"], [0.25, -0.5])
    # a negative value (a failed gamma) is not taken as converged
    n = o.nelder_mead(0, [0.25, -0.5], [1.0, 1.0], coef=[-1, -1, 0.0, 0.0])
    assert n["evals"] > 3


# ---------------------------------------------------------------- 5-point gradient smoothing

MASK1D = np.array([1, 4, 6, 4, 1], float) / 16.0


def test_smoothing_impulse_reproduces_binomial_mask():
    gx = np.zeros((12, 15), np.float32)
    gy = np.zeros((12, 15), np.float32)
    gx[6, 7] = 1.0
    gy[5, 9] = 2.0
    sx, sy = o.smooth_gradients(gx, gy)
    mask = np.outer(MASK1D, MASK1D)
    np.testing.assert_array_equal(sx[4:9, 5:10], mask.astype(np.float32))
    np.testing.assert_array_equal(sy[3:8, 7:12], (2 * mask).astype(np.float32))
    assert np.count_nonzero(sx) == 25 and np.count_nonzero(sy) == 25


def test_smoothing_constant_and_linear_fields_unchanged_inside():
    h, w = 20, 24
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    const = np.full((h, w), 3.5, np.float32)
    lin = (2.0 + 0.25 * x - 0.5 * y).astype(np.float32)
    sx, sy = o.smooth_gradients(const, lin)
    np.testing.assert_array_equal(sx, const)
    np.testing.assert_allclose(sy[2:-2, 2:-2], lin[2:-2, 2:-2], rtol=0, atol=1e-5)


def test_smoothing_leaves_the_border_band():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(16, 18)).astype(np.float32)
    sx, _ = o.smooth_gradients(g, g)
    np.testing.assert_array_equal(sx[:2], g[:2])
    np.testing.assert_array_equal(sx[-2:], g[-2:])
    np.testing.assert_array_equal(sx[:, :2], g[:, :2])
    np.testing.assert_array_equal(sx[:, -2:], g[:, -2:])
    # inside: the separable binomial filter of the input (not of already-smoothed neighbours)
    ref = np.zeros_like(g, dtype=np.float64)
    for i, a in enumerate(MASK1D):
        for j, b in enumerate(MASK1D):
            ref[2:-2, 2:-2] += a * b * g[i:i + g.shape[0] - 4, j:j + g.shape[1] - 4]
    np.testing.assert_allclose(sx[2:-2, 2:-2], ref[2:-2, 2:-2], rtol=0, atol=1e-5)


# ---------------------------------------------------------------- uncertainty fields

def uncertainty_case(k, cx, cy, S=21):
    d = np.load(os.path.join(GOLD, "synthetic_speckled.npz"))
    ref = d["def0"].astype(np.float64)
    de = d["def%d" % k].astype(np.float64)
    p = o.default_params()
    p.optimization_method = 1
    p.compute_uncertainty_fields = 1
    p.fast_solver_tolerance = 1e-8  # converge onto the integer shift
    params, res = o.objective(p, ref, de, S, cx, cy, 2, [k - 0.3, k + 0.2, 0, 0, 0, 0])
    assert res[0] == 0, res
    gx, gy = o.fd_gradients(ref)
    h = S // 2
    sl = (slice(cy - h, cy + h + 1), slice(cx - h, cx + h + 1))
    g = np.stack([gx[sl].ravel(), gy[sl].ravel()], axis=1)
    return params, res[3:10], g, ref[sl].ravel(), de[sl].ravel()


@pytest.mark.parametrize("k", [1, 2, 3, 5])
@pytest.mark.parametrize("centre", [(35, 40), (50, 50), (62, 44)])
def test_uncertainty_fields_closed_forms_on_integer_shifts(k, centre):
    params, fields, g, r, dd = uncertainty_case(k, *centre)
    u, v = params[0], params[1]
    assert abs(u - k) < 1e-6 and abs(v - k) < 1e-6
    f1, f2, f3, f4, f5, f6, f7 = fields
    n = len(g)
    g2 = np.sum(g * g, axis=1)
    keep = g2 >= 1e-4
    w = np.where(keep, 1.0 / np.where(keep, g2, 1.0), 0.0)
    assert f2 == 1.0
    np.testing.assert_allclose(f3, np.hypot(u, v) * np.sqrt(keep.sum()), rtol=1e-12)
    f5_cf = np.sqrt(np.sum((u * g[:, 0] + v * g[:, 1]) ** 2 * w))
    np.testing.assert_allclose(f5, f5_cf, rtol=1e-12)
    np.testing.assert_allclose(f1, f5_cf / (np.hypot(u, v) * np.sqrt(n)), rtol=1e-12)
    f6_cf = np.sqrt(np.sum((r - dd) ** 2 * w))
    np.testing.assert_allclose(f6, f6_cf, rtol=1e-5)
    np.testing.assert_allclose(f4, f6 / f1, rtol=1e-12)
    np.testing.assert_allclose(f7, (1.0 + g2.sum()) / n, rtol=1e-12)


def test_uncertainty_fields_zero_shift():
    """u = v = 0 (to the solver tolerance): FIELD_3 = FIELD_5 = 0; FIELD_6 = 0 (the deformed image is the
    reference), so FIELD_4 = 0; FIELD_1 stays a cosine."""
    params, fields, g, r, dd = uncertainty_case(0, 50, 50)
    f1, f2, f3, f4, f5, f6, f7 = fields
    assert abs(params[0]) < 1e-9 and abs(params[1]) < 1e-9
    assert f2 == 1.0
    assert abs(f3) < 1e-6 and abs(f5) < 1e-6 and abs(f6) < 1e-6 and abs(f4) < 1e-6
    assert 0.0 <= f1 <= 1.0 + 1e-12


# ---------------------------------------------------------------- blocking subsets

POLY = ([30, 52, 61, 40, 27], [20, 18, 39, 47, 35])


def test_deformed_polygon_identity_and_integer_translation():
    """Polygon::get_owned_pixels(shape_function, ...) (DICe_Shape.cpp:136-219): zero parameters and skin 1
    give the polygon's own pixels; an integer translation moves every vertex, and so every pixel, by it."""
    px, py = o.polygon_pixels(*POLY)
    dx, dy = o.deformed_polygon_pixels(*POLY, [0, 0, 0, 0, 0, 0], 44, 33, 1.0)
    assert sorted(zip(dy, dx)) == sorted(zip(py, px))
    tx, ty = o.deformed_polygon_pixels(*POLY, [3, -2, 0, 0, 0, 0], 44, 33, 1.0)
    assert sorted(zip(ty, tx)) == sorted(zip(py - 2, px + 3))


def test_deformed_polygon_skin_and_rotation():
    """The skin factor stretches the vertices about their integer centroid: skin 2 is the polygon whose
    vertices are 2 v - c (c = (int)(mean)); a quarter turn about a vertex maps (x, y) to
    (cx - (y - cy), cy + (x - cx)), the vertices land on integers."""
    vx, vy = np.array(POLY[0]), np.array(POLY[1])
    cx, cy = int(vx.mean()), int(vy.mean())
    sx, sy = o.deformed_polygon_pixels(vx, vy, [0, 0, 0, 0, 0, 0], 44, 33, 2.0)
    ex, ey = o.polygon_pixels(2 * vx - cx, 2 * vy - cy)
    assert sorted(zip(sy, sx)) == sorted(zip(ey, ex))
    rx, ry = o.deformed_polygon_pixels(vx, vy, [0, 0, np.pi / 2, 0, 0, 0], 30, 20, 1.0)
    qx, qy = o.polygon_pixels(30 - (vy - 20), 20 + (vx - 30))
    assert sorted(zip(ry, rx)) == sorted(zip(qy, qx))
