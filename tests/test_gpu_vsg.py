"""The VSG strain post-processor on the GPU (dice_hip_vsg_strain, include/dice_hip.h) against the oracle's
restatement of VSG_Strain_Post_Processor (DICe_PostProcessor.cpp:877-1136), which tests/test_host_cpp.py
pins to the dic_challenge_12_sorted gold.

Bar: MATCH identical; the seven VSG fields within one float ulp, or 1e-12 where they cancel to zero (the
reference stores them as scalar_t; the GPU sums the neighbours in another order than the oracle's sorted
one, in double)."""
import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = dh.Context(0)
    yield c
    c.close()


def check(ctx, cx, cy, fields, window):
    ctx.set_subsets(cx, cy, 21)
    fg = fields.copy()
    vg = ctx.vsg_strain(fg, window)
    fo = fields.copy()
    vo = o.vsg_strain(fo, window)
    assert np.array_equal(fg[dh.F["MATCH"]], fo[o.F["MATCH"]])
    # one float ulp, and 1e-12 absolute for strains that cancel to zero in one summation order and to
    # a few 1e-17 in the other
    scale = np.abs(vo)
    assert np.all(np.abs(vg - vo) <= 1.2e-7 * scale + 1e-12), np.abs(vg - vo).max()
    return vg, fg


@pytest.mark.parametrize("window", [41, 60, 121, 360])
def test_vsg_grid_parity(ctx, window):
    """A 1024^2 grid at step 10 (10,000 subsets) with a smooth nonlinear field, 2 % failed subsets and
    windows whose radius lands exactly on grid distances (even sizes) or between them (odd)."""
    rng = np.random.default_rng(window)
    cx, cy = dh.regular_grid(1024, 1024, 21, 10)
    f = dh.new_fields(cx, cy)
    f[dh.F["DISP_X"]] = (0.5 + 0.002 * cx + 1e-6 * cx * cy + 0.2 * np.sin(cy / 80)).astype(np.float32)
    f[dh.F["DISP_Y"]] = (-0.3 + 0.001 * cy - 2e-6 * cx * cx).astype(np.float32)
    f[dh.F["SIGMA"]] = 0.01
    f[dh.F["SIGMA"], rng.random(len(cx)) < 0.02] = -1.0
    vg, fg = check(ctx, cx, cy, f, window)
    assert np.mean(fg[dh.F["MATCH"]] == -1) < 0.05


def test_vsg_scattered_points_and_failures(ctx):
    """Irregular centres (conformal-style, with duplicates removed), sparse regions with fewer than three
    neighbours, and a row of collinear centres (singular normal matrix)."""
    rng = np.random.default_rng(3)
    pts = {(int(x), int(y)) for x, y in rng.uniform(30, 2000, size=(6000, 2))}
    pts |= {(3000 + 7 * k, 500) for k in range(40)}       # collinear row, far from the rest
    pts |= {(5000, 5000), (5003, 5001)}                    # two lonely subsets
    cx = np.array([p[0] for p in sorted(pts)], np.float64)
    cy = np.array([p[1] for p in sorted(pts)], np.float64)
    f = dh.new_fields(cx, cy)
    f[dh.F["DISP_X"]] = (0.1 * np.cos(cx / 300) + 0.003 * cy).astype(np.float32)
    f[dh.F["DISP_Y"]] = (0.2 * np.sin(cy / 250)).astype(np.float32)
    f[dh.F["SIGMA"]] = rng.uniform(0.001, 0.05, len(cx))
    f[dh.F["SIGMA"], rng.random(len(cx)) < 0.05] = -1.0
    vg, fg = check(ctx, cx, cy, f, 101)
    row = (cy == 500) & (cx >= 3000)
    assert np.all(fg[dh.F["MATCH"]][row] == -1) and np.all(vg[:, row] == 0)
    assert np.all(fg[dh.F["MATCH"]][cx >= 5000] == -1)


def test_vsg_errors(ctx):
    cx, cy = dh.regular_grid(256, 256, 21, 20)
    ctx.set_subsets(cx, cy, 21)
    f = dh.new_fields(cx, cy)
    for bad in (0, -65):
        with pytest.raises(dh.DiceHipError):
            ctx.vsg_strain(f, bad)


def test_uncertainty_post_processor_parity(ctx):
    """dice_hip_uncertainty against the oracle, bit for bit: UNCERTAINTY = SIGMA / FIELD_1 (0 for failed
    subsets and for FIELD_1 = 0), UNCERTAINTY_ANGLE = FIELD_1."""
    rng = np.random.default_rng(11)
    cx, cy = dh.regular_grid(512, 512, 21, 8)
    ctx.set_subsets(cx, cy, 21)
    f = dh.new_fields(cx, cy)
    f[dh.F["SIGMA"]] = rng.uniform(0.001, 0.05, len(cx))
    f[dh.F["SIGMA"], ::13] = -1.0
    f[dh.F["FIELD_1"]] = rng.uniform(0.2, 1.0, len(cx)).astype(np.float32)
    f[dh.F["FIELD_1"], ::17] = 0.0
    ug = ctx.uncertainty(f)
    uo = o.uncertainty(f)
    assert np.array_equal(ug, uo)
    assert np.all(ug[0, ::13] == 0) and np.all(ug[0, ::17] == 0)


@pytest.mark.parametrize("k", [9, 13, 20])
def test_vsg_k_nearest_parity(ctx, k):
    """k-nearest windows (negative sizes): scattered centres, so that no two neighbours tie at the k-th
    distance (the reference's kd-tree orders ties its own way, the GPU and the oracle by id)."""
    rng = np.random.default_rng(k)
    pts = sorted({(int(x), int(y)) for x, y in rng.uniform(20, 1500, size=(4000, 2))})
    cx = np.array([p[0] for p in pts], np.float64)
    cy = np.array([p[1] for p in pts], np.float64)
    f = dh.new_fields(cx, cy)
    f[dh.F["DISP_X"]] = (0.4 * np.sin(cx / 200) + 0.002 * cy).astype(np.float32)
    f[dh.F["DISP_Y"]] = (0.3 * np.cos(cy / 150)).astype(np.float32)
    f[dh.F["SIGMA"]] = 0.01
    f[dh.F["SIGMA"], rng.random(len(cx)) < 0.03] = -1.0
    check(ctx, cx, cy, f, -k)


def test_vsg_grid_edge_cases(ctx):
    """Edge cases of the grid built on the device, against the oracle: one subset; every centre at one
    point (a one-cell grid, singular plane fits); centres spread over 10^6 px, so that the 4000-cell cap
    sets the cell side instead of the radius; and a non-finite centre, which is an error."""
    one_x, one_y = np.array([100.0]), np.array([200.0])
    f = dh.new_fields(one_x, one_y)
    f[dh.F["SIGMA"]] = 0.01
    vg, fg = check(ctx, one_x, one_y, f, 41)
    assert fg[dh.F["MATCH"], 0] == -1 and np.all(vg == 0)

    same_x, same_y = np.full(50, 300.0), np.full(50, 400.0)
    f = dh.new_fields(same_x, same_y)
    f[dh.F["SIGMA"]] = 0.01
    f[dh.F["DISP_X"]] = np.linspace(0, 1, 50).astype(np.float32)
    check(ctx, same_x, same_y, f, 41)
    check(ctx, same_x, same_y, f, -9)

    rng = np.random.default_rng(5)
    base = rng.uniform(0, 1.0e6, size=(200, 2))
    pts = sorted({(int(x + dx), int(y + dy)) for x, y in base for dx in range(0, 40, 8) for dy in range(0, 40, 8)})
    wx = np.array([p[0] for p in pts], np.float64)
    wy = np.array([p[1] for p in pts], np.float64)
    f = dh.new_fields(wx, wy)
    f[dh.F["SIGMA"]] = 0.01
    f[dh.F["DISP_X"]] = (1e-3 * (wx % 997) + 2e-3 * (wy % 991)).astype(np.float32)
    f[dh.F["DISP_Y"]] = (-1e-3 * (wy % 983)).astype(np.float32)
    vg, fg = check(ctx, wx, wy, f, 41)
    assert np.mean(fg[dh.F["MATCH"]] == -1) < 0.5

    f = dh.new_fields(wx, wy)
    f[dh.F["COORD_X"], 7] = np.nan
    with pytest.raises(dh.DiceHipError):
        ctx.vsg_strain(f, 41)
