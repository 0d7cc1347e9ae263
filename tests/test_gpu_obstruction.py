"""GPU parity, through the C ABI, against the CPU oracle for the per-subset options of the generic
correlation routine and for sub-images:

* blocked image pixels (Subset::obstructed_coords_ from OBSTRUCTED shapes, DICe_Subset.cpp:162-167, and
  the TRACKING routine's pixels_blocked_by_other_subsets_): a warped subset pixel rounding onto one is
  deactivated (Subset::initialize, DICe_Subset.cpp:404-419; is_obstructed_pixel :447-460);
* force_simplex (FORCE_SIMPLEX and the blocking subsets, generic_correlation_routine,
  DICe_Schema.cpp:2160-2185, 2197-2204, 2226-2229): computeUpdateRobust with no fallback optimizer, also
  inside a Gauss-Newton run;
* image offsets (Image_::offset_x / offset_y, DICe_Image.h:168-176): sub-images of a frame with global
  subset coordinates give the results of the same crop correlated as a whole image with coordinates
  shifted by the offsets.
Gauss-Newton subsets at the bar of tests/gpu_common.py; Nelder-Mead subsets against the float/double
spread of the reference algorithm."""
import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth
from gpu_common import compare, compare_spread, run_oracle, to_oracle

pytestmark = pytest.mark.gpu

MODES = [dh.MODE_BATCHED, dh.MODE_PLAIN]
MODE_NAMES = {dh.MODE_PLAIN: "plain", dh.MODE_BATCHED: "batched"}


def gn6(opt=dh.GRADIENT_BASED):
    p = dh.default_params()
    p.optimization_method = opt
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    return p


def obstruction_masks(cx, cy, S, seed):
    """Per subset: a static band of image pixels over a part of it (a bar 5 px wide at a random offset
    and direction), or nothing for every fourth subset."""
    rng = np.random.default_rng(seed)
    off, xs, ys = [0], [], []
    h = S // 2
    for i, (x, y) in enumerate(zip(cx.astype(int), cy.astype(int))):
        if i % 4 == 3:
            off.append(off[-1])
            continue
        a = int(rng.integers(-h, h - 5))
        if rng.random() < 0.5:
            yy, xx = np.mgrid[y - h - 2:y + h + 3, x + a:x + a + 5]
        else:
            yy, xx = np.mgrid[y + a:y + a + 5, x - h - 2:x + h + 3]
        xs.append(xx.ravel())
        ys.append(yy.ravel())
        off.append(off[-1] + xx.size)
    return (np.array(off, np.int32), np.concatenate(xs).astype(np.int32), np.concatenate(ys).astype(np.int32))


@pytest.fixture(scope="module")
def scene():
    ref, de = synth.config1(seed=11)
    S = 31
    cx, cy = dh.regular_grid(512, 512, S, 24)
    f0 = dh.new_fields(cx, cy)
    f0[dh.F["DISP_X"]] = 0.5 + 0.01 * cx + 0.3
    f0[dh.F["DISP_Y"]] = -0.3 + 0.005 * cy - 0.3
    return ref, de, cx, cy, S, f0


def run_case(scene, p, mode, masks=None, force=None):
    ref, de, cx, cy, S, f0 = scene
    with dh.Context(0) as ctx:
        ctx.set_params(p)
        ctx.set_ref_image(ref)
        ctx.set_def_image(de)
        ctx.set_subsets(cx, cy, S)
        ctx.set_subset_masks(masks)
        ctx.set_force_simplex(force)
        fg = ctx.correlate(f0.copy(), mode=mode)
    kw = dict(fields=f0, force_simplex=force, masks=masks)
    fd = run_oracle(p, ref, de, cx, cy, S, **kw)
    ff = run_oracle(p, ref, de, cx, cy, S, precision="float", **kw)
    return fg, fd, ff


def sub(f, idx):
    return np.ascontiguousarray(f[:, idx])


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_obstructed_pixels_gauss_newton(scene, mode):
    ref, de, cx, cy, S, f0 = scene
    masks = obstruction_masks(cx, cy, S, 3)
    p = gn6()
    fg, fd, ff = run_case(scene, p, mode, masks=masks)
    agree, both = compare(fg, fd, label="obstructed")
    assert both.mean() > 0.95
    # the blocked pixels are left out of the subset: fewer active pixels where a band crosses it
    full, _, _ = run_case(scene, p, mode)
    banded = np.diff(masks[0]) > 0
    assert np.all(fg[dh.F["ACTIVE_PIXELS"]][banded] < full[dh.F["ACTIVE_PIXELS"]][banded])
    assert np.array_equal(fg[dh.F["ACTIVE_PIXELS"]][~banded], full[dh.F["ACTIVE_PIXELS"]][~banded])


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("opt", [dh.GRADIENT_BASED, dh.GRADIENT_BASED_THEN_SIMPLEX,
                                 dh.SIMPLEX_THEN_GRADIENT_BASED], ids=["gn", "gn_then_simplex", "simplex_then_gn"])
def test_force_simplex(scene, mode, opt):
    ref, de, cx, cy, S, f0 = scene
    n = len(cx)
    force = (np.arange(n) % 3 == 0).astype(np.uint8)
    masks = obstruction_masks(cx, cy, S, 5)
    p = gn6(opt)
    p.max_solver_iterations_fast = 3  # some Gauss-Newton subsets fail and take the fallback
    fg, fd, ff = run_case(scene, p, mode, masks=masks, force=force)
    forced = force == 1
    # forced subsets ran Nelder-Mead only: its iteration counts, never a Gauss-Newton fallback
    compare_spread(sub(fg, forced), sub(fd, forced), sub(ff, forced), label="forced", nelder_mead=True)
    it = fg[dh.F["ITERATIONS"]][forced]
    ok = np.isin(fg[dh.F["STATUS"]][forced], dh.CONVERGED_STATUSES)
    assert np.median(it[ok]) > 10  # Nelder-Mead iteration counts, not Gauss-Newton's
    rest = ~forced
    if opt == dh.GRADIENT_BASED:
        compare(sub(fg, rest), sub(fd, rest), label="not forced")
    else:
        compare_spread(sub(fg, rest), sub(fd, rest), sub(ff, rest), label="not forced", nelder_mead=True)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_sub_image_offsets(scene, mode):
    """The crop [x0, x0 + w) x [y0, y0 + h) of both images with offsets (x0, y0) and global subset
    coordinates against the oracle on the crop as a whole image with the coordinates shifted; subsets
    near the crop's edge lose pixels to the 4-px guard of the sub-image, some of them fail the
    field-of-view test."""
    ref, de, cx, cy, S, f0 = scene
    x0, y0, w, h = 97, 61, 300, 340
    rc = np.ascontiguousarray(ref[y0:y0 + h, x0:x0 + w])
    dc = np.ascontiguousarray(de[y0:y0 + h, x0:x0 + w])
    inside = (cx > x0 + 10) & (cx < x0 + w - 10) & (cy > y0 + 10) & (cy < y0 + h - 10)
    gx, gy = cx[inside], cy[inside]
    fi = sub(f0, inside)
    p = gn6()
    with dh.Context(0) as ctx:
        ctx.set_params(p)
        ctx.set_ref_image(rc)
        ctx.set_def_image(dc)
        ctx.set_image_offsets(x0, y0, x0, y0)
        ctx.set_subsets(gx, gy, S)
        fg = ctx.correlate(fi.copy(), mode=mode)
    assert np.array_equal(fg[dh.F["COORD_X"]], gx)
    fl = fi.copy()
    fl[o.F["COORD_X"]] -= x0
    fl[o.F["COORD_Y"]] -= y0
    fd = run_oracle(p, rc, dc, gx - x0, gy - y0, S, fields=fl)
    fd[o.F["COORD_X"]] += x0
    fd[o.F["COORD_Y"]] += y0
    # the noise window clamps to the global pixel 0 past the far edge (DICe_Subset.cpp:663-664), which
    # a shifted whole image would place differently: sigma and noise are compared on the other subsets
    agree, both = compare(fg, fd, label="sub-image")
    assert both.sum() > 50 and (~both).sum() > 0
    assert np.any(fg[dh.F["STATUS"]] == 12)  # field of view at the crop's edge
    near_far_edge = (gx + S // 2 + 3 >= x0 + w) | (gy + S // 2 + 3 >= y0 + h)
    ok = both & ~near_far_edge
    rel = np.abs(fg[dh.F["SIGMA"]] - fd[o.F["SIGMA"]])[ok] / np.abs(fd[o.F["SIGMA"]][ok])
    assert rel.max() < 1e-4
    assert np.array_equal(fg[dh.F["ACTIVE_PIXELS"]][both], fd[o.F["ACTIVE_PIXELS"]][both])
