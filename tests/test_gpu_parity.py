"""GPU parity: the HIP library (through its C ABI) against the CPU oracle and the
reference's own golds, on the same seeded inputs.

Bar (floating point, scalar_t = float on the GPU vs the oracle's double build):
displacement and rotation within 1e-4 px / rad, gamma within 1e-5 relative on
subsets where both converge, identical status on >= 99.9 % of subsets.
"""
import math
import os

import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = [dh.MODE_PLAIN, dh.MODE_BATCHED]
MODE_NAMES = {dh.MODE_PLAIN: "plain", dh.MODE_BATCHED: "batched"}

U_TOL = 1e-4
GAMMA_RTOL = 1e-5
STATUS_AGREE = 0.999


def load(name):
    return np.load(os.path.join(GOLD, name))


def gold_table(name):
    rows = [ln.split() for ln in open(os.path.join(GOLD, name)) if ln[:1].isdigit()]
    return np.array([[float(x) for x in r] for r in rows])


def to_oracle(p):
    q = o.Params()
    for name, _ in dh.Params._fields_:
        setattr(q, name, getattr(p, name))
    return q


@pytest.fixture(scope="module")
def ctx():
    c = dh.Context(0)
    yield c
    c.close()


def run_gpu(ctx, p, ref, de, cx, cy, subset_size, mode, fields=None, conformal=None, first_frame=True):
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(de)
    ctx.set_subsets(cx, cy, subset_size, conformal)
    f = dh.new_fields(cx, cy) if fields is None else fields.copy()
    ctx.correlate(f, mode=mode, first_frame=first_frame)
    return f


def run_oracle(p, ref, de, cx, cy, subset_size, fields=None, conformal=None, first_frame=True,
               precision="double"):
    f = o.new_fields(cx, cy) if fields is None else fields.copy()
    o.correlate(to_oracle(p), ref, de, f, subset_size=subset_size, conformal=conformal, first_frame=first_frame,
                precision=precision, threads=8)
    return f


def compare(fg, fo, u_tol=U_TOL, gamma_rtol=GAMMA_RTOL, status_agree=STATUS_AGREE, label=""):
    sg, so = fg[dh.F["STATUS"]], fo[o.F["STATUS"]]
    agree = float(np.mean(sg == so))
    if agree < status_agree:
        pairs, counts = np.unique(np.stack([sg[sg != so], so[sg != so]]), axis=1, return_counts=True)
        detail = ", ".join("gpu %d / oracle %d: %d" % (a, b, c) for (a, b), c in zip(pairs.T, counts))
        raise AssertionError("%s status agreement %.4f (%s)" % (label, agree, detail))
    both = np.isin(sg, dh.CONVERGED_STATUSES) & np.isin(so, dh.CONVERGED_STATUSES)
    for name in ("DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY"):
        d = np.abs(fg[dh.F[name]] - fo[o.F[name]])[both]
        assert d.size == 0 or d.max() <= u_tol, "%s %s max diff %.3g" % (label, name, d.max())
    g, go = fg[dh.F["GAMMA"]][both], fo[o.F["GAMMA"]][both]
    rel = np.abs(g - go) / np.maximum(np.abs(go), 1e-30)
    assert rel.size == 0 or rel.max() <= gamma_rtol, "%s gamma max rel diff %.3g" % (label, rel.max())
    return agree, both


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_parallel_subsets_gold_gpu(ctx, mode):
    """tests/regression/parallel_subsets gold through the GPU path."""
    d = load("speckle_512.npz")
    g = gold_table("parallel_subsets_gold.txt")
    p = dh.default_params()
    fg = run_gpu(ctx, p, d["ref"], d["defm"], g[:, 1], g[:, 2], 27, mode)
    assert np.abs(fg[dh.F["DISP_X"]] - g[:, 3]).max() < 5e-5
    assert np.abs(fg[dh.F["DISP_Y"]] - g[:, 4]).max() < 5e-5
    assert np.abs(fg[dh.F["GAMMA"]] - g[:, 8]).max() < 1e-6
    assert np.array_equal(fg[dh.F["ITERATIONS"]], g[:, 9])
    assert np.array_equal(fg[dh.F["STATUS"]], g[:, 10])
    fo = run_oracle(p, d["ref"], d["defm"], g[:, 1], g[:, 2], 27)
    compare(fg, fo, status_agree=1.0, label="parallel_subsets")
    # sigma / noise / contrast / active pixels of the epilogue
    for name, tol in (("SIGMA", 1e-4), ("NOISE", 1e-4), ("CONTRAST", 1e-4)):
        rel = np.abs(fg[dh.F[name]] - fo[o.F[name]]) / np.maximum(np.abs(fo[o.F[name]]), 1e-12)
        assert rel.max() < tol, name
    assert np.array_equal(fg[dh.F["ACTIVE_PIXELS"]], fo[o.F["ACTIVE_PIXELS"]])


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_single_speckled_shift_sigma_gpu(ctx, mode):
    """tests/regression/single_speckled_shift: u, v and SIGMA."""
    d = load("speckle_512.npz")
    g = gold_table("single_speckled_shift_gold.txt")
    fg = run_gpu(ctx, dh.default_params(), d["ref"], d["defm"], g[:, 1], g[:, 2], 27, mode)
    assert np.abs(fg[dh.F["DISP_X"]] - g[:, 3]).max() < 5e-5
    assert np.abs(fg[dh.F["DISP_Y"]] - g[:, 4]).max() < 5e-5
    assert np.abs(fg[dh.F["SIGMA"]] - g[:, 9]).max() < 2e-6
    assert np.all(fg[dh.F["STATUS"]] == 1)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_config1_affine_gn_parity(ctx, mode):
    """512^2 speckle, S=31 step 10 (2116 subsets), 6-param affine GN, Keys-4: vs the double oracle."""
    ref, de = synth.config1()
    cx, cy = dh.regular_grid(512, 512, 31, 10)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    fg = run_gpu(ctx, p, ref, de, cx, cy, 31, mode)
    fo = run_oracle(p, ref, de, cx, cy, 31)
    agree, both = compare(fg, fo, label="config1")
    assert both.mean() > 0.97
    # accuracy vs the imposed field (subset-centre displacement)
    ut = 0.5 + 0.01 * cx
    vt = -0.3 + 0.005 * cy
    assert np.abs(fg[dh.F["DISP_X"]] - ut)[both].mean() <= 0.01
    assert np.abs(fg[dh.F["DISP_Y"]] - vt)[both].mean() <= 0.01
    same_it = both & (fg[dh.F["ITERATIONS"]] == fo[o.F["ITERATIONS"]])
    assert same_it.sum() >= 0.99 * both.sum()
    # uncertainty fields and the condition number
    for name in ("FIELD_1", "FIELD_3", "FIELD_5", "FIELD_6", "FIELD_7", "CONDITION"):
        a, b = fg[dh.F[name]][same_it], fo[o.F[name]][same_it]
        rel = np.abs(a - b) / np.maximum(np.abs(b), 1e-6)
        assert np.median(rel) < 1e-3 and rel.max() < 5e-2, name


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("opt", [0, 3, 1, 4])
@pytest.mark.parametrize("interp", [0, 2])
def test_correlate_params_known_answer_gpu(ctx, mode, opt, interp):
    """DICe_TestCorrelateParams.cpp:78-140 through the GPU path."""
    d = load("subset_construction.npz")
    p = dh.default_params()
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.robust_solver_tolerance = 1e-4
    p.optimization_method = opt
    p.interpolation_method = interp
    f = dh.new_fields([30.0], [30.0])
    f[dh.F["DISP_X"]] = 9.8 - 0.25
    f[dh.F["DISP_Y"]] = -7.62 - 0.25
    f[dh.F["ROT_Z"]] = 0.262 - 0.05
    fg = run_gpu(ctx, p, d["ref"], d["defm"], [30.0], [30.0], 21, mode, fields=f)
    v = fg[:, 0]
    assert abs(v[dh.F["DISP_X"]] - 9.8) < 1e-2
    assert abs(v[dh.F["DISP_Y"]] + 7.62) < 1e-2
    assert abs(v[dh.F["ROT_Z"]] - 0.262) < 1e-2
    assert abs(v[dh.F["NS_XX"]] - 0.121) < 1e-2
    assert abs(v[dh.F["NS_YY"]] - 0.121) < 1e-2
    assert abs(v[dh.F["SS_XY"]]) < 1e-2
    assert v[dh.F["STATUS"]] == 1
    assert abs(v[dh.F["SIGMA"]] + 1) >= 1e-2


def test_correlate_points_api_integer_shifts():
    """DICe_TestAPI.cpp:70-150 through the dice_correlate drop-in (static context)."""
    d = load("synthetic_speckled.npz")
    pts = np.zeros((4, 9), np.float32)
    pts[:, 0] = [47, 51, 51, 47]
    pts[:, 1] = [47, 47, 51, 51]
    for k in range(6):
        rc, pts = dh.correlate_points(pts, 21, d["def0"], d["def%d" % k])
        assert rc == 0
        assert np.all(np.abs(pts[:, 2] - k) <= 1e-4)
        assert np.all(np.abs(pts[:, 3] - k) <= 1e-4)
        assert np.all(np.abs(pts[:, 4]) <= 1e-4)
        assert np.all(np.abs(pts[:, 6]) <= 1e-4)
        assert np.all(pts[:, 8] == 1)
    # size mismatch is rejected as in the reference
    rc, _ = dh.correlate_points(pts, 21, d["def0"], d["def0"][:50])
    assert rc == -1


def texture_u16(w, h, seed, n_sin=20, contrast=0.05, noise=0.005, shift=(0.0, 0.0), shear=0.0):
    """12-bit low-contrast sinusoid texture (config-5 family), sampled at the reference point
    X = x - u(X) of a translation + shear field, rounded to uint16."""
    rng = np.random.default_rng(seed)
    kx = rng.uniform(0.05, 0.6, n_sin) * rng.choice([-1, 1], n_sin)
    ky = rng.uniform(0.05, 0.6, n_sin) * rng.choice([-1, 1], n_sin)
    ph = rng.uniform(0, 2 * np.pi, n_sin)
    y, x = np.mgrid[0:h, 0:w].astype(np.float64)
    X = x - shift[0] - shear * y
    Y = y - shift[1]
    t = sum(np.sin(kx[i] * X + ky[i] * Y + ph[i]) for i in range(n_sin)) / math.sqrt(n_sin / 2.0)
    img = 2048.0 + contrast * 4095.0 * 0.5 * t
    img += np.random.default_rng(seed + 7).normal(0, noise * 4095.0, img.shape)
    return np.clip(np.round(img), 0, 4095).astype(np.uint16)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_simplex_uint16_bilinear_parity(ctx, mode):
    """Gradient-free path on 12-bit uint16 storage: 9x9 subsets, Nelder-Mead, bilinear, 6 params."""
    w, h = 160, 120
    ref = texture_u16(w, h, 11)
    de = texture_u16(w, h, 11, shift=(0.7, -0.4), shear=0.01)
    cx, cy = dh.regular_grid(w, h, 9, 4)
    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.interpolation_method = dh.BILINEAR
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.max_solver_iterations_robust = 200
    fg = run_gpu(ctx, p, ref, de, cx, cy, 9, mode)
    fo = run_oracle(p, ref, de, cx, cy, 9, precision="float")
    fd = run_oracle(p, ref, de, cx, cy, 9, precision="double")
    # Nelder-Mead branches on comparisons of nearly equal gamma values, so trajectories
    # fork on rounding: the float and double builds of the reference algorithm already
    # disagree on a few percent of these 200-iteration-capped solves.  The GPU must not
    # disagree with the float build more than that inherent spread (plus 1 %).
    simplex_spread_check(fg, fo, fd)


def simplex_spread_check(fg, fo, fd):
    """Nelder-Mead branches on comparisons of nearly equal gamma values, so its path
    forks on rounding: the float and double builds of the reference algorithm already
    take different iteration counts on most subsets and disagree on the status of a few
    iteration-capped ones.  The GPU result must sit inside that inherent spread: status
    agreement with the float build no worse than float-vs-double minus 1 %, and
    displacement differences to the double build of the same size as float-vs-double."""
    sg, so, sd = fg[dh.F["STATUS"]], fo[o.F["STATUS"]], fd[o.F["STATUS"]]
    inherent = float(np.mean(so == sd))
    assert np.mean(sg == so) >= min(0.99, inherent - 0.01)
    ok = np.isin(sg, dh.CONVERGED_STATUSES) & np.isin(so, dh.CONVERGED_STATUSES) & np.isin(sd, dh.CONVERGED_STATUSES)
    assert ok.mean() > 0.9
    for name in ("DISP_X", "DISP_Y"):
        gpu_vs_d = np.abs(fg[dh.F[name]] - fd[o.F[name]])[ok]
        f_vs_d = np.abs(fo[o.F[name]] - fd[o.F[name]])[ok]
        assert np.median(gpu_vs_d) <= 2.0 * np.median(f_vs_d) + 1e-5, name
        assert np.percentile(gpu_vs_d, 95) <= 2.0 * np.percentile(f_vs_d, 95) + 1e-4, name
    it_g = fg[dh.F["ITERATIONS"]][ok].mean()
    it_o = fo[o.F["ITERATIONS"]][ok].mean()
    assert abs(it_g - it_o) <= 0.1 * it_o


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("nparams", [2, 3, 6])
def test_simplex_speckle_parity(ctx, mode, nparams):
    """SIMPLEX on the reference's speckle pair (S=27, 144 subsets) with 2, 3 and 6 parameters."""
    d = load("speckle_512.npz")
    cx, cy = dh.regular_grid(512, 512, 27, 40)
    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.enable_rotation = 1 if nparams >= 3 else 0
    p.enable_normal_strain = p.enable_shear_strain = 1 if nparams == 6 else 0
    fg = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, 27, mode)
    fo = run_oracle(p, d["ref"], d["defm"], cx, cy, 27, precision="float")
    fd = run_oracle(p, d["ref"], d["defm"], cx, cy, 27, precision="double")
    simplex_spread_check(fg, fo, fd)
    # against the imposed shift of this pair (u, v ~ 0.4 px) all three agree to the
    # simplex tolerance
    assert np.abs(fg[dh.F["DISP_X"]] - fd[o.F["DISP_X"]]).max() < 2e-2


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
@pytest.mark.parametrize("nparams", [2, 6])
def test_simplex_lane_path_speckle(ctx, mode, nparams):
    """SIMPLEX on 15 x 15 subsets (225 px): small enough for the BATCHED mode to run the Nelder-Mead
    iterations 16 lanes per subset, the persistent kernel doing the checks and the epilogue."""
    d = load("speckle_512.npz")
    cx, cy = dh.regular_grid(512, 512, 15, 24)
    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1 if nparams == 6 else 0
    fg = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, 15, mode)
    fo = run_oracle(p, d["ref"], d["defm"], cx, cy, 15, precision="float")
    fd = run_oracle(p, d["ref"], d["defm"], cx, cy, 15, precision="double")
    simplex_spread_check(fg, fo, fd)
    for name in ("SIGMA", "NOISE", "CONTRAST"):
        ok = np.isin(fg[dh.F["STATUS"]], dh.CONVERGED_STATUSES) & np.isin(fd[o.F["STATUS"]], dh.CONVERGED_STATUSES)
        rel = np.abs(fg[dh.F[name]] - fd[o.F[name]])[ok] / np.maximum(np.abs(fd[o.F[name]][ok]), 1e-12)
        assert np.median(rel) < 1e-3, name


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_simplex_lane_path_conformal(ctx, mode):
    """SIMPLEX on polygon subsets of under 256 pixels (with excluded pixels) through the 16-lanes-per-subset
    iterations."""
    ref, de = synth.config1(seed=5)
    cx, cy, conf = conformal_set(512, 512, 120, 4, rmin=7, rmax=8.8)
    assert np.diff(conf[0]).max() <= 256
    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    fg = run_gpu(ctx, p, ref, de, cx, cy, 0, mode, conformal=conf)
    fo = run_oracle(p, ref, de, cx, cy, 0, conformal=conf, precision="float")
    fd = run_oracle(p, ref, de, cx, cy, 0, conformal=conf, precision="double")
    simplex_spread_check(fg, fo, fd)


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_fallback_lane_path(ctx, mode):
    """GRADIENT_BASED_THEN_SIMPLEX on 15 x 15 subsets: the simplex fallback pass over the device list runs
    16 lanes per subset; FOV and initial-guess failures keep their statuses."""
    d = load("speckle_512.npz")
    cx = np.array([200.0, 4.0, 300.0, 150.0, 260.0, 330.0])
    cy = np.array([200.0, 200.0, 300.0, 150.0, 120.0, 400.0])
    f0 = dh.new_fields(cx, cy)
    f0[dh.F["SIGMA"], 2] = -1.0
    p = dh.default_params()
    p.max_solver_iterations_fast = 0
    p.optimization_method = dh.GRADIENT_BASED_THEN_SIMPLEX
    fg = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, 15, mode, fields=f0)
    fo = run_oracle(p, d["ref"], d["defm"], cx, cy, 15, fields=f0)
    assert np.array_equal(fg[dh.F["STATUS"]], fo[o.F["STATUS"]])
    assert fg[dh.F["STATUS"], 1] == 12 and fg[dh.F["STATUS"], 2] == 5
    ok = np.isin(fg[dh.F["STATUS"]], dh.CONVERGED_STATUSES)
    assert ok.sum() == 4
    assert np.abs(fg[dh.F["DISP_X"]] - fo[o.F["DISP_X"]])[ok].max() < 2e-2


@pytest.mark.parametrize("size,delta", [(9, None), (15, None), (9, 2.0)], ids=["9x9", "15x15", "9x9-wide-steps"])
@pytest.mark.parametrize("interp", [dh.BILINEAR, dh.KEYS_FOURTH], ids=["bilinear", "keys"])
def test_simplex_plain_and_batched_paths_agree(ctx, size, delta, interp):
    """Subsets of at most 256 pixels sum the Nelder-Mead objective in one order in both modes (8 strided
    lane partials for subsets of at most 88 pixels, else 16, then a pairwise tree), so the BATCHED
    iterations in rows of 8 or 16 lanes and the PLAIN workgroup follow the same path: identical statuses,
    iteration counts and final parameters on every subset (the epilogue's other sums may differ in
    rounding).  The wide-steps case starts the simplex with strain and rotation steps of 2, so that the
    first strain and shear vertices already stretch a 9 x 9 subset past the 256-cell window of its row
    (27 x 11 cells bilinear): bilinear rows then read those points' taps from the image, Keys rows hand
    the subset to the PLAIN kernel through the overflow list."""
    d = load("speckle_512.npz")
    cx, cy = dh.regular_grid(512, 512, size, 12)
    p = dh.default_params()
    p.optimization_method = dh.SIMPLEX
    p.interpolation_method = interp
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    if delta is not None:
        p.robust_delta_theta = delta
    fb = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, size, dh.MODE_BATCHED)
    handed_over = ctx.pass_counts()[1]
    fp = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, size, dh.MODE_PLAIN)
    if delta is not None and interp == dh.KEYS_FOURTH:
        assert handed_over > 0
    assert np.array_equal(fb[dh.F["STATUS"]], fp[dh.F["STATUS"]])
    assert np.array_equal(fb[dh.F["ITERATIONS"]], fp[dh.F["ITERATIONS"]])
    for name in ("DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY"):
        assert np.array_equal(fb[dh.F[name]], fp[dh.F[name]]), name


def conformal_set(w, h, n, seed, rmin=10, rmax=20):
    """Random convex polygons (vertices on a jittered circle) rasterised by the
    Polygon::get_owned_pixels rule (oracle restatement used only to build inputs here)."""
    rng = np.random.default_rng(seed)
    offs, xs, ys, acts, cxs, cys = [0], [], [], [], [], []
    for _ in range(n):
        cx = int(rng.uniform(rmax + 10, w - rmax - 10))
        cy = int(rng.uniform(rmax + 10, h - rmax - 10))
        r = rng.uniform(rmin, rmax)
        nv = int(rng.integers(5, 9))
        ang = np.sort(rng.uniform(0, 2 * np.pi, nv))
        vx = np.round(cx + r * np.cos(ang)).astype(np.int32)
        vy = np.round(cy + r * np.sin(ang)).astype(np.int32)
        px, py = o.polygon_pixels(vx, vy)
        order = np.lexsort((px, py))
        px, py = px[order], py[order]
        xs.append(px)
        ys.append(py)
        a = np.ones(len(px), np.uint8)
        a[(np.abs(px - cx) <= 1) & (np.abs(py - cy) <= 2)] = 0  # a small excluded area
        acts.append(a)
        offs.append(offs[-1] + len(px))
        cxs.append(cx)
        cys.append(cy)
    return (np.array(cxs, float), np.array(cys, float),
            (np.array(offs, np.int32), np.concatenate(xs), np.concatenate(ys), np.concatenate(acts)))


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_conformal_subsets_parity(ctx, mode):
    """Arbitrary-shape subsets (CSR pixel lists with excluded pixels), GN 6 params, Keys-4."""
    ref, de = synth.config1(seed=5)
    cx, cy, conf = conformal_set(512, 512, 60, 3)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    fg = run_gpu(ctx, p, ref, de, cx, cy, 0, mode, conformal=conf)
    fo = run_oracle(p, ref, de, cx, cy, 0, conformal=conf)
    compare(fg, fo, label="conformal")


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_fallback_and_statuses(ctx, mode):
    """GN failure -> simplex fallback, FOV failure, failed initial guess, flat subsets."""
    d = load("speckle_512.npz")
    cx = np.array([200.0, 5.0, 300.0, 150.0])
    cy = np.array([200.0, 200.0, 300.0, 150.0])
    f0 = dh.new_fields(cx, cy)
    f0[dh.F["SIGMA"], 2] = -1.0
    p = dh.default_params()
    p.max_solver_iterations_fast = 0
    for opt, expect0 in ((dh.GRADIENT_BASED, 11), (dh.GRADIENT_BASED_THEN_SIMPLEX, 1)):
        p.optimization_method = opt
        fg = run_gpu(ctx, p, d["ref"], d["defm"], cx, cy, 27, mode, fields=f0)
        fo = run_oracle(p, d["ref"], d["defm"], cx, cy, 27, fields=f0)
        assert fg[dh.F["STATUS"], 0] == expect0
        assert np.array_equal(fg[dh.F["STATUS"]], fo[o.F["STATUS"]])
        assert fg[dh.F["STATUS"], 1] == 12 and fg[dh.F["STATUS"], 2] == 5
    for value, status in ((0.0, 24), (100.0, 23)):
        flat = np.full((64, 64), value, np.float32)
        q = dh.default_params()
        q.optimization_method = dh.GRADIENT_BASED
        fg = run_gpu(ctx, q, flat, flat, [32.0], [32.0], 21, mode)
        assert fg[dh.F["STATUS"], 0] == status
        assert fg[dh.F["CONDITION"], 0] == -1.0


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_def_gradient_mode_parity(ctx, mode):
    """compute_def_gradients: GN with interpolated deformed-image gradients."""
    ref, de = synth.config1(seed=9)
    cx, cy = dh.regular_grid(512, 512, 31, 20)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = 1
    p.compute_def_gradients = 1
    p.compute_uncertainty_fields = 0
    fg = run_gpu(ctx, p, ref, de, cx, cy, 31, mode)
    fo = run_oracle(p, ref, de, cx, cy, 31)
    compare(fg, fo, label="defgrad")


def test_sequence_graph_matches_eager(ctx):
    """A 3-frame sequence replayed from a hipGraph equals frame-by-frame eager calls (twice)."""
    torch = pytest.importorskip("torch")
    sp = synth.Speckle(256, 256, seed=4)
    ref = sp.frame(None)
    frames = [sp.frame(synth.sinusoid_field(0.3 * k, period=128.0), noise_sigma=2.55, noise_seed=40 + k)
              for k in range(1, 4)]
    cx, cy = dh.regular_grid(256, 256, 21, 8)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_subsets(cx, cy, 21)
    # eager, frame by frame, host fields
    eager = []
    f = dh.new_fields(cx, cy)
    for k, fr in enumerate(frames):
        ctx.set_def_image(fr)
        ctx.correlate(f, mode=dh.MODE_BATCHED, first_frame=(k == 0))
        eager.append(f.copy())
    dev = torch.device("cuda:0")
    d_frames = [torch.from_numpy(fr).to(dev) for fr in frames]
    n = len(cx)
    for use_graph in (False, True, True):
        d_fields = torch.from_numpy(dh.new_fields(cx, cy)).to(dev)
        d_hist = torch.zeros((len(frames), dh.NUM_FIELDS, n), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.set_def_image_device(d_frames[0].data_ptr(), 256, 256, dh.STORAGE_F32)
        ctx.correlate_sequence([t.data_ptr() for t in d_frames], d_fields.data_ptr(), d_hist.data_ptr(),
                               mode=dh.MODE_BATCHED, use_graph=use_graph)
        ctx.synchronize()
        hist = d_hist.cpu().numpy()
        for k in range(len(frames)):
            assert np.array_equal(hist[k], eager[k]), "frame %d graph=%s" % (k, use_graph)


def test_autotune_keeps_results(ctx):
    """The autotuner only changes the launch shape.  Statuses and iterations are identical and values agree
    to float rounding: a subset whose warped window outgrows a smaller LDS slot on some iteration is solved
    by the workgroup kernel (FP64 per-pixel sums) instead of the wave (FP32 lane partials), which moves the
    last bits of its fields."""
    torch = pytest.importorskip("torch")
    ref, de = synth.config1(seed=2)
    cx, cy = dh.regular_grid(512, 512, 31, 10)
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(de)
    ctx.set_subsets(cx, cy, 31)
    f1 = ctx.correlate(dh.new_fields(cx, cy), mode=dh.MODE_BATCHED)
    d_f = torch.from_numpy(dh.new_fields(cx, cy)).to("cuda:0")
    torch.cuda.synchronize()
    ctx.autotune(d_f.data_ptr())
    shape = ctx.launch_shape()
    assert shape[0] >= 1
    f2 = ctx.correlate(dh.new_fields(cx, cy), mode=dh.MODE_BATCHED)
    for name in ("STATUS", "ITERATIONS"):
        assert np.array_equal(f1[dh.F[name]], f2[dh.F[name]]), name
    rel = np.abs(f1 - f2) / np.maximum(np.abs(f1), 1e-6)
    assert rel.max() < 1e-5 and np.mean(rel > 0) < 0.01
    assert np.array_equal(d_f.cpu().numpy(), dh.new_fields(cx, cy))
