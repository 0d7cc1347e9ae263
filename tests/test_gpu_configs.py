"""GPU parity at the full sizes of the benchmark workloads, through the C ABI, against the CPU oracle
(double build) on the same seeded inputs.

* 2048^2 speckle, 41 x 41 subsets at step 5 (155,236 subsets), sinusoidal field, two frames of the
  hipGraph sequence (frame 2 starts from frame 1's fields), 6-parameter Gauss-Newton, Keys-4: every
  31st subset of each frame against the oracle started from the same fields.
* 1024^2 speckle, 5,000 convex polygon subsets of 200-3000 pixels, two frames of rigid + affine motion,
  initial guesses = truth + U(+-2 px), at most 40 Gauss-Newton iterations, the default launch shape (so
  the largest subsets overflow the BATCHED wave's LDS slot and the PLAIN kernel takes them over): every
  subset against the oracle.
* 1536^2 speckle under a quadratic field, 21 x 21 subsets at step 6 (62,500 subsets): the forward gamma
  against the oracle's Objective::gamma on a 2,000-subset sample, and the adjoint against central finite
  differences of the forward over all subsets (dL/dp) and at sampled pixels (dL/dI_def).

Bar: status identical on >= 99.9 % of subsets, u / v / theta / strains within 1e-4, gamma within 1e-5
relative (on the polygon subsets: on 99.9 % of them, and on all within the reference's own float-build
deviation from its double build, tests/gpu_common.py); adjoint within 1e-3 relative of the central
differences."""
import numpy as np
import pytest

import dice_amd as dh
import oracle_py as o
from dice_amd import synth
from gpu_common import compare, csr_take, run_oracle, to_oracle

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MODES = [dh.MODE_BATCHED, dh.MODE_PLAIN]
MODE_NAMES = {dh.MODE_PLAIN: "plain", dh.MODE_BATCHED: "batched"}


def gn6_params(max_fast=None):
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.interpolation_method = dh.KEYS_FOURTH
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    if max_fast is not None:
        p.max_solver_iterations_fast = max_fast
    return p


@pytest.fixture(scope="module")
def c2_scene():
    ref, frames = synth.sinusoid_sequence(size=2048, frames=2, seed=1)
    cx, cy = dh.regular_grid(2048, 2048, 41, 5)
    assert len(cx) == 155236
    return ref, frames, cx, cy


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_c2_sequence_against_oracle(c2_scene, mode):
    ref, frames, cx, cy = c2_scene
    n = len(cx)
    p = gn6_params()
    dev = torch.device("cuda:0")
    with dh.Context(0) as ctx:
        ctx.set_params(p)
        ctx.set_ref_image(ref)
        ctx.set_subsets(cx, cy, 41)
        d_frames = [torch.from_numpy(f).to(dev) for f in frames]
        d_fields = torch.from_numpy(dh.new_fields(cx, cy)).to(dev)
        d_hist = torch.zeros((len(frames), dh.NUM_FIELDS, n), dtype=torch.float64, device=dev)
        torch.cuda.synchronize()
        ctx.set_def_image_device(d_frames[0].data_ptr(), 2048, 2048, dh.STORAGE_F32)
        ctx.correlate_sequence([t.data_ptr() for t in d_frames], d_fields.data_ptr(), d_hist.data_ptr(), mode=mode,
                               use_graph=True)
        ctx.synchronize()
        hist = d_hist.cpu().numpy()
    sample = np.arange(0, n, 31)
    start = o.new_fields(cx[sample], cy[sample])
    for k in range(len(frames)):
        fg = np.ascontiguousarray(hist[k][:, sample])
        fo = run_oracle(p, ref, frames[k], cx[sample], cy[sample], 41, fields=start, first_frame=(k == 0))
        agree, both = compare(fg, fo, label="2048^2 frame %d" % (k + 1))
        assert both.mean() > 0.99
        # the epilogue of the converged subsets: sigma, noise, contrast.  noise_std_dev shifts its integer
        # window by int += scalar_t (DICe_Subset.cpp:654-656), which truncates: where u or v sits within
        # rounding of an integer (the sinusoid's zero crossings) float and double can take windows one
        # pixel apart, so those subsets are left out of the sigma and noise comparison
        u, v = fo[o.F["DISP_X"]], fo[o.F["DISP_Y"]]
        clear = both & (np.abs(u - np.round(u)) > 1e-4) & (np.abs(v - np.round(v)) > 1e-4)
        assert clear.sum() > 0.95 * both.sum()
        for name in ("SIGMA", "NOISE", "CONTRAST"):
            sel = both if name == "CONTRAST" else clear
            rel = np.abs(fg[dh.F[name]] - fo[o.F[name]])[sel] / np.abs(fo[o.F[name]][sel])
            assert rel.max() < 1e-4, (k, name, rel.max())
        # the next frame starts from this frame's GPU fields, as in the sequence
        start = np.ascontiguousarray(hist[k][:, sample])
    # the field itself: A = 0.2 px in frame 2
    conv = np.isin(hist[1][dh.F["STATUS"]], dh.CONVERGED_STATUSES)
    ut = 0.2 * np.sin(2 * np.pi * cx / 256.0)
    assert np.abs(hist[1][dh.F["DISP_X"]] - ut)[conv].mean() < 0.01


@pytest.fixture(scope="module")
def c3_scene():
    size, frames = 1024, 2
    ref, seq, motion, subs = synth.config3(n_subsets=5000, frames=frames, size=size)
    cx, cy, off, px, py, act = subs
    assert len(cx) == 5000
    cnt = np.diff(off)
    assert cnt.min() >= 200 and cnt.max() <= 3000 and cnt.max() > 2500
    c = (size - 1) / 2.0
    rng = np.random.default_rng(21)
    inits = []
    for F, t in motion:
        d = F - np.eye(2)
        f = dh.new_fields(cx, cy)
        f[dh.F["DISP_X"]] = d[0, 0] * (cx - c) + d[0, 1] * (cy - c) + t[0] + rng.uniform(-2.0, 2.0, len(cx))
        f[dh.F["DISP_Y"]] = d[1, 0] * (cx - c) + d[1, 1] * (cy - c) + t[1] + rng.uniform(-2.0, 2.0, len(cx))
        inits.append(f)
    return ref, seq, (off, px, py, act), cx, cy, inits


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_c3_uneven_conformal_against_oracle(c3_scene, mode):
    ref, seq, conf, cx, cy, inits = c3_scene
    p = gn6_params(max_fast=40)
    with dh.Context(0) as ctx:
        ctx.set_params(p)
        ctx.set_ref_image(ref)
        ctx.set_subsets(cx, cy, 0, conformal=conf)
        for k, de in enumerate(seq):
            ctx.set_def_image(de)
            fg = ctx.correlate(inits[k].copy(), mode=mode, first_frame=False)
            fb, ov0, ov1 = ctx.pass_counts()
            if mode == dh.MODE_BATCHED:
                # the largest polygons do not fit the default 2,560-float slot: the PLAIN kernel solves them
                assert ov0 > 0, "no subset overflowed the LDS slot"
            fo = run_oracle(p, ref, de, cx, cy, 0, fields=inits[k], conformal=conf, first_frame=False)
            ff = run_oracle(p, ref, de, cx, cy, 0, fields=inits[k], conformal=conf, first_frame=False,
                            precision="float")
            agree, both = compare(fg, fo, label="conformal frame %d" % (k + 1), ff=ff)
            assert both.mean() > 0.99
            it_g, it_o = fg[dh.F["ITERATIONS"]][both], fo[o.F["ITERATIONS"]][both]
            assert np.mean(it_g == it_o) >= 0.99
            assert it_o.max() - it_o.min() >= 4  # the uneven iteration counts this workload is about


def test_c3_overflow_handoff_matches_in_slot_solve(c3_scene):
    """The same subsets solved with a slot too small for most of them (overflow to the PLAIN kernel) and
    with one large enough for all: identical statuses, values within float rounding."""
    ref, seq, conf, cx, cy, inits = c3_scene
    idx = np.arange(0, len(cx), 10)
    sub = csr_take(conf, idx)
    p = gn6_params(max_fast=40)
    out = []
    for slot in (None, 16384):
        with dh.Context(0) as ctx:
            ctx.set_params(p)
            ctx.set_ref_image(ref)
            ctx.set_subsets(cx[idx], cy[idx], 0, conformal=sub)
            if slot:
                ctx.set_launch_shape(1, 4, slot)
            ctx.set_def_image(seq[0])
            out.append(ctx.correlate(np.ascontiguousarray(inits[0][:, idx]), mode=dh.MODE_BATCHED, first_frame=False))
            out.append(ctx.pass_counts()[1])
    # a subset whose warp runs away mid-solve can outgrow even the large slot
    assert out[1] > 0 and out[3] <= out[1] // 20
    a, b = out[0], out[2]
    assert np.array_equal(a[dh.F["STATUS"]], b[dh.F["STATUS"]])
    conv = np.isin(a[dh.F["STATUS"]], dh.CONVERGED_STATUSES)
    for name in ("DISP_X", "DISP_Y", "ROT_Z"):
        assert np.abs(a[dh.F[name]] - b[dh.F[name]])[conv].max() < 1e-4, name


def params_from_field(field, cx, cy):
    u, v, ux, uy, vx, vy = field(cx, cy)
    F = np.stack([np.stack([1 + ux, uy], -1), np.stack([vx, 1 + vy], -1)], -2)
    th, ex, ey, g = synth.polar_params(F)
    return np.stack([u, v, th, ex, ey, g], axis=1)


@pytest.fixture(scope="module")
def c4_scene():
    ref, de, field = synth.config4(size=1536, peak=1.5, seed=1)
    cx, cy = dh.regular_grid(1536, 1536, 21, 6)
    assert len(cx) == 62500
    rng = np.random.default_rng(3)
    P = params_from_field(field, cx, cy)
    P += rng.normal(0, 1, P.shape) * np.array([0.3, 0.3, 0.01, 0.01, 0.01, 0.01])
    ctx = dh.Context(0)
    p = dh.default_params()
    p.interpolation_method = dh.KEYS_FOURTH
    p.compute_def_gradients = 0
    # dice_hip_gamma evaluates all six parameters; the oracle's objective reads the enabled ones
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    ctx.set_params(p)
    ctx.set_ref_image(ref)
    ctx.set_def_image(de)
    ctx.set_subsets(cx, cy, 21)
    yield ctx, p, ref, de, cx, cy, P
    ctx.close()


def _gamma(ctx, P, mode):
    d = torch.from_numpy(np.ascontiguousarray(P, dtype=np.float64)).cuda()
    g = torch.empty(len(P), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    ctx.gamma(d.data_ptr(), g.data_ptr(), mode)
    ctx.synchronize()
    return g.cpu().numpy()


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_c4_forward_gamma_against_oracle(c4_scene, mode):
    ctx, p, ref, de, cx, cy, P = c4_scene
    g = _gamma(ctx, P, mode)
    assert np.all(g > 0)
    q = to_oracle(p)
    ref64, de64 = ref.astype(np.float64), de.astype(np.float64)
    sample = np.random.default_rng(5).choice(len(cx), 2000, replace=False)
    go = np.array([o.objective(q, ref64, de64, 21, cx[i], cy[i], 0, P[i])[1][0] for i in sample])
    rel = np.abs(g[sample] - go) / np.abs(go)
    assert rel.max() <= 1e-5, rel.max()


@pytest.mark.parametrize("mode", MODES, ids=lambda m: MODE_NAMES[m])
def test_c4_adjoint_against_central_differences(c4_scene, mode):
    ctx, p, ref, de, cx, cy, P = c4_scene
    rng = np.random.default_rng(7)
    w = rng.uniform(0.5, 1.5, len(P))
    d = torch.from_numpy(np.ascontiguousarray(P)).cuda()
    dg = torch.from_numpy(w).cuda()
    dp = torch.empty_like(d)
    dimg = torch.zeros(de.shape, dtype=torch.float32, device="cuda")
    torch.cuda.synchronize()
    ctx.gamma_adjoint(d.data_ptr(), dg.data_ptr(), dp.data_ptr(), dimg.data_ptr(), mode)
    ctx.synchronize()
    dp, dimg = dp.cpu().numpy(), dimg.cpu().numpy()
    # dL/dp: central differences of every subset at once (the subsets are independent)
    for k, h in enumerate([1e-2, 1e-2, 1e-3, 1e-3, 1e-3, 1e-3]):
        Pp, Pm = P.copy(), P.copy()
        Pp[:, k] += h
        Pm[:, k] -= h
        fd = w * (_gamma(ctx, Pp, mode) - _gamma(ctx, Pm, mode)) / (2 * h)
        err = np.linalg.norm(dp[:, k] - fd) / np.linalg.norm(fd)
        assert err < 1e-3, (k, err)
    # dL/dI_def at pixels under subsets
    pick = rng.choice(len(cx), 16, replace=False)
    fd, an = [], []
    for i in pick:
        yy, xx = int(cy[i]) + int(rng.integers(-8, 9)), int(cx[i]) + int(rng.integers(-8, 9))
        hp, hm = de.copy(), de.copy()
        hp[yy, xx] += 0.5
        hm[yy, xx] -= 0.5
        ctx.set_def_image(hp)
        gp = _gamma(ctx, P, mode)
        ctx.set_def_image(hm)
        gm = _gamma(ctx, P, mode)
        fd.append(np.sum(w * (gp - gm)) / 1.0)
        an.append(dimg[yy, xx])
    ctx.set_def_image(de)
    fd, an = np.array(fd), np.array(an)
    assert np.linalg.norm(an - fd) / np.linalg.norm(fd) < 1e-3, (an, fd)
