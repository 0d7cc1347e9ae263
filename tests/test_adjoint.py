"""Differentiation path through the C ABI: dice_hip_gamma (ZNSSD gamma of every subset at given affine
parameters) and dice_hip_gamma_adjoint (dL/dp and dL/dI_def).

Checked against central finite differences of the forward, against the Gauss-Newton gradient direction
of computeUpdateFast (DICe_Objective.cpp:500-546, built here in numpy from the oracle's interpolant and
the deformed image's FD gradients), and through an outer least-gamma fit of a 12-coefficient quadratic
displacement field compared with the per-subset Gauss-Newton solution.  The torch.autograd.Function
lives only in this harness; the library itself has no torch types."""
import math

import numpy as np
import pytest

import dice_amd as dh
from dice_amd import synth
import oracle_py as o

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

MODES = [dh.MODE_BATCHED, dh.MODE_PLAIN]
SIZE, S, STEP = 512, 21, 24


@pytest.fixture(scope="module")
def scene():
    ref, de, field = synth.config4(size=SIZE, peak=1.5, seed=1)
    cx, cy = dh.regular_grid(SIZE, SIZE, S, STEP)
    return ref, de, field, cx, cy


def params_from_field(field, cx, cy):
    """Affine shape parameters of the field at the centres: (u, v) and the polar split of F = I + grad u
    into R(theta) [[1 + ex, g], [g, 1 + ey]]."""
    u, v, ux, uy, vx, vy = field(cx, cy)
    f11, f12, f21, f22 = 1 + ux, uy, vx, 1 + vy
    th = np.arctan2(f21 - f12, f11 + f22)
    c, s = np.cos(th), np.sin(th)
    u11 = c * f11 + s * f21
    u12 = c * f12 + s * f22
    u22 = -s * f12 + c * f22
    return np.stack([u, v, th, u11 - 1, u22 - 1, u12], axis=1)


def make_ctx(ref, de, cx, cy, interp=dh.KEYS_FOURTH):
    c = dh.Context(0)
    p = dh.default_params()
    p.interpolation_method = interp
    p.compute_def_gradients = 0
    c.set_params(p)
    c.set_ref_image(ref)
    c.set_def_image(de)
    c.set_subsets(cx, cy, S)
    return c


def gamma(c, P, mode):
    d = torch.from_numpy(np.ascontiguousarray(P, dtype=np.float64)).cuda()
    g = torch.empty(len(P), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    c.gamma(d.data_ptr(), g.data_ptr(), mode)
    c.synchronize()
    return g.cpu().numpy()


def adjoint(c, P, w, mode, shape=None):
    d = torch.from_numpy(np.ascontiguousarray(P, dtype=np.float64)).cuda()
    dg = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float64)).cuda()
    dp = torch.empty_like(d)
    img = torch.zeros(shape, dtype=torch.float32, device="cuda") if shape else None
    torch.cuda.synchronize()
    c.gamma_adjoint(d.data_ptr(), dg.data_ptr(), dp.data_ptr(), img.data_ptr() if img is not None else 0, mode)
    c.synchronize()
    return dp.cpu().numpy(), (img.cpu().numpy() if img is not None else None)


@pytest.mark.parametrize("mode", MODES, ids=["batched", "plain"])
@pytest.mark.parametrize("interp", [dh.KEYS_FOURTH, dh.BILINEAR], ids=["keys", "bilinear"])
def test_adjoint_params_match_central_differences(scene, mode, interp):
    ref, de, field, cx, cy = scene
    rng = np.random.default_rng(7)
    P = params_from_field(field, cx, cy)
    P += rng.normal(0, 1, P.shape) * np.array([0.3, 0.3, 0.01, 0.01, 0.01, 0.01])  # away from the optimum
    w = rng.uniform(0.5, 1.5, len(P))
    c = make_ctx(ref, de, cx, cy, interp)
    g0 = gamma(c, P, mode)
    assert np.all(g0 > 0)
    dp, _ = adjoint(c, P, w, mode)
    # central differences, all subsets at once (they are independent); steps move pixels by ~1e-2 px
    steps = [1e-2, 1e-2, 1e-3, 1e-3, 1e-3, 1e-3]
    for k, h in enumerate(steps):
        Pp, Pm = P.copy(), P.copy()
        Pp[:, k] += h
        Pm[:, k] -= h
        fd = w * (gamma(c, Pp, mode) - gamma(c, Pm, mode)) / (2 * h)
        err = np.linalg.norm(dp[:, k] - fd) / np.linalg.norm(fd)
        # bilinear is piecewise linear: its one-sided slopes differ at integer positions, which the
        # central steps straddle for a share of the pixels
        tol = 1e-3 if interp == dh.KEYS_FOURTH else 1.5e-2
        assert err < tol, (k, err)
    c.close()


@pytest.mark.parametrize("mode", MODES, ids=["batched", "plain"])
def test_adjoint_image_matches_central_differences(scene, mode):
    ref, de, field, cx, cy = scene
    rng = np.random.default_rng(8)
    P = params_from_field(field, cx, cy)
    P[:, :2] += rng.normal(0, 0.2, (len(P), 2))
    w = rng.uniform(0.5, 1.5, len(P))
    c = make_ctx(ref, de, cx, cy)
    _, dimg = adjoint(c, P, w, mode, shape=de.shape)
    # pixels under subsets: the centres' neighbourhoods
    pick = rng.choice(len(cx), 24, replace=False)
    pts = [(int(cy[i]) + int(rng.integers(-8, 9)), int(cx[i]) + int(rng.integers(-8, 9))) for i in pick]
    h = 0.5
    fd, an = [], []
    for (yy, xx) in pts:
        dp_img, dm_img = de.copy(), de.copy()
        dp_img[yy, xx] += h
        dm_img[yy, xx] -= h
        c.set_def_image(dp_img)
        gp = gamma(c, P, mode)
        c.set_def_image(dm_img)
        gm = gamma(c, P, mode)
        fd.append(np.sum(w * (gp - gm)) / (2 * h))
        an.append(dimg[yy, xx])
    fd, an = np.array(fd), np.array(an)
    assert np.linalg.norm(an - fd) / np.linalg.norm(fd) < 1e-3, (an, fd)
    c.close()


@pytest.mark.parametrize("strain", [0.0, 0.6], ids=["classes", "atomics"])
def test_adjoint_image_batched_matches_plain(scene, strain):
    """The whole dL/dI_def image of the BATCHED adjoint against the PLAIN kernel's (global atomics per tap).
    strain 0: every subset takes the class-by-class LDS read-modify-write scatter; strain 0.6: the warp
    moves pixels too far for that (||M - I|| > 0.45) and the scatter uses LDS atomics.  Only the order
    of the float sums differs."""
    ref, de, field, cx, cy = scene
    rng = np.random.default_rng(10)
    P = params_from_field(field, cx, cy)
    P[:, :2] += rng.normal(0, 0.2, (len(P), 2))
    P[:, 3] += strain
    w = rng.uniform(0.5, 1.5, len(P))
    c = make_ctx(ref, de, cx, cy)
    dpb, ib = adjoint(c, P, w, dh.MODE_BATCHED, shape=de.shape)
    dpp, ip = adjoint(c, P, w, dh.MODE_PLAIN, shape=de.shape)
    assert np.count_nonzero(ib) > 0.5 * len(cx) * S * S
    assert np.linalg.norm(ib - ip) / np.linalg.norm(ip) < 1e-5
    assert np.abs(ib - ip).max() <= 1e-4 * np.abs(ip).max()
    assert np.linalg.norm(dpb - dpp) / np.linalg.norm(dpp) < 1e-6
    c.close()


def test_adjoint_image_dense_overlap_batched_matches_plain():
    """Windows that overlap as in the 1536^2 differentiation workload (21 x 21 subsets every 6 px, so
    about 20 subsets add into each pixel, from waves on every XCD): the BATCHED image gradient (LDS
    accumulator, one device-scope atomic per window cell) against the PLAIN kernel's (one atomic per
    tap).  A lost update would move a pixel by a whole subset's contribution."""
    ref, de, field = synth.config4(size=SIZE, peak=1.5, seed=2)
    cx, cy = dh.regular_grid(SIZE, SIZE, S, 6)
    rng = np.random.default_rng(11)
    P = params_from_field(field, cx, cy)
    P[:, :2] += rng.normal(0, 0.2, (len(P), 2))
    w = rng.uniform(0.5, 1.5, len(P))
    c = make_ctx(ref, de, cx, cy)
    dpb, ib = adjoint(c, P, w, dh.MODE_BATCHED, shape=de.shape)
    dpp, ip = adjoint(c, P, w, dh.MODE_PLAIN, shape=de.shape)
    assert len(cx) > 6000
    assert np.count_nonzero(ip) > 0.8 * ip.size
    assert np.linalg.norm(ib - ip) / np.linalg.norm(ip) < 1e-5
    assert np.abs(ib - ip).max() <= 1e-4 * np.abs(ip).max()
    assert np.linalg.norm(dpb - dpp) / np.linalg.norm(dpp) < 1e-6
    c.close()


def test_adjoint_direction_matches_gauss_newton_gradient(scene):
    """Near the optimum the ZNSSD gradient points along the Gauss-Newton gradient q of computeUpdateFast
    (ZSSD residual, the deformed image's interpolated FD gradients): dgamma/dp ~ (2 / |d - dbar|^2) q.
    Its magnitude is checked against the exact ZNSSD gradient (2 / nd) sum (rho dhat - rhat) r_j built
    from the same interpolated FD gradients; the adjoint differs from both only by using the derivative
    of the interpolant itself."""
    ref, de, field, cx, cy = scene
    c = make_ctx(ref, de, cx, cy)
    # the per-subset GN solution is the optimum
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.compute_uncertainty_fields = 0
    c.set_params(p)
    f = dh.new_fields(cx, cy)
    P0 = params_from_field(field, cx, cy)
    for k, name in enumerate(["DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY"]):
        f[dh.F[name]] = P0[:, k]
    f = c.correlate(f, mode=dh.MODE_BATCHED)
    ok = f[dh.F["STATUS"]] == 1
    Popt = np.stack([f[dh.F[n]] for n in ["DISP_X", "DISP_Y", "ROT_Z", "NS_XX", "NS_YY", "SS_XY"]], axis=1)
    rng = np.random.default_rng(9)
    P = Popt + rng.normal(0, 1, Popt.shape) * np.array([0.1, 0.1, 0.003, 0.003, 0.003, 0.003])
    p.optimization_method = dh.GRADIENT_BASED_THEN_SIMPLEX
    c.set_params(p)
    dp, _ = adjoint(c, P, np.ones(len(P)), dh.MODE_BATCHED)
    gx, gy = o.fd_gradients(de.astype(np.float64))
    cos_gn, cos_ex, ratio = [], [], []
    for i in np.where(ok)[0][:40]:
        h = S // 2
        ys, xs = np.mgrid[int(cy[i]) - h:int(cy[i]) + h + 1, int(cx[i]) - h:int(cx[i]) + h + 1]
        dx, dy = (xs - cx[i]).ravel(), (ys - cy[i]).ravel()
        u, v, th, ex, ey, g = P[i]
        cs, sn = math.cos(th), math.sin(th)
        Dx, Dy = (1 + ex) * dx + g * dy, (1 + ey) * dy + g * dx
        X = cx[i] + u + cs * Dx - sn * Dy
        Y = cy[i] + v + sn * Dx + cs * Dy
        d = o.interp(2, de.astype(np.float64), X, Y)
        Gx = o.interp(2, gx, X, Y)
        Gy = o.interp(2, gy, X, Y)
        r = ref[ys, xs].ravel().astype(np.float64)
        gmf = (d - d.mean()) - (r - r.mean())
        rows = np.stack([Gx, Gy, Gx * (-sn * Dx - cs * Dy) + Gy * (cs * Dx - sn * Dy), (Gx * cs + Gy * sn) * dx,
                         (-Gx * sn + Gy * cs) * dy, Gx * (cs * dy - sn * dx) + Gy * (sn * dy + cs * dx)], axis=1)
        q = rows.T @ gmf
        nd = np.sqrt(np.sum((d - d.mean()) ** 2))
        nr = np.sqrt(np.sum((r - r.mean()) ** 2))
        rho = np.sum((d - d.mean()) * (r - r.mean())) / (nd * nr)
        gn_grad = 2.0 * q / nd ** 2
        exact = (2.0 / nd) * rows.T @ (rho * (d - d.mean()) / nd - (r - r.mean()) / nr)
        cos_gn.append(np.dot(gn_grad, dp[i]) / (np.linalg.norm(gn_grad) * np.linalg.norm(dp[i])))
        cos_ex.append(np.dot(exact, dp[i]) / (np.linalg.norm(exact) * np.linalg.norm(dp[i])))
        ratio.append(np.linalg.norm(dp[i]) / np.linalg.norm(exact))
    print("cos vs GN q: median %.4f min %.4f; vs exact: median %.4f min %.4f; |adjoint|/|exact| %.3f..%.3f"
          % (np.median(cos_gn), min(cos_gn), np.median(cos_ex), min(cos_ex), min(ratio), max(ratio)))
    # the interpolated 4th-order FD image gradients differ from the interpolant's own derivative by a few
    # per cent, which is what separates these from the adjoint (the exact checks are the central
    # differences above)
    assert np.median(cos_gn) > 0.98 and min(cos_gn) > 0.9
    assert np.median(cos_ex) > 0.99 and min(cos_ex) > 0.93
    assert 0.85 < min(ratio) and max(ratio) < 1.2 and 0.95 < np.median(ratio) < 1.05
    c.close()


class ZnssdGamma(torch.autograd.Function):
    """gamma(p) per subset through dice_hip_gamma; backward through dice_hip_gamma_adjoint."""

    @staticmethod
    def forward(fctx, params, c):
        params = params.detach().contiguous()
        g = torch.empty(params.shape[0], dtype=torch.float64, device=params.device)
        torch.cuda.synchronize()
        c.gamma(params.data_ptr(), g.data_ptr(), dh.MODE_BATCHED)
        c.synchronize()
        fctx.save_for_backward(params)
        fctx.c = c
        return g

    @staticmethod
    def backward(fctx, dg):
        (params,) = fctx.saved_tensors
        dg = dg.contiguous()
        dp = torch.empty_like(params)
        torch.cuda.synchronize()
        fctx.c.gamma_adjoint(params.data_ptr(), dg.data_ptr(), dp.data_ptr(), 0, dh.MODE_BATCHED)
        fctx.c.synchronize()
        return dp, None


def test_outer_fit_of_quadratic_field_matches_gauss_newton(scene):
    """Least-gamma fit of the 12 field coefficients through the adjoint, against the per-subset GN."""
    ref, de, field, cx, cy = scene
    c = make_ctx(ref, de, cx, cy)
    half = (SIZE - 1) / 2.0
    X = torch.tensor((cx - half) / half, dtype=torch.float64, device="cuda")
    Y = torch.tensor((cy - half) / half, dtype=torch.float64, device="cuda")
    a = torch.zeros(12, dtype=torch.float64, device="cuda", requires_grad=True)

    def shape_params(a):
        u = a[0] + a[1] * X + a[2] * Y + a[3] * X * X + a[4] * X * Y + a[5] * Y * Y
        v = a[6] + a[7] * X + a[8] * Y + a[9] * X * X + a[10] * X * Y + a[11] * Y * Y
        ux = (a[1] + 2 * a[3] * X + a[4] * Y) / half
        uy = (a[2] + a[4] * X + 2 * a[5] * Y) / half
        vx = (a[7] + 2 * a[9] * X + a[10] * Y) / half
        vy = (a[8] + a[10] * X + 2 * a[11] * Y) / half
        f11, f12, f21, f22 = 1 + ux, uy, vx, 1 + vy
        th = torch.atan2(f21 - f12, f11 + f22)
        cs, sn = torch.cos(th), torch.sin(th)
        u11, u12, u22 = cs * f11 + sn * f21, cs * f12 + sn * f22, -sn * f12 + cs * f22
        return torch.stack([u, v, th, u11 - 1, u22 - 1, u12], dim=1)

    opt = torch.optim.LBFGS([a], lr=1.0, max_iter=60, line_search_fn="strong_wolfe")

    def closure():
        opt.zero_grad()
        loss = ZnssdGamma.apply(shape_params(a), c).sum()
        loss.backward()
        return loss

    for _ in range(4):
        opt.step(closure)
    fit = shape_params(a).detach().cpu().numpy()
    # the plain per-subset Gauss-Newton solution from a zero initial guess
    p = dh.default_params()
    p.optimization_method = dh.GRADIENT_BASED
    p.enable_rotation = p.enable_normal_strain = p.enable_shear_strain = 1
    p.compute_uncertainty_fields = 0
    c.set_params(p)
    f = c.correlate(dh.new_fields(cx, cy), mode=dh.MODE_PLAIN)
    ok = f[dh.F["STATUS"]] == 1
    assert ok.mean() > 0.95
    du = np.abs(fit[:, 0] - f[dh.F["DISP_X"]])[ok]
    dv = np.abs(fit[:, 1] - f[dh.F["DISP_Y"]])[ok]
    truth = params_from_field(field, cx, cy)
    print("fit vs GN: mean |du| %.4f |dv| %.4f; fit vs truth %.4f" %
          (du.mean(), dv.mean(), np.abs(fit[:, 0] - truth[:, 0]).mean()))
    assert du.mean() < 0.02 and dv.mean() < 0.02
    c.close()
