"""Seeded synthetic speckle workloads (reference image + analytically deformed frames).

Speckle pattern: a Poisson number (mean density * W * H) of Gaussian blobs with
uniform centres, radius r ~ U(r_lo, r_hi) (sigma = r / 2) and amplitude
~ U(0, 255), summed and linearly rescaled to [0, 255] (the rescale of the
reference frame is applied to every frame).  A deformed frame moves each blob
centre by the Lagrangian displacement u(X) and maps its covariance by
F = I + grad u(X) evaluated at the centre, so intensity is carried with the
material point (exact for affine fields, first order otherwise).  Gaussian
noise of the given sigma is added per frame with its own seed.

Rendering runs in libdice_speckle.so (C++/OpenMP).
"""
import ctypes
import os

import numpy as np

from . import LIB_DIR

_rl = None


def _render_lib():
    global _rl
    if _rl is None:
        path = os.path.join(LIB_DIR, "libdice_speckle.so")
        if not os.path.exists(path):
            raise RuntimeError("libdice_speckle.so not built (run __graft_entry__.build())")
        _rl = ctypes.CDLL(path)
        P = ctypes.c_void_p
        _rl.dice_render_blobs.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, P, P, P, P, P,
                                          ctypes.c_int]
    return _rl


def _ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


class Speckle:
    """A seeded blob population over a W x H reference image."""

    def __init__(self, width, height, density=0.5, r_lo=2.0, r_hi=4.0, seed=1, margin=8):
        rng = np.random.default_rng(seed)
        area = (width + 2 * margin) * (height + 2 * margin)
        nb = rng.poisson(density * area)
        self.w, self.h = width, height
        self.cx = rng.uniform(-margin, width + margin, nb)
        self.cy = rng.uniform(-margin, height + margin, nb)
        r = rng.uniform(r_lo, r_hi, nb)
        self.sigma = r / 2.0
        self.amp = rng.uniform(0.0, 255.0, nb)
        self.lo = None
        self.hi = None

    def _render(self, cx, cy, ixx, ixy, iyy, radius, threads):
        order = np.argsort(cy - radius, kind="stable")
        arrs = [np.ascontiguousarray(a[order], dtype=np.float64) for a in (cx, cy, ixx, ixy, iyy, self.amp, radius)]
        out = np.zeros((self.h, self.w), dtype=np.float32)
        rc = _render_lib().dice_render_blobs(self.w, self.h, len(order), *[_ptr(a) for a in arrs], _ptr(out),
                                             threads)
        if rc != 0:
            raise RuntimeError("render failed")
        return out

    def frame(self, disp=None, noise_sigma=0.0, noise_seed=2, threads=None):
        """Render a frame.  disp(X, Y) -> (u, v, dudx, dudy, dvdx, dvdy) at blob centres
        (None = reference).  Returns float32 intensities, rescaled with the reference's range."""
        threads = threads or max(1, min(16, os.cpu_count() or 1))
        s2 = self.sigma ** 2
        if disp is None:
            cx, cy = self.cx, self.cy
            ixx = 1.0 / s2
            ixy = np.zeros_like(s2)
            iyy = 1.0 / s2
            radius = 5.0 * self.sigma
        else:
            u, v, dudx, dudy, dvdx, dvdy = disp(self.cx, self.cy)
            cx, cy = self.cx + u, self.cy + v
            # covariance F S F^T with S = s2 I  ->  inverse = (F F^T)^-1 / s2
            f11, f12, f21, f22 = 1.0 + dudx, dudy, dvdx, 1.0 + dvdy
            c11 = (f11 * f11 + f12 * f12) * s2
            c12 = (f11 * f21 + f12 * f22) * s2
            c22 = (f21 * f21 + f22 * f22) * s2
            det = c11 * c22 - c12 * c12
            ixx, ixy, iyy = c22 / det, -c12 / det, c11 / det
            radius = 5.0 * self.sigma * np.maximum(np.abs(f11) + np.abs(f12), np.abs(f21) + np.abs(f22))
        img = self._render(cx, cy, ixx, ixy, iyy, radius, threads)
        if self.lo is None:
            ref = img if disp is None else self.frame(None, threads=threads)
            self.lo, self.hi = float(ref.min()), float(ref.max())
        img = (img - self.lo) * (255.0 / (self.hi - self.lo))
        if noise_sigma > 0:
            img = img + np.random.default_rng(noise_seed).normal(0.0, noise_sigma, img.shape)
        return img.astype(np.float32)


def affine_field(u0, v0, dudx, dudy, dvdx, dvdy):
    """u = u0 + dudx X + dudy Y, v = v0 + dvdx X + dvdy Y (Lagrangian)."""
    def f(X, Y):
        return (u0 + dudx * X + dudy * Y, v0 + dvdx * X + dvdy * Y,
                np.full_like(X, dudx), np.full_like(X, dudy), np.full_like(X, dvdx), np.full_like(X, dvdy))
    return f


def sinusoid_field(amp, period=256.0):
    """u = A sin(2 pi X / P), v = 0.5 A cos(2 pi Y / P)."""
    k = 2.0 * np.pi / period

    def f(X, Y):
        return (amp * np.sin(k * X), 0.5 * amp * np.cos(k * Y), amp * k * np.cos(k * X), np.zeros_like(X),
                np.zeros_like(X), -0.5 * amp * k * np.sin(k * Y))
    return f


def quadratic_field(coef, cx=0.0, cy=0.0, scale=1.0):
    """u, v quadratic in the normalised coordinates x = (X - cx) / scale, y = (Y - cy) / scale:
    u = c0 + c1 x + c2 y + c3 x^2 + c4 x y + c5 y^2 and v with c6..c11 (12 coefficients)."""
    c = np.asarray(coef, dtype=np.float64)

    def f(X, Y):
        x, y = (X - cx) / scale, (Y - cy) / scale
        u = c[0] + c[1] * x + c[2] * y + c[3] * x * x + c[4] * x * y + c[5] * y * y
        v = c[6] + c[7] * x + c[8] * y + c[9] * x * x + c[10] * x * y + c[11] * y * y
        dudx = (c[1] + 2 * c[3] * x + c[4] * y) / scale
        dudy = (c[2] + c[4] * x + 2 * c[5] * y) / scale
        dvdx = (c[7] + 2 * c[9] * x + c[10] * y) / scale
        dvdy = (c[8] + c[10] * x + 2 * c[11] * y) / scale
        return u, v, dudx, dudy, dvdx, dvdy
    return f


def quadratic_coefficients(peak=1.5, seed=4):
    """12 random coefficients scaled so that max |(u, v)| over the unit square [-1, 1]^2 is `peak`."""
    c = np.random.default_rng(seed).normal(size=12)
    g = np.linspace(-1, 1, 65)
    x, y = np.meshgrid(g, g)
    u, v = quadratic_field(c)(x, y)[:2]
    return c * (peak / max(np.abs(u).max(), np.abs(v).max()))


def config4(size=1536, peak=1.5, seed=1, noise_sigma=2.55):
    """size^2 speckle deformed by a quadratic field with 12 random coefficients (peak |u|, |v| = peak px):
    returns ref, def and the field function of pixel coordinates."""
    sp = Speckle(size, size, seed=seed)
    ref = sp.frame(None)
    half = (size - 1) / 2.0
    field = quadratic_field(quadratic_coefficients(peak, seed + 3), half, half, half)
    de = sp.frame(field, noise_sigma=noise_sigma, noise_seed=seed + 1)
    return ref, de, field


def polar_params(F):
    """Affine shape parameters (theta, ex, ey, gxy) of a deformation gradient F = R(theta) [[1+ex, g], [g, 1+ey]]
    (the parameterisation of DICe's affine shape function).  F: (..., 2, 2)."""
    f11, f12, f21, f22 = F[..., 0, 0], F[..., 0, 1], F[..., 1, 0], F[..., 1, 1]
    th = np.arctan2(f21 - f12, f11 + f22)
    c, s = np.cos(th), np.sin(th)
    return th, c * f11 + s * f21 - 1.0, -s * f12 + c * f22 - 1.0, c * f12 + s * f22


def _schema_lib():
    path = os.path.join(LIB_DIR, "libdice_schema.so")
    if not os.path.exists(path):
        raise RuntimeError("libdice_schema.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P = ctypes.c_void_p
    lib.dice_polygon_pixel_counts.argtypes = [ctypes.c_int, P, P, P, P]
    lib.dice_polygon_pixels.argtypes = [ctypes.c_int, P, P, P, P, P, P]
    return lib


def rasterize_polygons(vert_offsets, vx, vy):
    """Owned pixels of each polygon by Polygon::get_owned_pixels (the library's C++ rasteriser, OpenMP):
    returns (pixel_offsets int64, px int32, py int32), pixels of polygon i in (y, x) order."""
    lib = _schema_lib()
    vo = np.ascontiguousarray(vert_offsets, dtype=np.int32)
    vx = np.ascontiguousarray(vx, dtype=np.int32)
    vy = np.ascontiguousarray(vy, dtype=np.int32)
    n = len(vo) - 1
    counts = np.zeros(n, dtype=np.int32)
    if lib.dice_polygon_pixel_counts(n, _ptr(vo), _ptr(vx), _ptr(vy), _ptr(counts)) != 0:
        raise ValueError("polygon with fewer than three vertices")
    off = np.zeros(n + 1, dtype=np.int64)
    np.cumsum(counts, out=off[1:])
    px = np.zeros(off[-1], dtype=np.int32)
    py = np.zeros(off[-1], dtype=np.int32)
    lib.dice_polygon_pixels(n, _ptr(vo), _ptr(vx), _ptr(vy), _ptr(off), _ptr(px), _ptr(py))
    return off, px, py


def conformal_subsets(n, width, height, area_lo=200, area_hi=3000, seed=5):
    """n convex polygon subsets: 5 to 9 vertices at random angles on a circle (a jittered circle of random
    radius), rasterised by the winding rule, kept when their pixel count is in [area_lo, area_hi].
    Returns rounded pixel centroids (cx, cy) and the CSR (offsets int32, px, py, active) for Context.set_subsets."""
    rng = np.random.default_rng(seed)
    cxs, vo, vxs, vys = [], [0], [], []
    while len(cxs) < n:
        m = int((n - len(cxs)) * 1.3) + 16
        area = rng.uniform(area_lo, area_hi, m)
        r = np.sqrt(area / np.pi) * 1.08  # the inscribed polygon is a little smaller than its circle
        margin = r + 12
        cx = np.round(rng.uniform(margin, width - margin)).astype(np.int64)
        cy = np.round(rng.uniform(margin, height - margin)).astype(np.int64)
        k = rng.integers(5, 10, m)
        for i in range(m):
            ang = np.sort(rng.uniform(0, 2 * np.pi, k[i]))
            vxs.append(np.round(cx[i] + r[i] * np.cos(ang)).astype(np.int32))
            vys.append(np.round(cy[i] + r[i] * np.sin(ang)).astype(np.int32))
            vo.append(vo[-1] + k[i])
        cxs.extend(cx.tolist())
    vo = np.array(vo, dtype=np.int32)
    off, px, py = rasterize_polygons(vo, np.concatenate(vxs), np.concatenate(vys))
    cnt = np.diff(off)
    keep = np.where((cnt >= area_lo) & (cnt <= area_hi))[0][:n]
    sel_off = np.zeros(len(keep) + 1, dtype=np.int64)
    np.cumsum(cnt[keep], out=sel_off[1:])
    idx = np.concatenate([np.arange(off[i], off[i + 1]) for i in keep])
    px, py = px[idx], py[idx]
    # subset centre: the rounded pixel centroid (random vertex angles can leave the circle's centre
    # several pixels off the polygon's own centre)
    seg = np.repeat(np.arange(len(keep)), cnt[keep])
    cx = np.round(np.bincount(seg, px) / cnt[keep])
    cy = np.round(np.bincount(seg, py) / cnt[keep])
    return cx, cy, sel_off.astype(np.int32), px, py, np.ones(len(idx), dtype=np.uint8)


def affine_motion(frames, width, height, seed=6, t_max=3.0, theta_max_deg=2.0, strain_max=1e-2):
    """A smooth random rigid + affine motion about the image centre over `frames` frames: translation
    |t| <= t_max px, |theta| <= theta_max, strains <= strain_max.  Returns per frame (F 2x2, t)."""
    rng = np.random.default_rng(seed)
    k = np.arange(1, frames + 1) / frames
    def smooth(scale):
        a, f, ph = rng.uniform(0.5, 1.0, 3), rng.uniform(0.5, 2.0, 3), rng.uniform(0, 2 * np.pi, 3)
        v = sum(a[j] * np.sin(2 * np.pi * f[j] * k + ph[j]) for j in range(3))
        return scale * v / np.abs(v).max()
    tx, ty = smooth(t_max), smooth(t_max)
    th = smooth(np.deg2rad(theta_max_deg))
    ex, ey, g = smooth(strain_max), smooth(strain_max), smooth(strain_max)
    out = []
    for i in range(frames):
        c, s = np.cos(th[i]), np.sin(th[i])
        F = np.array([[c, -s], [s, c]]) @ np.array([[1 + ex[i], g[i]], [g[i], 1 + ey[i]]])
        out.append((F, np.array([tx[i], ty[i]])))
    return out


def affine_about(F, t, cx, cy):
    """Field x = c + F (X - c) + t as affine_field coefficients."""
    d = F - np.eye(2)
    u0 = t[0] - d[0, 0] * cx - d[0, 1] * cy
    v0 = t[1] - d[1, 0] * cx - d[1, 1] * cy
    return affine_field(u0, v0, d[0, 0], d[0, 1], d[1, 0], d[1, 1])


def config3(n_subsets=50000, frames=100, size=1024, seed=1):
    """1024^2 speckle, `frames` frames of smooth random rigid + affine motion, n conformal polygon subsets
    of 200-3000 pixels.  Returns ref, [def frames], motion [(F, t)], subsets tuple."""
    sp = Speckle(size, size, seed=seed)
    ref = sp.frame(None)
    c = (size - 1) / 2.0
    motion = affine_motion(frames, size, size, seed=seed + 5)
    seq = [sp.frame(affine_about(F, t, c, c), noise_sigma=2.55, noise_seed=seed + 200 + i)
           for i, (F, t) in enumerate(motion)]
    return ref, seq, motion, conformal_subsets(n_subsets, size, size, seed=seed + 4)


def texture_u16(width, height, field=None, n_waves=20, contrast=0.05, noise=0.005, seed=11, noise_seed=12):
    """12-bit low-contrast smooth texture as uint16: mid-grey plus a sum of n_waves random plane
    sinusoids (wavelengths 6-30 px) scaled to +-contrast * 4095, rendered at the reference position of
    each pixel (field: affine (F, t) about the image centre, inverted exactly), plus Gaussian noise of
    noise * 4095, rounded and clipped to [0, 4095]."""
    rng = np.random.default_rng(seed)
    lam = rng.uniform(6.0, 30.0, n_waves)
    ang = rng.uniform(0, np.pi, n_waves)
    kx, ky = 2 * np.pi * np.cos(ang) / lam, 2 * np.pi * np.sin(ang) / lam
    ph = rng.uniform(0, 2 * np.pi, n_waves)
    amp = rng.uniform(0.5, 1.0, n_waves)
    amp *= contrast * 4095.0 / amp.sum()
    ys, xs = np.mgrid[0:height, 0:width].astype(np.float64)
    if field is not None:
        F, t = field
        c = np.array([(width - 1) / 2.0, (height - 1) / 2.0])
        Fi = np.linalg.inv(F)
        dx, dy = xs - c[0] - t[0], ys - c[1] - t[1]
        xs, ys = c[0] + Fi[0, 0] * dx + Fi[0, 1] * dy, c[1] + Fi[1, 0] * dx + Fi[1, 1] * dy
    img = np.full((height, width), 2048.0)
    for j in range(n_waves):
        img += amp[j] * np.sin(kx[j] * xs + ky[j] * ys + ph[j])
    if noise > 0:
        img += np.random.default_rng(noise_seed).normal(0.0, noise * 4095.0, img.shape)
    return np.clip(np.round(img), 0, 4095).astype(np.uint16)


def config5(width=4096, height=3000, seed=11):
    """4096 x 3000 uint16 texture pair under translation + shear up to 2 px about the centre."""
    F = np.array([[1.0, 2.0 / height], [0.0, 1.0]])  # shear: u grows by 1 px over half the height
    t = np.array([1.0, 0.5])
    ref = texture_u16(width, height, None, seed=seed, noise_seed=seed + 1)
    de = texture_u16(width, height, (F, t), seed=seed, noise_seed=seed + 2)
    return ref, de, (F, t)


def config1(seed=1):
    """512 x 512 speckle, u = 0.5 + 0.01 X, v = -0.3 + 0.005 Y, noise sigma = 2.55 (1 % of range)."""
    sp = Speckle(512, 512, seed=seed)
    ref = sp.frame(None)
    de = sp.frame(affine_field(0.5, -0.3, 0.01, 0.0, 0.0, 0.005), noise_sigma=2.55, noise_seed=seed + 1)
    return ref, de


def sinusoid_sequence(size=2048, frames=10, seed=1, noise_sigma=2.55):
    """size^2 speckle, frame k (1..frames): sinusoidal field with A = 0.1 k."""
    sp = Speckle(size, size, seed=seed)
    ref = sp.frame(None)
    seq = [sp.frame(sinusoid_field(0.1 * k), noise_sigma=noise_sigma, noise_seed=seed + 100 + k)
           for k in range(1, frames + 1)]
    return ref, seq
