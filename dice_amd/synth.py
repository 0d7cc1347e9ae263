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
