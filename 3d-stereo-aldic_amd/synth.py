"""Seeded synthetic speckle pairs with known displacement fields.

Images are sums of Gaussian spots exp(-|X - c_k|^2 / r^2) with Poisson
centres, scaled to a grey-level range, plus Gaussian noise, rounded and
clipped.  The deformed image is rendered analytically from the same spot
list: g(x) = I(X(x)) where x = X + u(X), with X(x) found by fixed-point
iteration, so no resampling error enters the ground truth.

Arrays follow the reference's layout: shape (M, N), first axis = image x,
pixel (i, j) (1-based) sits at coordinates (x, y) = (i, j).
"""
from __future__ import annotations

import numpy as np


def spot_centres(M, N, density, seed, margin=12.0):
    rng = np.random.default_rng(seed)
    area = (M + 2 * margin) * (N + 2 * margin)
    k = rng.poisson(density * area)
    cx = rng.uniform(1 - margin, M + margin, k)
    cy = rng.uniform(1 - margin, N + margin, k)
    return np.stack([cx, cy], axis=1), rng


def render(M, N, centres, radius, Xmap=None, Ymap=None, chunk=20000):
    """Sum of spots sampled at reference coordinates (Xmap, Ymap) of each
    pixel (identity when None)."""
    out = np.zeros(M * N, dtype=np.float64)
    if Xmap is None:
        ii, jj = np.meshgrid(np.arange(1, M + 1, dtype=np.float64),
                             np.arange(1, N + 1, dtype=np.float64), indexing="ij")
        Xmap, Ymap = ii, jj
        disp = np.zeros((len(centres), 2))
    else:
        disp = None
    Xf = Xmap.reshape(-1, order="F")
    Yf = Ymap.reshape(-1, order="F")
    rr = int(np.ceil(4 * radius)) + 2
    offs = np.arange(-rr, rr + 1)
    ox, oy = np.meshgrid(offs, offs, indexing="ij")
    ox = ox.reshape(-1)
    oy = oy.reshape(-1)
    inv_r2 = 1.0 / (radius * radius)
    for s in range(0, len(centres), chunk):
        c = centres[s:s + chunk]
        if disp is None:
            # where does this spot land in the deformed image?  Search the
            # pixels around c + u(c) via the inverse map's local shift.
            ci = np.clip(np.round(c[:, 0]).astype(np.int64), 1, M)
            cj = np.clip(np.round(c[:, 1]).astype(np.int64), 1, N)
            k0 = (ci - 1) + (cj - 1) * M
            sx = ci - Xf[k0]
            sy = cj - Yf[k0]
            bx = np.round(c[:, 0] + sx).astype(np.int64)
            by = np.round(c[:, 1] + sy).astype(np.int64)
        else:
            bx = np.round(c[:, 0]).astype(np.int64)
            by = np.round(c[:, 1]).astype(np.int64)
        px = bx[:, None] + ox[None, :]
        py = by[:, None] + oy[None, :]
        ok = (px >= 1) & (px <= M) & (py >= 1) & (py <= N)
        kk = (px - 1) + (py - 1) * M
        kk = np.where(ok, kk, 0)
        dx = Xf[kk] - c[:, 0:1]
        dy = Yf[kk] - c[:, 1:2]
        val = np.exp(-(dx * dx + dy * dy) * inv_r2) * ok
        out += np.bincount(kk.reshape(-1), weights=val.reshape(-1), minlength=M * N)
    return out.reshape(M, N, order="F")


def inverse_map(M, N, ufun, vfun, iters=30):
    """Reference coordinates X(x) with x = X + u(X) for every pixel x."""
    x, y = np.meshgrid(np.arange(1, M + 1, dtype=np.float64),
                       np.arange(1, N + 1, dtype=np.float64), indexing="ij")
    X, Y = x.copy(), y.copy()
    for _ in range(iters):
        X = x - ufun(X, Y)
        Y = y - vfun(X, Y)
    return X, Y


def make_pair(M, N, ufun, vfun, density=0.02, radius=3.0, seed=1, bits=8, noise=2.0,
              lo=20.0, hi=235.0):
    """Reference and deformed speckle images (float64 grey levels, integer
    valued) plus the spot list.  ``bits`` 8 -> [lo, hi] grey levels; 12 ->
    the same range times 16; noise sigma is in 8-bit grey levels."""
    centres, rng = spot_centres(M, N, density, seed)
    I0 = render(M, N, centres, radius)
    X, Y = inverse_map(M, N, ufun, vfun)
    I1 = render(M, N, centres, radius, X, Y)
    a, b = I0.min(), I0.max()
    scale = 16.0 if bits == 12 else 1.0
    top = (2 ** bits) - 1

    def to_gl(I):
        g = (lo + (hi - lo) * (I - a) / (b - a)) * scale
        g = g + rng.normal(0.0, noise * scale, size=g.shape)
        return np.clip(np.round(g), 0, top)

    return np.asfortranarray(to_gl(I0)), np.asfortranarray(to_gl(I1)), centres


# ---------------------------------------------------------------------------
# Displacement fields of the benchmark configurations
# ---------------------------------------------------------------------------
def field_c1():
    """u = 0.5 + 0.002 x + 0.3 sin(2 pi y / 128), v = -0.3 + 0.001 y."""
    u = lambda X, Y: 0.5 + 0.002 * X + 0.3 * np.sin(2 * np.pi * Y / 128.0)
    v = lambda X, Y: -0.3 + 0.001 * Y
    return u, v


def field_stereo(M, N, shift=40.0, spread=6.0, curvature=0.005):
    """Disparity of a tilted plane seen by a second camera: a 40 px x-shift,
    a linear foreshortening about the image centre that adds +/- ``spread``
    px at the left/right edges, and a small projective curvature term
    (``curvature`` px at the edges).  The spread is kept so the whole field
    lies inside a radius-50 search: the 6% foreshortening of a full 20 deg
    tilt of a 2048 px image would alone add +/- 61 px."""
    xc = (M + 1) / 2.0
    half = xc - 1.0

    def u(X, Y):
        d = (X - xc) / half
        return shift + spread * d + curvature * d * d

    def v(X, Y):
        return 0.0 * X

    return u, v


def field_translation_stretch(stretch=0.001, ty=0.7, M=2048):
    xc = (M + 1) / 2.0
    return (lambda X, Y: stretch * (X - xc)), (lambda X, Y: ty + 0.0 * X)


def field_bulge(M, N, shift=25.0, peak=15.0, sigma=600.0):
    """Disparity of a plane with a Gaussian bulge seen by a second camera:
    a uniform x-shift plus ``peak`` px at the centre decaying with
    ``sigma`` px (up to shift + peak = 40 px), no y component."""
    xc, yc = (M + 1) / 2.0, (N + 1) / 2.0

    def u(X, Y):
        return shift + peak * np.exp(-((X - xc) ** 2 + (Y - yc) ** 2) / (2.0 * sigma ** 2))

    def v(X, Y):
        return 0.0 * X

    return u, v
