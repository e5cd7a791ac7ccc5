"""FP64 numpy restatement of the drivers' NaN fill and of the interpolant it
uses, scatteredInterpolant(x, y, v, 'natural', 'nearest').

TEST INFRASTRUCTURE ONLY: imported by tests/ (never by the product package).

The reference's interpolant is a MATLAB builtin that is absent here, so the
restatement follows the published definition (Sibson 1980, "A vector
identity for the Dirichlet tessellation"): inside the convex hull of the
sites, the weight of site i at q is the area that q's Voronoi cell, in the
diagram of the sites plus q, takes from the Voronoi cell of i, over the area
of q's cell.  On the hull boundary the weights reduce to linear
interpolation between the ends of the hull edge; outside the hull the value
is the nearest site's ('nearest' extrapolation; the first site in input
order on a tie).

It is written to be independent of the library's C++ (natural.cpp):
  * the hull is a Jarvis march (the library: monotone chain);
  * q's cell is clipped from a box by the bisectors of EVERY site, and the
    area taken from site i is that cell clipped by the bisectors of i and
    every other site nearby (the library: sites visited bucket ring by
    bucket ring until no further site can cut the cell, then only the
    natural neighbours);
  * the mask regions are labelled by a breadth-first flood fill over pixels
    (the library: union-find over the runs of each column).

Pinned by tests/test_natural.py: linear precision (a linear field is
reproduced exactly, a property of Sibson's coordinates), symmetric
configurations whose weights are known (the centre of a square: 1/4 each;
of a regular hexagon: 1/6 each), and stolen areas against a raster count
of nearest sites on a fine grid.  The NaN-fill control flow follows
func_quadtree/LocalICGNQuadtree.m:200-247 and :290-325 and
func_quadtree/Subpb1Quadtree.m:189-254; parity with MATLAB's own
scatteredInterpolant, bwconncomp and tie rules is unpinned (no reference
fixture holds their outputs).
"""
from __future__ import annotations

from collections import deque

import numpy as np


def _clip(P, a, b, c):
    """Convex polygon P (list of (x, y), CCW) clipped to a*x + b*y <= c."""
    out = []
    n = len(P)
    for k in range(n):
        x0, y0 = P[k]
        x1, y1 = P[(k + 1) % n]
        d0 = a * x0 + b * y0 - c
        d1 = a * x1 + b * y1 - c
        if d0 <= 0:
            out.append((x0, y0))
        if (d0 <= 0) != (d1 <= 0):
            t = d0 / (d0 - d1)
            out.append((x0 + t * (x1 - x0), y0 + t * (y1 - y0)))
    return out


def _area(P):
    s = 0.0
    for k in range(len(P)):
        x0, y0 = P[k]
        x1, y1 = P[(k + 1) % len(P)]
        s += x0 * y1 - x1 * y0
    return 0.5 * s


def _jarvis(x, y):
    """Convex hull vertex indices, counter-clockwise, no collinear points."""
    n = len(x)
    start = min(range(n), key=lambda i: (x[i], y[i]))
    hull = [start]
    p = start
    while True:
        q = (p + 1) % n
        for r in range(n):
            if r == p:
                continue
            cr = (x[q] - x[p]) * (y[r] - y[p]) - (y[q] - y[p]) * (x[r] - x[p])
            d_q = (x[q] - x[p]) ** 2 + (y[q] - y[p]) ** 2
            d_r = (x[r] - x[p]) ** 2 + (y[r] - y[p]) ** 2
            if cr < 0 or (cr == 0 and d_r > d_q):
                q = r
        p = q
        if p == start:
            break
        hull.append(p)
        if len(hull) > n:
            raise RuntimeError("hull did not close")
    return hull


class Natural:
    """scatteredInterpolant(x, y, v, 'natural', 'nearest')."""

    def __init__(self, x, y, v):
        x = np.asarray(x, dtype=np.float64).ravel()
        y = np.asarray(y, dtype=np.float64).ravel()
        v = np.asarray(v, dtype=np.float64).reshape(len(x), -1)
        # duplicates: one site, the mean value, the first index
        key = {}
        order = []
        for i in range(len(x)):
            k = (x[i], y[i])
            if k not in key:
                key[k] = len(order)
                order.append([i])
            else:
                order[key[k]].append(i)
        self.x = np.array([x[g[0]] for g in order])
        self.y = np.array([y[g[0]] for g in order])
        self.v = np.array([v[g].mean(axis=0) for g in order])
        if len(self.x) < 3:
            raise ValueError("fewer than 3 distinct sites")
        self.hull = _jarvis(self.x, self.y)
        if len(self.hull) < 3:
            raise ValueError("collinear sites")
        self.scale = max(np.ptp(self.x), np.ptp(self.y))

    def weights(self, qx, qy):
        """Sibson weights (site index -> weight) of an interior q."""
        x, y = self.x - qx, self.y - qy
        # near the hull q's cell reaches far out: grow the box until the
        # cell is clear of it
        for grow in (10.0, 1e3, 1e5, 1e7):
            B = grow * self.scale
            V = [(-B, -B), (B, -B), (B, B), (-B, B)]
            for s in np.argsort(np.hypot(x, y), kind="stable"):
                V = _clip(V, x[s], y[s], 0.5 * (x[s] ** 2 + y[s] ** 2))
            R = max(np.hypot(px, py) for px, py in V)
            if R <= 0.4 * B:
                break
        else:
            raise ValueError("q's cell reaches the box: not inside the hull")
        d = np.hypot(x, y)
        cand = np.flatnonzero(d <= 2.0 * R * (1 + 1e-9))
        comp = np.flatnonzero(d <= 4.0 * R * (1 + 1e-9))
        w = {}
        for i in cand:
            T = list(V)
            for j in comp:
                if j == i:
                    continue
                T = _clip(T, x[j] - x[i], y[j] - y[i],
                          0.5 * ((x[j] ** 2 + y[j] ** 2) - (x[i] ** 2 + y[i] ** 2)))
                if not T:
                    break
            a = _area(T) if T else 0.0
            if a > 0:
                w[int(i)] = a
        tot = sum(w.values())
        return {i: a / tot for i, a in w.items()}

    def _locate(self, qx, qy):
        h = self.hull
        on = None
        for k in range(len(h)):
            a, b = h[k], h[(k + 1) % len(h)]
            ex, ey = self.x[b] - self.x[a], self.y[b] - self.y[a]
            cr = ex * (qy - self.y[a]) - ey * (qx - self.x[a])
            tol = 1e-12 * np.hypot(ex, ey) * self.scale
            if cr < -tol:
                return -1, None
            if cr <= tol:
                t = ((qx - self.x[a]) * ex + (qy - self.y[a]) * ey) / (ex * ex + ey * ey)
                on = (a, b, min(max(t, 0.0), 1.0))
        return (0, on) if on is not None else (1, None)

    def __call__(self, qx, qy):
        qx = np.asarray(qx, dtype=np.float64)
        qy = np.asarray(qy, dtype=np.float64)
        out = np.zeros(qx.shape + (self.v.shape[1],))
        for idx in np.ndindex(qx.shape):
            out[idx] = self.eval1(float(qx[idx]), float(qy[idx]))
        return out

    def eval1(self, qx, qy):
        hit = np.flatnonzero((self.x == qx) & (self.y == qy))
        if hit.size:
            return self.v[hit[0]]
        where, on = self._locate(qx, qy)
        if where < 0:
            d = (self.x - qx) ** 2 + (self.y - qy) ** 2
            return self.v[int(np.argmin(d))]  # argmin: first index on a tie
        if where == 0:
            a, b, t = on
            return (1 - t) * self.v[a] + t * self.v[b]
        w = self.weights(qx, qy)
        return sum(wi * self.v[i] for i, wi in w.items())


def label8(mask):
    """8-connected components of mask != 0 by breadth-first flood fill."""
    m = np.asarray(mask) != 0
    lab = np.zeros(m.shape, dtype=np.int64)
    K = 0
    M, N = m.shape
    for j in range(N):
        for i in range(M):
            if not m[i, j] or lab[i, j]:
                continue
            K += 1
            lab[i, j] = K
            dq = deque([(i, j)])
            while dq:
                a, b = dq.popleft()
                for da in (-1, 0, 1):
                    for db in (-1, 0, 1):
                        c, d = a + da, b + db
                        if 0 <= c < M and 0 <= d < N and m[c, d] and not lab[c, d]:
                            lab[c, d] = K
                            dq.append((c, d))
    return lab, K


def _round(v):
    return np.where(v < 0, np.ceil(v - 0.5), np.floor(v + 0.5)).astype(np.int64)


def nan_fill(coords, U, F=None, mask=None):
    """LocalICGNQuadtree.m:200-247 + :290-325 (F given) and
    Subpb1Quadtree.m:189-254 (F None).  Returns a filled copy of U."""
    coords = np.asarray(coords, dtype=np.float64)
    n = coords.shape[0]
    U = np.array(U, dtype=np.float64).reshape(-1)
    nanidx = np.flatnonzero(np.isnan(U[0::2]))
    good = np.setdiff1d(np.arange(n), nanidx)
    if nanidx.size and mask is not None:
        lab, K = label8(mask)
        rx, ry = _round(coords[:, 0]), _round(coords[:, 1])
        node_lab = lab[rx - 1, ry - 1]
        for r in range(1, K + 1):
            src = good[node_lab[good] == r]
            qry = nanidx[node_lab[nanidx] == r]
            if len(src) < 3 or len(qry) == 0:
                continue
            try:
                I = Natural(coords[src, 0], coords[src, 1],
                            np.stack([U[2 * src], U[2 * src + 1]], axis=1))
            except ValueError:
                continue
            val = I(coords[qry, 0], coords[qry, 1])
            U[2 * qry] = val[:, 0]
            U[2 * qry + 1] = val[:, 1]
    nanidx = np.flatnonzero(np.isnan(U[0::2]))
    good = np.setdiff1d(np.arange(n), nanidx)
    fnan = F is not None and np.any(np.isnan(np.asarray(F).reshape(-1)[0::4]))
    if nanidx.size == 0 and not fnan:
        return U
    I = Natural(coords[good, 0], coords[good, 1], np.stack([U[2 * good], U[2 * good + 1]], axis=1))
    val = I(coords[:, 0], coords[:, 1])
    out = np.empty_like(U)
    out[0::2] = val[:, 0]
    out[1::2] = val[:, 1]
    return out
