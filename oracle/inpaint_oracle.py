"""Dense restatement of the search's outlier pass: inpaint_nans methods 3
and 4 (func/inpaint_nans.m:290-398), medFilt2 and funRemoveOutliers
(func/funRemoveOutliers.m:1-87, :374-444), for small grids.

TEST INFRASTRUCTURE ONLY: imported by tests/, never by the product package.

Every equation is built by explicit loops over the grid points from the
stencils written as 2-D weight arrays, and the least-squares problems are
solved densely (numpy lstsq).  What pins it (tests/test_outliers.py):
  * method 4: a lone NaN takes the mean of its four neighbours; a NaN run
    in a one-row grid takes the linear interpolant of its two ends;
  * method 3: linear fields are reproduced exactly, holes at the border
    included; quadratic fields are reproduced exactly when every equation
    touching the hole is the del^4 one (a hole at least 5 points from the
    border), which method 4 does not do; its 5 x 5 template is the discrete
    Laplacian convolved with itself;
  * medFilt2: the median of the point, the one below and the diagonal one,
    against a direct numpy median over a mirrored pad.
"""
from __future__ import annotations

import numpy as np

# del^4 template centred at [2, 2], rows = dim 1 (MATLAB rows), columns = dim 2
DEL4 = np.array([[0, 0, 1, 0, 0],
                 [0, 2, -8, 2, 0],
                 [1, -8, 20, -8, 1],
                 [0, 2, -8, 2, 0],
                 [0, 0, 1, 0, 0]], dtype=np.float64)
DEL2 = np.array([[0, 1, 0], [1, -4, 1], [0, 1, 0]], dtype=np.float64)


def _add(E, n, m, i, j, W, ci, cj):
    """Add template W (centre at [ci, cj]) at point (i, j) to row lin(i, j)."""
    row = i + j * n
    for a in range(W.shape[0]):
        for b in range(W.shape[1]):
            if W[a, b] != 0.0:
                E[row, (i + a - ci) + (j + b - cj) * n] += W[a, b]


def inpaint_nans(A, method):
    A = np.array(A, dtype=np.float64)
    shape = A.shape
    if A.ndim == 1:
        A = A.reshape(-1, 1)
    n, m = A.shape
    nm = n * m
    flat = A.reshape(-1, order="F")
    nan = [k for k in range(nm) if np.isnan(flat[k])]
    known = [k for k in range(nm) if not np.isnan(flat[k])]
    if not nan:
        return A.reshape(shape)
    if method == 4:
        springs = set()
        for k in nan:
            i, j = k % n, k // n
            for di, dj in ((-1, 0), (1, 0), (0, -1), (0, 1)):
                i2, j2 = i + di, j + dj
                if 0 <= i2 < n and 0 <= j2 < m:
                    springs.add(tuple(sorted((k, i2 + j2 * n))))
        E = np.zeros((len(springs), nm))
        for r, (a, b) in enumerate(sorted(springs)):
            E[r, a] = 1.0
            E[r, b] = -1.0
        rows = list(range(E.shape[0]))
    elif method == 3:
        nanset = set(nan)
        pts = set(nan)
        for k in nan:
            i, j = k % n, k // n
            for a in range(5):
                for b in range(5):
                    if DEL4[a, b] == 0.0 or (a, b) == (2, 2):
                        continue
                    i2, j2 = i + a - 2, j + b - 2
                    if 0 <= i2 < n and 0 <= j2 < m and (i2 + j2 * n) not in nanset:
                        pts.add(i2 + j2 * n)
        E = np.zeros((nm, nm))
        for k in sorted(pts):
            i, j = k % n, k // n
            r, c = i + 1, j + 1  # 1-based, as the reference tests them
            if 3 <= r <= n - 2 and 3 <= c <= m - 2:
                _add(E, n, m, i, j, DEL4, 2, 2)
            if ((r in (2, n - 1) and 2 <= c <= m - 1) or (c in (2, m - 1) and 2 <= r <= n - 1)):
                _add(E, n, m, i, j, DEL2, 1, 1)
            if r in (1, n) and 2 <= c <= m - 1:
                _add(E, n, m, i, j, np.array([[1.0, -2.0, 1.0]]), 0, 1)
            if c in (1, m) and 2 <= r <= n - 1:
                _add(E, n, m, i, j, np.array([[1.0], [-2.0], [1.0]]), 1, 0)
        rows = [r for r in range(nm) if np.any(E[r, nan] != 0.0)]
    else:
        raise ValueError("methods 3 and 4 only")
    Er = E[rows]
    rhs = -Er[:, known] @ flat[known]
    x = np.linalg.lstsq(Er[:, nan], rhs, rcond=None)[0]
    out = flat.copy()
    out[nan] = x
    return out.reshape((n, m), order="F").reshape(shape)


def medFilt2(V0):
    V = np.asarray(V0, dtype=np.float32)
    n, m = V.shape
    out = np.zeros((n, m), dtype=np.float32)
    for i in range(n):
        for j in range(m):
            i2 = i + 1 if i + 1 < n else n - 1  # mirror pad of one row after the data
            j2 = j + 1 if j + 1 < m else m - 1
            trio = sorted([V[i, j], V[i2, j], V[i2, j2]])
            out[i, j] = trio[1]
    return out


def funRemoveOutliers(u, v, Thr0=100.0, epsilon=0.1):
    u = np.array(u, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    hole = np.isnan(u)

    def fluct(a):
        a3 = inpaint_nans(a, 3)
        med = medFilt2(a3)
        fl = (a3.astype(np.float32) - med).astype(np.float32)
        res = medFilt2(np.abs(fl))
        return np.abs(fl / (res + np.float32(epsilon)))

    mag = np.sqrt(fluct(u) ** 2 + fluct(v) ** 2)
    out = mag > Thr0
    u2, v2 = u.copy(), v.copy()
    u2[out] = np.nan
    v2[out] = np.nan
    u2 = inpaint_nans(u2, 4)
    v2 = inpaint_nans(v2, 4)
    u2[hole] = np.nan
    v2[hole] = np.nan
    return u2, v2, out
