"""Host post-processing of the integer search: funRemoveOutliers with its
median test and John D'Errico's inpaint_nans (methods 3 and 4, the two the
search path calls), vendored by the reference as func/inpaint_nans.m and
func/funRemoveOutliers.m.

Grids are 2-D arrays used in MATLAB's column-major order (linear index =
row + col * n for an n x m grid), as the reference indexes them.  The
sparse least-squares solves that MATLAB's backslash does with a QR
factorisation are done here through the augmented system
[[I, A], [A', 0]] [r; x] = [b; 0] by a sparse LU: the same minimiser of
||A x - b|| for a full-column-rank A, without squaring A's condition number.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl


def _lsq(A, b):
    """argmin ||A x - b|| for a sparse A of full column rank."""
    A = sp.csr_matrix(A)
    r, c = A.shape
    if c == 0:
        return np.zeros(0)
    K = sp.bmat([[sp.identity(r, format="csr"), A], [A.T, None]], format="csc")
    rhs = np.concatenate([np.asarray(b, dtype=np.float64).reshape(-1), np.zeros(c)])
    sol = spl.spsolve(K, rhs)
    return np.asarray(sol[r:], dtype=np.float64)


# del^4 template of inpaint_nans method 3 (inpaint_nans.m:294-315): the
# linear offsets -2n, -n-1, ..., 2n as (row, column) steps, and their weights
_DEL4 = [((0, -2), 1), ((-1, -1), 2), ((0, -1), -8), ((1, -1), 2), ((-2, 0), 1), ((-1, 0), -8),
         ((0, 0), 20), ((1, 0), -8), ((2, 0), 1), ((-1, 1), 2), ((0, 1), -8), ((1, 1), 2),
         ((0, 2), 1)]
_DEL4_TALKS = [(-2, 0), (-1, -1), (-1, 0), (-1, 1), (0, -2), (0, -1), (0, 1), (0, 2), (1, -1),
               (1, 0), (1, 1), (2, 0)]


def _stencil_rows(nodes_r, nodes_c, n, m, offsets):
    """COO entries of one equation per node (row index = the node's linear
    index) with the given (dr, dc, weight) stencil."""
    lin = nodes_r + nodes_c * n
    rows, cols, vals = [], [], []
    for (dr, dc), w in offsets:
        rows.append(lin)
        cols.append(lin + dr + dc * n)
        vals.append(np.full(lin.size, float(w)))
    if not rows:
        return np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0)
    return np.concatenate(rows), np.concatenate(cols), np.concatenate(vals)


def inpaint_nans(A, method=3):
    """B = inpaint_nans(A, method) for method 3 (del^4 operator, least
    squares over the NaNs and their del^4 neighbours, reduced order at the
    edges; inpaint_nans.m:290-367) and method 4 (zero-length springs
    between each NaN and its horizontal and vertical neighbours;
    inpaint_nans.m:369-398)."""
    A = np.array(A, dtype=np.float64)
    shape = A.shape
    if A.ndim == 1:
        A = A.reshape(-1, 1)
    n, m = A.shape
    nm = n * m
    flat = A.reshape(-1, order="F")
    isn = np.isnan(flat)
    nan_list = np.flatnonzero(isn)
    known = np.flatnonzero(~isn)
    if nan_list.size == 0:
        return A.reshape(shape)
    nr, nc = nan_list % n, nan_list // n  # 0-based row, column
    if method == 3:
        # neighbours of the NaNs under the del^4 template, not NaN themselves
        rr = (nr[None, :] + np.array([t[0] for t in _DEL4_TALKS])[:, None]).reshape(-1)
        cc = (nc[None, :] + np.array([t[1] for t in _DEL4_TALKS])[:, None]).reshape(-1)
        keep = (rr >= 0) & (rr < n) & (cc >= 0) & (cc < m)
        nb = np.unique(rr[keep] + cc[keep] * n)
        nb = nb[~isn[nb]]
        all_lin = np.concatenate([nan_list, nb])
        ar, ac = all_lin % n + 1, all_lin // n + 1  # 1-based as the reference tests them
        R, Cc, V = [], [], []
        sel = (ar >= 3) & (ar <= n - 2) & (ac >= 3) & (ac <= m - 2)
        part = _stencil_rows(ar[sel] - 1, ac[sel] - 1, n, m, _DEL4)
        R.append(part[0]); Cc.append(part[1]); V.append(part[2])
        sel = ((((ar == 2) | (ar == n - 1)) & (ac >= 2) & (ac <= m - 1)) |
               (((ac == 2) | (ac == m - 1)) & (ar >= 2) & (ar <= n - 1)))
        part = _stencil_rows(ar[sel] - 1, ac[sel] - 1, n, m,
                             [((0, -1), 1), ((-1, 0), 1), ((0, 0), -4), ((1, 0), 1), ((0, 1), 1)])
        R.append(part[0]); Cc.append(part[1]); V.append(part[2])
        sel = ((ar == 1) | (ar == n)) & (ac >= 2) & (ac <= m - 1)
        part = _stencil_rows(ar[sel] - 1, ac[sel] - 1, n, m, [((0, -1), 1), ((0, 0), -2),
                                                             ((0, 1), 1)])
        R.append(part[0]); Cc.append(part[1]); V.append(part[2])
        sel = ((ac == 1) | (ac == m)) & (ar >= 2) & (ar <= n - 1)
        part = _stencil_rows(ar[sel] - 1, ac[sel] - 1, n, m, [((-1, 0), 1), ((0, 0), -2),
                                                             ((1, 0), 1)])
        R.append(part[0]); Cc.append(part[1]); V.append(part[2])
        fda = sp.csr_matrix((np.concatenate(V), (np.concatenate(R), np.concatenate(Cc))),
                            shape=(nm, nm))  # duplicates sum, as sparse() does
        rhs = -(fda[:, known] @ flat[known])
        sub = fda[:, nan_list]
        rows = np.flatnonzero(np.asarray(abs(sub).sum(axis=1)).reshape(-1) > 0)
        B = flat.copy()
        B[nan_list] = _lsq(sub[rows], rhs[rows])
    elif method == 4:
        springs = []
        for dr, dc in ((-1, 0), (1, 0), (0, -1), (0, 1)):
            r2, c2 = nr + dr, nc + dc
            ok = (r2 >= 0) & (r2 < n) & (c2 >= 0) & (c2 < m)
            springs.append(np.stack([nan_list[ok], (r2 + c2 * n)[ok]], 1))
        pairs = np.unique(np.sort(np.concatenate(springs), axis=1), axis=0)
        nh = pairs.shape[0]
        S = sp.csr_matrix((np.concatenate([np.ones(nh), -np.ones(nh)]),
                           (np.concatenate([np.arange(nh)] * 2),
                            np.concatenate([pairs[:, 0], pairs[:, 1]]))), shape=(nh, nm))
        rhs = -(S[:, known] @ flat[known])
        B = flat.copy()
        B[nan_list] = _lsq(S[:, nan_list], rhs)
    else:
        raise NotImplementedError("inpaint_nans: only methods 3 and 4 are on this path")
    return B.reshape((n, m), order="F").reshape(shape)


def medFilt2(V0):
    """medFilt2(V0, [2 2], 'symmetric', 2) (funRemoveOutliers.m:393-444): the
    grid padded by one mirrored row and column after the data, the 2 x 2
    neighbourhood anchored at each point without its second element (the
    right neighbour), and the middle of the remaining three -- the point,
    the one below it and the diagonal one -- in single precision."""
    V = np.asarray(V0, dtype=np.float32)
    P = np.pad(V, ((0, 1), (0, 1)), mode="symmetric")
    a = P[:-1, :-1]
    b = P[1:, :-1]
    c = P[1:, 1:]
    return np.median(np.stack([a, b, c], axis=2), axis=2).astype(np.float32)


def funMedRemoveOutliers(u, epsilon=0.1):
    """[medianU, normFluct] = funMedRemoveOutliers(u, epsilon)
    (funRemoveOutliers.m:374-390): NaNs filled by inpaint_nans(., 3), then
    |u - med| / (med(|u - med|) + epsilon) in single precision, as the
    reference's mixed double/single arithmetic gives."""
    uf = inpaint_nans(np.asarray(u, dtype=np.float64), 3)
    med = medFilt2(uf)
    fluct = (uf.astype(np.float32) - med).astype(np.float32)
    res = medFilt2(np.abs(fluct))
    return med, np.abs(fluct / (res + np.float32(epsilon))).astype(np.float32)


def funRemoveOutliers(u, v, cc, qDICOrNot=0, Thr0=100):
    """[u, v, cc, [], [], RemoveOutliersList] = funRemoveOutliers(u, v, cc,
    0, Thr0) (funRemoveOutliers.m:1-87) with a preset threshold: points
    whose normalised median fluctuation magnitude exceeds Thr0 are set to
    NaN and filled by inpaint_nans(., 4) together with the holes, and the
    holes are NaN again afterwards.  RemoveOutliersList is 0-based in the
    column-major order.  The q-factor removal (qDICOrNot = 1) and the
    interactive threshold prompt are not on this path."""
    if qDICOrNot:
        raise NotImplementedError("qDICOrNot = 1 is not used by IntegerSearchQuadtree (:159)")
    if Thr0 is None or Thr0 == 0:
        raise ValueError("the median-test threshold must be given (the reference would prompt)")
    u = np.array(u, dtype=np.float64)
    v = np.array(v, dtype=np.float64)
    hole = np.isnan(u)
    _, fu = funMedRemoveOutliers(u)
    _, fv = funMedRemoveOutliers(v)
    mag = np.sqrt(fu.astype(np.float32) ** 2 + fv.astype(np.float32) ** 2)
    out = mag > Thr0
    rm = np.flatnonzero(out.reshape(-1, order="F"))
    if rm.size == 0:
        # inpaint_nans(., 4) changes nothing but the holes, which are reset
        return u, v, cc, rm
    u2 = u.copy()
    v2 = v.copy()
    u2[out] = np.nan
    v2[out] = np.nan
    u2 = inpaint_nans(u2, 4)
    v2 = inpaint_nans(v2, 4)
    u2[hole] = np.nan
    v2[hole] = np.nan
    return u2, v2, cc, rm
