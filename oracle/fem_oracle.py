"""FP64 numpy/scipy restatement of Subpb2Quadtree, the global FE step of the
ALM loop (func_quadtree/Subpb2Quadtree.m:1-397).

TEST INFRASTRUCTURE ONLY: imported by tests/ (never by the product package).

It follows the reference's own structure rather than the library's: the
full 16 x 16 element matrices over interleaved (x, y) degrees of freedom,
the padded dummy node for missing midpoints (:118-128), the F index table
of :108-113, the 'involved' degrees of freedom of :314-318 (the first entry
of unique(elementsFEM) dropped), the Neumann loop of :336-341, the Dirichlet
lift b - A Uhat (:345-349) and a direct sparse solve (scipy's spsolve in
place of MATLAB's backslash).  The library solves one scalar matrix for both
components by conjugate gradients; agreement with this file therefore also
checks that the components decouple.

Pinned by tests/test_fem.py: on a square element without hanging nodes the
element matrices equal the closed-form bilinear mass and stiffness matrices
(2 x 2 Gauss integrates both exactly); every element's stiffness rows sum to
zero and its mass sums to its area (partition of unity, with and without
hanging nodes); and the patch test -- U linear, F its gradient, Dirichlet
nodes on the boundary -- returns U at every node, on meshes with hanging
nodes, for any beta and mu.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla


def gauss(order):
    """Subpb2Quadtree.m:49-64, literals as written."""
    if order == 2:
        return np.array([-1 / np.sqrt(3), 1 / np.sqrt(3)]), np.array([1.0, 1.0])
    if order == 3:
        g = np.sqrt(3 / 5)
        return np.array([0.0, g, -g]), np.array([8 / 9, 5 / 9, 5 / 9])
    if order == 4:
        return (np.array([0.339981, -0.339981, 0.861136, -0.861136]),
                np.array([0.652145, 0.652145, 0.347855, 0.347855]))
    if order == 5:
        return (np.array([0.0, 0.538469, -0.538469, 0.90618, -0.90618]),
                np.array([0.568889, 0.478629, 0.478629, 0.236927, 0.236927]))
    raise ValueError("GaussPtOrder must be 2..5")


def shape_functions(ksi, eta, delta):
    """N (8,), dN/dksi (8,), dN/deta (8,) of Subpb2Quadtree.m:170-180 and
    the funDN* subroutines (:355-397); delta[k] = 1 when midpoint 5+k exists."""
    d5, d6, d7, d8 = delta
    sg = np.sign  # MATLAB sign(0) = 0, as numpy's
    N5 = d5 * 0.5 * (1 + ksi) * (1 - abs(eta))
    N6 = d6 * 0.5 * (1 + eta) * (1 - abs(ksi))
    N7 = d7 * 0.5 * (1 - ksi) * (1 - abs(eta))
    N8 = d8 * 0.5 * (1 - eta) * (1 - abs(ksi))
    N1 = (1 - ksi) * (1 - eta) * 0.25 - 0.5 * (N7 + N8)
    N2 = (1 + ksi) * (1 - eta) * 0.25 - 0.5 * (N8 + N5)
    N3 = (1 + ksi) * (1 + eta) * 0.25 - 0.5 * (N5 + N6)
    N4 = (1 - ksi) * (1 + eta) * 0.25 - 0.5 * (N6 + N7)
    dk5, de5 = d5 * 0.5 * (1 - abs(eta)), d5 * 0.5 * (1 + ksi) * sg(-eta)
    dk6, de6 = d6 * 0.5 * (1 + eta) * sg(-ksi), d6 * 0.5 * (1 - abs(ksi))
    dk7, de7 = d7 * 0.5 * (-1) * (1 - abs(eta)), d7 * 0.5 * (1 - ksi) * sg(-eta)
    dk8, de8 = d8 * 0.5 * (1 - eta) * sg(-ksi), d8 * 0.5 * (-1) * (1 - abs(ksi))
    dk1, de1 = -0.25 * (1 - eta) - 0.5 * (dk7 + dk8), -0.25 * (1 - ksi) - 0.5 * (de7 + de8)
    dk2, de2 = 0.25 * (1 - eta) - 0.5 * (dk8 + dk5), -0.25 * (1 + ksi) - 0.5 * (de8 + de5)
    dk3, de3 = 0.25 * (1 + eta) - 0.5 * (dk5 + dk6), 0.25 * (1 + ksi) - 0.5 * (de5 + de6)
    dk4, de4 = -0.25 * (1 + eta) - 0.5 * (dk6 + dk7), 0.25 * (1 - ksi) - 0.5 * (de6 + de7)
    return (np.array([N1, N2, N3, N4, N5, N6, N7, N8]),
            np.array([dk1, dk2, dk3, dk4, dk5, dk6, dk7, dk8]),
            np.array([de1, de2, de3, de4, de5, de6, de7, de8]))


def element(pts, delta, order, beta, mu, alpha, Fe, Ue):
    """Element matrix (16 x 16) and vector (16,) over interleaved dofs.
    pts: 8 x 2 node coordinates (0 for missing); Fe: 16 x 4 gathered F
    (rows per dof, columns F(4j-3), F(4j-1), F(4j-2), F(4j)); Ue: (16,)."""
    pt, wt = gauss(order)
    K, E = np.meshgrid(pt, pt, indexing="ij")
    WK, WE = np.meshgrid(wt, wt, indexing="ij")
    A = np.zeros((16, 16))
    b = np.zeros(16)
    for ksi, eta, wk, we in zip(K.ravel(order="F"), E.ravel(order="F"), WK.ravel(order="F"),
                                WE.ravel(order="F")):
        Nv, dk, de = shape_functions(ksi, eta, delta)
        N = np.zeros((2, 16))
        N[0, 0::2] = Nv
        N[1, 1::2] = Nv
        J = np.array([[dk @ pts[:, 0], dk @ pts[:, 1]], [de @ pts[:, 0], de @ pts[:, 1]]])
        det = np.linalg.det(J)
        invJ = np.array([[J[1, 1], -J[0, 1]], [-J[1, 0], J[0, 0]]]) / det
        G = np.zeros((4, 16))
        G[0, 0::2], G[1, 0::2] = dk, de
        G[2, 1::2], G[3, 1::2] = dk, de
        DN = np.kron(np.eye(2), invJ) @ G
        w = det * wk * we
        A += w * ((beta + alpha) * DN.T @ DN + mu * N.T @ N)
        b += w * (beta * np.diag(DN.T @ Fe.T) + mu * N.T @ N @ Ue + alpha * DN.T @ DN @ Ue)
    return A, b


def subpb2(coords, elements, dirichlet, neumann, order, beta, mu, U, F, alpha, winstepsize):
    """Uhat = Subpb2Quadtree(DICmesh, order, beta, mu, U, F, ~, ~, alpha,
    winstepsize, 0).  elements: ne x 8, 1-based, 0 = no midpoint."""
    coords = np.asarray(coords, dtype=np.float64)
    el = np.asarray(elements, dtype=np.int64)
    if el.shape[1] == 4:
        el = np.concatenate([el, np.zeros_like(el)], axis=1)
    n = coords.shape[0]
    size = 2 * n + 8
    Ux = np.concatenate([np.asarray(U, dtype=np.float64).reshape(-1), np.zeros(8)])
    Fx = np.concatenate([np.asarray(F, dtype=np.float64).reshape(-1), np.zeros(16)])
    rows, cols, vals, bi, bv = [], [], [], [], []
    for e in range(el.shape[0]):
        nodes = el[e].copy()
        delta = (nodes[4:] != 0).astype(float)
        nd = np.where(nodes == 0, n + 1, nodes)  # missing -> the dummy node n+1
        iu = np.repeat(2 * nd, 2)
        iu[0::2] -= 1  # 1-based dofs
        j = np.repeat(nd, 2)
        iF = np.stack([4 * j - 3, 4 * j - 1, 4 * j - 2, 4 * j], axis=1)  # 16 x 4, 1-based
        pts = np.zeros((8, 2))
        for a in range(8):
            if nodes[a] != 0:
                pts[a] = coords[nodes[a] - 1]
        A, b = element(pts, delta, order, beta, mu, alpha, Fx[iF - 1], Ux[iu - 1])
        I, Jm = np.meshgrid(iu - 1, iu - 1, indexing="ij")
        rows.append(I.ravel())
        cols.append(Jm.ravel())
        vals.append(A.ravel())
        bi.append(iu - 1)
        bv.append(b)
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))),
                      shape=(size, size))
    b = np.bincount(np.concatenate(bi), weights=np.concatenate(bv), minlength=size)
    involved = np.unique(el[:, :8])[1:]
    uinv = np.sort(np.concatenate([2 * involved - 1, 2 * involved]))
    d = np.unique(np.asarray(dirichlet, dtype=np.int64).reshape(-1))
    dtemp = np.concatenate([2 * d, 2 * d - 1])
    free = np.setdiff1d(uinv, dtemp) - 1  # 0-based
    if neumann is not None and np.size(neumann):
        nm = np.asarray(neumann, dtype=np.float64).reshape(-1, 4)
        BC = -1.0 / winstepsize * Fx
        for t in range(nm.shape[0]):
            a, c = int(nm[t, 0]), int(nm[t, 1])
            L = 0.5 * np.linalg.norm(coords[a - 1] - coords[c - 1])
            idx = np.array([a, c])
            b[2 * idx - 2] += L * (BC[4 * idx - 4] * nm[t, 2] + BC[4 * idx - 2] * nm[t, 3])
            b[2 * idx - 1] += L * (BC[4 * idx - 3] * nm[t, 2] + BC[4 * idx - 1] * nm[t, 3])
    Uhat = np.zeros(size)
    for c in range(2):
        Uhat[2 * d - 1 - c] = Ux[2 * d - 1 - c]
    b = b - A @ Uhat
    if free.size:
        Uhat[free] = spla.spsolve(A[free][:, free].tocsc(), b[free])
    return Uhat[:2 * n]


def element_matrices(pts, delta, order=2):
    """Scalar mass and stiffness matrices of one element (8 x 8 over its
    node slots), from element() with (beta, mu) = (1, 0) and (0, 1)."""
    z16 = np.zeros(16)
    zF = np.zeros((16, 4))
    K, _ = element(pts, delta, order, 1.0, 0.0, 0.0, zF, z16)
    Mm, _ = element(pts, delta, order, 0.0, 1.0, 0.0, zF, z16)
    return Mm[0::2, 0::2], K[0::2, 0::2], Mm, K
