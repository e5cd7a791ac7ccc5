"""The quadtree FE mesh of the stereo and temporal drivers (host code).

* ``MeshSetUp(x, y, DICpara)``                  func/MeshSetUp.m:1-86
* ``GenerateQuadtreeMesh(U0, Df, mask, DICmesh, DICpara)``
                                                func_quadtree/GenerateQuadtreeMesh.m:1-199
* ``QrefineR(coordinates, elements, irregular, marked)``
                                                func_quadtree/refinement/QrefineR.m:1-131
* ``funMarkEdge``, ``funMarkInside``            func_quadtree/funMarkEdge.m, funMarkInside.m

QrefineR is the red refinement with the 1-irregular rule of Funken and
Schmidt ("Adaptive mesh refinement in 2D: an efficient implementation in
Matlab", CMAM 20:459-479, 2020); the reference vendors it with its helper
``provideGeometricData`` (func_quadtree/refinement/provideGeometricData.m),
which :func:`geometric_data` restates: the irregular list [a b h] is taken
as triangles (edges a-b, b-h, h-a), the elements as quadrilaterals (edges
1-2, 2-3, 3-4, 4-1), and the edges are numbered in the lexicographic order
of their sorted node pairs (``unique(sort(edges,2),'rows')``).  The steps
of QrefineR and that numbering are the published algorithm's, kept alike so
that nodes and elements come out in the reference's order.

The per-element mask tests read one summed-area table of the mask instead
of slicing the image per element; the rules (range of the mask over the
element's bounding box, the 50% rule of funMarkInside) are unchanged.
"""
from __future__ import annotations

import numpy as np


# ---------------------------------------------------------------------------
# uniform mesh
# ---------------------------------------------------------------------------
def MeshSetUp(x, y, DICpara):
    """DICmesh = MeshSetUp(x, y, DICpara): x, y are the nY x nX node grids of
    the integer search; nodes x-index fastest, elements [(i,j) (i+1,j)
    (i+1,j+1) (i,j+1)] (MeshSetUp.m:23-40)."""
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    N, M = x.shape  # N along image y, M along image x
    coords = np.stack([x.reshape(-1), y.reshape(-1)], axis=1)  # x.T(:) = x fastest
    i = np.arange(1, M)
    j = np.arange(1, N)
    I, J = np.meshgrid(i, j, indexing="xy")
    I, J = I.reshape(-1), J.reshape(-1)
    el = np.stack([(J - 1) * M + I, (J - 1) * M + I + 1, J * M + I + 1, J * M + I], axis=1)
    ImgSize = DICpara.get("ImgSize")
    mesh = {"coordinatesFEM": np.asfortranarray(coords), "elementsFEM": el.astype(np.int64),
            "dirichlet": np.zeros(0, np.int64), "neumann": np.zeros((0, 4)),
            "x0": x.T, "y0": y.T, "M": M, "N": N}
    if ImgSize is not None:
        mesh["coordinatesFEMWorld"] = np.stack([coords[:, 0], ImgSize[1] + 1 - coords[:, 1]], 1)
    return mesh


# ---------------------------------------------------------------------------
# refinement
# ---------------------------------------------------------------------------
def geometric_data(irregular, elements):
    """[edge2nodes, irregular2edges, element2edges] =
    provideGeometricData(irregular, elements) as QrefineR.m:43-45 calls it:
    the edges of the irregular triangles (a-b, b-h, h-a for each row, all
    first edges, then all second, then all third) followed by those of the
    quadrilaterals (1-2, 2-3, 3-4, 4-1, likewise), numbered by
    unique(sort(edges,2),'rows') (provideGeometricData.m:44-57).
    edge2nodes is E x 2 (sorted pairs, 1-based nodes); edge numbers are
    0-based."""
    el = np.asarray(elements, dtype=np.int64).reshape(-1, 4)
    irr = np.asarray(irregular, dtype=np.int64).reshape(-1, 3)
    tri = np.concatenate([irr[:, [0, 1]], irr[:, [1, 2]], irr[:, [2, 0]]])
    quad = np.concatenate([el[:, [0, 1]], el[:, [1, 2]], el[:, [2, 3]], el[:, [3, 0]]])
    edges = np.sort(np.concatenate([tri, quad]), axis=1)
    edge2nodes, ie = np.unique(edges, axis=0, return_inverse=True)
    ie = ie.reshape(-1)
    nt, ne = len(irr), len(el)
    irregular2edges = ie[:3 * nt].reshape(3, nt).T
    element2edges = ie[3 * nt:3 * nt + 4 * ne].reshape(4, ne).T
    return edge2nodes, irregular2edges, element2edges


def QrefineR(coordinates, elements, irregular, marked):
    """[coordinates, elements4, irregular] = QrefineR(coordinates,
    elements4, irregular, marked): red refinement of the marked elements,
    closed by the 1-irregular rule (QrefineR.m:38-131).  1-based indices,
    marked included."""
    coords = np.asarray(coordinates, dtype=np.float64).reshape(-1, 2)
    el = np.asarray(elements, dtype=np.int64).reshape(-1, 4)
    irr = np.asarray(irregular, dtype=np.int64).reshape(-1, 3)
    marked = np.asarray(marked, dtype=np.int64).reshape(-1) - 1
    ne = len(el)
    edge2nodes, irr2e, el2e = geometric_data(irr, el)
    e2n = np.zeros(len(edge2nodes), dtype=np.int64)
    e2n[el2e[marked].reshape(-1)] = 1
    if len(irr):
        e2n[irr2e[:, 0]] = 1
    kdx = np.array([0])
    swap = np.zeros(0, np.int64)
    first = True
    while first or kdx.size or swap.size:
        first = False
        me = e2n[el2e]
        s = np.abs(me).sum(axis=1)
        kdx = np.flatnonzero((s < 4) & ((s > 2) | (me.min(axis=1) < 0)))
        sub = el2e[kdx][me[kdx] == 0]
        e2n[sub] = 1
        if len(irr):
            mi = e2n[irr2e]
            flag = irr2e[np.any(mi[:, 1:] != 0, axis=1), 0]
            swap = np.flatnonzero(e2n[flag] != -1)
            e2n[flag[swap]] = -1
        else:
            swap = np.zeros(0, np.int64)
    # new nodes on marked edges, in edge order
    if len(irr):
        e2n[irr2e[:, 0]] = -1
    idx = e2n > 0
    nc = len(coords)
    e2n[idx] = nc + 1 + np.arange(int(idx.sum()))
    newc = (coords[edge2nodes[idx, 0] - 1] + coords[edge2nodes[idx, 1] - 1]) / 2
    coords = np.vstack([coords, newc])
    if len(irr):
        e2n[irr2e[:, 0]] = irr[:, 2]
    newNodes = e2n[el2e]  # ne x 4
    reftyp = (newNodes != 0).astype(np.int64) @ np.array([1, 2, 4, 8])
    none = reftyp < 15
    red = reftyp == 15
    ridx = np.flatnonzero(red)
    mid = np.zeros(ne, dtype=np.int64)
    mid[ridx] = len(coords) + 1 + np.arange(len(ridx))
    c = coords
    centre = (((c[el[ridx, 0] - 1] + c[el[ridx, 1] - 1]) + c[el[ridx, 2] - 1]) +
              c[el[ridx, 3] - 1]) / 4
    coords = np.vstack([coords, centre])
    cnt = np.where(none, 1, 4)
    start = np.concatenate([[0], np.cumsum(cnt)])[:-1]
    out = np.zeros((int(cnt.sum()), 4), dtype=np.int64)
    out[start[none]] = el[none]
    nn = newNodes[red]
    e = el[red]
    m = mid[red]
    s0 = start[red]
    out[s0] = np.stack([e[:, 0], nn[:, 0], m, nn[:, 3]], axis=1)
    out[s0 + 1] = np.stack([e[:, 1], nn[:, 1], m, nn[:, 0]], axis=1)
    out[s0 + 2] = np.stack([e[:, 2], nn[:, 2], m, nn[:, 1]], axis=1)
    out[s0 + 3] = np.stack([e[:, 3], nn[:, 3], m, nn[:, 2]], axis=1)
    # hanging nodes of the unrefined elements, then the halves of old ones
    part = np.flatnonzero((reftyp > 0) & (reftyp < 15))
    sub = newNodes[part]
    jj, ii = np.nonzero(sub.T)  # MATLAB find: column-major
    edx = el2e[part[ii], jj]
    newirr = np.concatenate([edge2nodes[edx], sub[ii, jj][:, None]], axis=1)
    if len(irr):
        halves = e2n[irr2e[:, 1:3]]
        k2 = np.flatnonzero(halves.sum(axis=1) != 0)
        sub2 = halves[k2]
        jj, ii = np.nonzero(sub2.T)
        edx = irr2e[k2[ii], jj + 1]
        newirr = np.concatenate([newirr, np.concatenate([edge2nodes[edx], sub2[ii, jj][:, None]],
                                                        axis=1)])
    return coords, out, newirr.astype(np.int64)


# ---------------------------------------------------------------------------
# mask tests on element bounding boxes
# ---------------------------------------------------------------------------
def _sat(mask):
    m = np.asarray(mask, dtype=np.float64)
    S = np.zeros((m.shape[0] + 1, m.shape[1] + 1))
    S[1:, 1:] = m.cumsum(0).cumsum(1)
    return S


def _boxes(coordinates, elements4, shape):
    c = np.asarray(coordinates, dtype=np.float64)
    el = np.asarray(elements4, dtype=np.int64)[:, :4] - 1
    xs, ys = c[el, 0], c[el, 1]
    lo_x, hi_x, lo_y, hi_y = xs.min(1), xs.max(1), ys.min(1), ys.max(1)
    for v in (lo_x, hi_x, lo_y, hi_y):
        if np.any(v != np.round(v)):
            raise ValueError("an element corner is not on a pixel: the reference indexes "
                             "the mask with it")
    if lo_x.min() < 1 or lo_y.min() < 1 or hi_x.max() > shape[0] or hi_y.max() > shape[1]:
        raise ValueError("an element reaches outside the mask")
    return lo_x.astype(np.int64), hi_x.astype(np.int64), lo_y.astype(np.int64), hi_y.astype(np.int64)


def _box_sum(S, a, b, c, d):
    """sum of mask(a:b, c:d), 1-based inclusive."""
    return S[b, d] - S[a - 1, d] - S[b, c - 1] + S[a - 1, c - 1]


def funMarkEdge(coordinates, elements4, ImgRefMask, elementMinSize):
    """[mark3, mark4] = funMarkEdge(...): elements whose bounding box sees
    both values of the mask and whose smaller side exceeds elementMinSize."""
    if len(elements4) == 0:
        return [], np.zeros(0, bool)
    m = np.asarray(ImgRefMask, dtype=np.float64)
    vals = np.unique(m)
    if not np.all(np.isin(vals, [0.0, 1.0])):
        raise ValueError("funMarkEdge: the mask must hold 0 and 1 only")
    a, b, c, d = _boxes(coordinates, elements4, m.shape)
    S = _sat(m)
    tot = _box_sum(S, a, b, c, d)
    count = (b - a + 1) * (d - c + 1)
    rng = (tot > 0) & (tot < count)  # max - min > 0 on a 0/1 mask
    size = np.minimum(np.abs(b - a), np.abs(d - c))
    return [], rng & (size > elementMinSize)


def funMarkInside(coordinates, elements4, ImgRefMask):
    """[markInside4, markOutside4] (1-based): inside if the mask sums to less
    than half of (xmax - xmin) * (ymax - ymin) over the bounding box."""
    if len(elements4) == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    m = np.asarray(ImgRefMask, dtype=np.float64)
    a, b, c, d = _boxes(coordinates, elements4, m.shape)
    tot = _box_sum(_sat(m), a, b, c, d)
    inside = tot < 0.5 * (b - a) * (d - c)
    return np.flatnonzero(inside) + 1, np.flatnonzero(~inside) + 1


# ---------------------------------------------------------------------------
# GenerateQuadtreeMesh
# ---------------------------------------------------------------------------
def _reorder(coords, el):
    """Corner order (xmin,ymin) (xmax,ymin) (xmax,ymax) (xmin,ymax), the
    reference's sort of (x - xmin)^2 + 2 (y - ymin)^2 (:36-47)."""
    cx = coords[el[:, :4] - 1, 0]
    cy = coords[el[:, :4] - 1, 1]
    t = (cx - cx.min(1, keepdims=True)) ** 2 + 2 * (cy - cy.min(1, keepdims=True)) ** 2
    o = np.argsort(t, axis=1, kind="stable")[:, [0, 1, 3, 2]]
    out = el.copy()
    out[:, :4] = np.take_along_axis(el[:, :4], o, axis=1)
    return out


def _hanging_columns(el8, irregular):
    """Columns 5..8 from the hanging nodes (:56-75): first element whose
    edge (in either direction) is the irregular edge."""
    el8 = el8.copy()
    pairs = [((0, 1), 7), ((1, 2), 4), ((2, 3), 5), ((3, 0), 6),
             ((1, 0), 7), ((2, 1), 4), ((3, 2), 5), ((0, 3), 6)]
    lookup = []
    for (p, q), colx in pairs:
        d = {}
        for r in range(len(el8) - 1, -1, -1):  # first row wins
            d[(int(el8[r, p]), int(el8[r, q]))] = r
        lookup.append((d, colx))
    for a, b, h in np.asarray(irregular, dtype=np.int64).reshape(-1, 3):
        for d, colx in lookup:
            r = d.get((int(a), int(b)))
            if r is not None:
                el8[r, colx] = h
    return el8


def _unique_nodes(rows):
    u = np.unique(rows)
    return u[1:] if u.size and u[0] == 0 else u


def GenerateQuadtreeMesh(U0, Df, fNormalizedMask, DICmesh, DICpara, U0_method="linear"):
    """[DICmesh, DICpara, U0] = GenerateQuadtreeMesh(U0, Df, fNormalizedMask,
    DICmesh, DICpara) (func_quadtree/GenerateQuadtreeMesh.m:1-199).

    ``DICmesh`` needs coordinatesFEM, elementsFEM (the uniform mesh) and
    elementMinSize; ``DICpara`` winstepsize (and winsize).  U0 (2n over the
    uniform mesh, NaN allowed) is carried to the new nodes by piecewise
    linear interpolation on a Delaunay triangulation of its valid nodes,
    linear outside their hull (scatteredInterpolant 'linear', 'linear').
    ``Df`` is unused, as in the reference."""
    mask = np.asarray(fNormalizedMask, dtype=np.float64)
    coords0 = np.asarray(DICmesh["coordinatesFEM"], dtype=np.float64)[:, :2]
    c = coords0.copy()
    el = np.asarray(DICmesh["elementsFEM"], dtype=np.int64)[:, :4]
    irregular = np.zeros((0, 3), dtype=np.int64)
    minsize = DICmesh["elementMinSize"]
    while True:
        _, mark4 = funMarkEdge(c, el, mask, minsize)
        mark4 = np.flatnonzero(mark4)
        c, el, irregular = QrefineR(c, el, irregular, mark4 + 1)
        if mark4.size == 0:
            break
    el = _reorder(c, el)
    el8 = np.concatenate([el, np.zeros((len(el), 4), np.int64)], axis=1)
    el8 = _hanging_columns(el8, irregular)
    _, outside = funMarkInside(c, el8, mask)
    el8 = el8[outside - 1]
    # nodes near the hole edges and the boundary (:104-150)
    ws = np.max(np.atleast_1d(np.asarray(DICpara["winstepsize"], dtype=np.float64)))
    diag = np.hypot(c[el8[:, 0] - 1, 0] - c[el8[:, 2] - 1, 0],
                    c[el8[:, 0] - 1, 1] - c[el8[:, 2] - 1, 1])
    refined = diag < 0.99 * np.sqrt(2) * max(ws, 0.0)
    xMin, xMax = coords0[:, 0].min(), coords0[:, 0].max()
    yMin, yMax = coords0[:, 1].min(), coords0[:, 1].max()
    step = np.mean(np.atleast_1d(DICpara["winstepsize"]))
    near = ((c[el8[:, 0] - 1, 0] < xMin + 1.01 * step) | (c[el8[:, 2] - 1, 0] > xMax - 1.01 * step) |
            (c[el8[:, 0] - 1, 1] < yMin + 1.01 * step) | (c[el8[:, 2] - 1, 1] > yMax - 1.01 * step))
    marked = refined | near
    edge = _unique_nodes(el8[marked])
    for _ in range(2):
        touch = np.isin(el8, edge).any(axis=1)
        edge = _unique_nodes(el8[touch])
    out = dict(DICmesh)
    out["coordinatesFEM"] = np.asfortranarray(c)
    out["elementsFEM"] = el8
    out["irregular"] = irregular
    out["markCoordHoleEdge"] = edge
    out["dirichlet"] = edge
    out["coordinatesFEMWorld"] = np.stack([c[:, 0], mask.shape[1] + 1 - c[:, 1]], axis=1)
    U0 = np.asarray(U0, dtype=np.float64).reshape(-1)
    ok = np.flatnonzero(~np.isnan(U0[0::2]))
    if len(ok) >= 3:
        vals = np.stack([U0[2 * ok], U0[2 * ok + 1]], axis=1)
        Ui = linear_interp(coords0[ok], vals, c)
        Unew = np.empty(2 * len(c))
        Unew[0::2] = Ui[:, 0]
        Unew[1::2] = Ui[:, 1]
    else:
        Unew = np.zeros(2 * len(c))
    return out, DICpara, Unew


def linear_interp(xy, vals, q):
    """scatteredInterpolant(x, y, v, 'linear', 'linear') at q: barycentric
    on the Delaunay triangulation of the sites; outside their hull, the
    plane of the triangle whose boundary edge is nearest to the query.
    Where the triangulation is not unique (cocircular sites) the choice of
    diagonal is scipy's (Qhull), which may differ from MATLAB's."""
    from scipy.spatial import Delaunay

    tri = Delaunay(xy)
    simp = tri.find_simplex(q)
    out = np.empty((len(q), vals.shape[1]))
    inside = simp >= 0
    if np.any(inside):
        T = tri.transform[simp[inside]]
        b2 = np.einsum("nij,nj->ni", T[:, :2], q[inside] - T[:, 2])
        bc = np.concatenate([b2, 1 - b2.sum(1, keepdims=True)], axis=1)
        out[inside] = np.einsum("nk,nkc->nc", bc, vals[tri.simplices[simp[inside]]])
    if np.any(~inside):
        # hull edges and the triangle each belongs to
        hull_tri, hull_edge = [], []
        for t, nb in enumerate(tri.neighbors):
            for k in range(3):
                if nb[k] == -1:
                    s = tri.simplices[t]
                    hull_tri.append(t)
                    hull_edge.append([s[(k + 1) % 3], s[(k + 2) % 3]])
        hull_tri = np.array(hull_tri)
        hull_edge = np.array(hull_edge)
        A, B = xy[hull_edge[:, 0]], xy[hull_edge[:, 1]]
        for r in np.flatnonzero(~inside):
            p = q[r]
            d = B - A
            t = np.clip(((p - A) * d).sum(1) / (d * d).sum(1), 0, 1)
            dist = np.hypot(*(A + t[:, None] * d - p).T)
            ti = hull_tri[int(np.argmin(dist))]
            s = tri.simplices[ti]
            P = np.c_[xy[s], np.ones(3)]
            coef = np.linalg.solve(P, vals[s])
            out[r] = np.array([p[0], p[1], 1.0]) @ coef
    return out
