"""The quadtree mesh (quadtree.py: MeshSetUp, QrefineR, GenerateQuadtreeMesh)
against the cell-tree restatement (oracle/quadtree_oracle.py), CPU only.

Parity is geometric: the same element boxes, the same hanging nodes per
side, the same Dirichlet node positions.  The node numbering depends on the
reference's absent provideGeometricData and is not compared.  The U0
transfer (scatteredInterpolant 'linear') is pinned by linear precision:
a linear U0 is carried exactly, whichever diagonals the triangulation
picks on the grid.
"""
import numpy as np
import pytest

import quadtree_oracle as qo
from _cases import hole_mask


@pytest.fixture(scope="module")
def qt(pkg):
    from aldic_amd import quadtree
    return quadtree


def _uniform(qt, M, N, step, ws):
    from aldic_amd import mesh as m
    coords, nX, nY = m.uniform_grid([1, M], [1, N], ws, step)
    x = coords[:, 0].reshape(nY, nX)
    y = coords[:, 1].reshape(nY, nX)
    return qt.MeshSetUp(x, y, {"ImgSize": (M, N), "winstepsize": step}), x, y


def _boxes(c, el):
    out = {}
    for e in el:
        xs, ys = c[e[:4] - 1, 0], c[e[:4] - 1, 1]
        hang = [tuple(c[h - 1]) if h else None for h in e[4:8]]
        out[(xs.min(), ys.min(), xs.max(), ys.max())] = hang
    return out


def test_meshsetup_layout(qt):
    DICmesh, x, y = _uniform(qt, 200, 160, 16, 16)
    c, el = DICmesh["coordinatesFEM"], DICmesh["elementsFEM"]
    M = x.shape[1]
    assert c[1, 0] - c[0, 0] == 16 and c[M, 1] - c[0, 1] == 16  # x fastest
    # element corners (xmin,ymin) (xmax,ymin) (xmax,ymax) (xmin,ymax)
    p = c[el[0] - 1]
    assert np.array_equal(p, [[p[0, 0], p[0, 1]], [p[0, 0] + 16, p[0, 1]],
                              [p[0, 0] + 16, p[0, 1] + 16], [p[0, 0], p[0, 1] + 16]])
    assert len(el) == (x.shape[0] - 1) * (x.shape[1] - 1)


def test_qrefine_single_element_and_closure(qt):
    # 3 x 3 uniform elements; refine the centre: its four neighbours get one
    # hanging node each, nothing else changes
    c = np.array([[x, y] for y in range(4) for x in range(4)], dtype=float) * 8
    el = np.array([[j * 4 + i + 1, j * 4 + i + 2, (j + 1) * 4 + i + 2, (j + 1) * 4 + i + 1]
                   for j in range(3) for i in range(3)])
    c2, el2, irr = qt.QrefineR(c, el, np.zeros((0, 3)), [5])
    assert len(el2) == 12 and len(c2) == 16 + 5 and len(irr) == 4
    # refine one child of the centre, at the corner next to element 1: the
    # 1-irregular rule refines its coarse neighbours across the hanging sides
    child = [k for k in range(len(el2)) if c2[el2[k] - 1, 0].min() == 8 and
             c2[el2[k] - 1, 1].min() == 8][0]
    c3, el3, irr3 = qt.QrefineR(c2, el2, irr, [child + 1])
    sizes = sorted({np.ptp(c3[e - 1, 0]) for e in el3})
    assert sizes == [2.0, 4.0, 8.0]
    tree = qo.Tree(0, 0, 8, 3, 3)
    tree.refine([(0, 1, 1)])
    tree.refine([(1, 2, 2)])
    want = {tree.box(cc): tree.hanging(cc) for cc in tree.leaves}
    assert set(_boxes(c3, np.concatenate([el3, np.zeros((len(el3), 4), int)], 1))) == set(want)


@pytest.mark.parametrize("case", ["hole", "two-holes", "edge-strip"])
def test_generate_matches_tree(qt, case):
    M, N, step, ws = 256, 224, 16, 16
    if case == "hole":
        mask = hole_mask(M, N, (120.0, 110.0), 45.0)
    elif case == "two-holes":
        mask = hole_mask(M, N, (70.0, 70.0), 25.0) * hole_mask(M, N, (180.0, 150.0), 30.0)
    else:
        mask = np.ones((M, N))
        mask[150:, :] = 0
        mask[:, :20] = 0
    DICmesh, x, y = _uniform(qt, M, N, step, ws)
    DICmesh["elementMinSize"] = 2
    n0 = len(DICmesh["coordinatesFEM"])
    U0 = np.zeros(2 * n0)
    out, _, U = qt.GenerateQuadtreeMesh(U0, None, mask, DICmesh, {"winstepsize": step,
                                                                  "winsize": ws})
    c, el = out["coordinatesFEM"], out["elementsFEM"]
    got = _boxes(c, el)
    want, tree = qo.generate(mask, x.min(), y.min(), step, x.shape[1] - 1, x.shape[0] - 1, 2)
    assert set(got) == set(want)
    for b in got:
        assert got[b] == [tuple(h) if h else None for h in want[b]], b
    # no duplicated node positions; every node used
    assert len(np.unique(c, axis=0)) == len(c)
    assert len(np.unique(el[el > 0])) <= len(c)
    assert U.shape == (2 * len(c),)


def test_generate_dirichlet_and_u0(qt):
    M, N, step, ws = 256, 224, 16, 16
    mask = hole_mask(M, N, (130.0, 100.0), 40.0)
    DICmesh, x, y = _uniform(qt, M, N, step, ws)
    DICmesh["elementMinSize"] = 4
    c0 = DICmesh["coordinatesFEM"]
    U0 = np.empty(2 * len(c0))
    U0[0::2] = 0.3 + 0.01 * c0[:, 0] - 0.02 * c0[:, 1]
    U0[1::2] = -0.1 + 0.005 * c0[:, 0] + 0.001 * c0[:, 1]
    U0[2 * 7] = U0[2 * 7 + 1] = np.nan  # a bad node is left out of the transfer
    out, _, U = qt.GenerateQuadtreeMesh(U0, None, mask, DICmesh, {"winstepsize": step,
                                                                  "winsize": ws})
    c, el = out["coordinatesFEM"], out["elementsFEM"]
    assert np.allclose(U[0::2], 0.3 + 0.01 * c[:, 0] - 0.02 * c[:, 1], atol=1e-10)
    assert np.allclose(U[1::2], -0.1 + 0.005 * c[:, 0] + 0.001 * c[:, 1], atol=1e-10)
    # Dirichlet nodes: the nodes of refined or near-boundary elements,
    # grown twice by element adjacency (GenerateQuadtreeMesh.m:104-150)
    xs, ys = c[el[:, :4] - 1, 0], c[el[:, :4] - 1, 1]
    diag = np.hypot(np.ptp(xs, 1), np.ptp(ys, 1))
    mk = ((diag < 0.99 * np.sqrt(2) * step) | (xs.min(1) < c0[:, 0].min() + 1.01 * step) |
          (xs.max(1) > c0[:, 0].max() - 1.01 * step) | (ys.min(1) < c0[:, 1].min() + 1.01 * step) |
          (ys.max(1) > c0[:, 1].max() - 1.01 * step))
    S = set(el[mk].ravel()) - {0}
    for _ in range(2):
        S = set(el[[bool(set(e) & S) for e in el]].ravel()) - {0}
    assert set(out["dirichlet"]) == S
    # the hole's edge is surrounded by refined elements down to the minimum
    assert np.min(np.ptp(xs, 1)) == 4


def test_oracle_pins():
    """The cell tree: no mask edge -> the uniform mesh; a single masked
    pixel refines the cells around it down to elementMinSize; every leaf
    is at most one level coarser than any leaf across a side."""
    M, N = 120, 100
    want, t = qo.generate(np.ones((M, N)), 9.0, 9.0, 16.0, 6, 5, 2)
    assert len(want) == 30 and all(lv == 0 for lv, _, _ in t.leaves)
    mask = np.ones((M, N))
    mask[50, 40] = 0  # pixel (51, 41)
    want, t = qo.generate(mask, 9.0, 9.0, 16.0, 6, 5, 2)
    small = [b for b in want if b[2] - b[0] == 2]
    assert any(b[0] <= 51 <= b[2] and b[1] <= 41 <= b[3] for b in small)
    for c in t.leaves:
        for side in range(4):
            for f in t.finer_along(c, side):
                assert f[0] == c[0] + 1, (c, f)


def test_geometric_data_numbering():
    """provideGeometricData as QrefineR calls it (provideGeometricData.m:
    44-57): irregular rows are triangles a-b, b-h, h-a; element edges 1-2,
    2-3, 3-4, 4-1; edges numbered in the lexicographic order of their
    sorted node pairs, edge2nodes holding those sorted pairs."""
    from aldic_amd import quadtree as q
    #  4---3---6
    #  |   |   |
    #  1---2---5
    el = np.array([[1, 2, 3, 4], [2, 5, 6, 3]])
    irr = np.array([[5, 6, 7]])  # a hanging node 7 on edge 5-6
    e2n, i2e, el2e = q.geometric_data(irr, el)
    want = [(1, 2), (1, 4), (2, 3), (2, 5), (3, 4), (3, 6), (5, 6), (5, 7), (6, 7)]
    assert [tuple(r) for r in e2n] == want
    idx = {p: k for k, p in enumerate(want)}
    assert list(el2e[0]) == [idx[(1, 2)], idx[(2, 3)], idx[(3, 4)], idx[(1, 4)]]
    assert list(el2e[1]) == [idx[(2, 5)], idx[(5, 6)], idx[(3, 6)], idx[(2, 3)]]
    assert list(i2e[0]) == [idx[(5, 6)], idx[(6, 7)], idx[(5, 7)]]


def test_qrefiner_node_order():
    """One red refinement: the new edge nodes are numbered in edge order
    (the lexicographic numbering above: edges 1-2, 1-4, 2-3, 3-4), then the
    midpoint, and the four children follow QrefineR.m's layout."""
    from aldic_amd import quadtree as q
    c = np.array([[0.0, 0.0], [8.0, 0.0], [8.0, 8.0], [0.0, 8.0]])
    el = np.array([[1, 2, 3, 4]])
    c2, el2, irr2 = q.QrefineR(c, el, np.zeros((0, 3), np.int64), [1])
    assert np.allclose(c2[4:], [[4, 0], [0, 4], [8, 4], [4, 8], [4, 4]])
    assert el2.tolist() == [[1, 5, 9, 6], [2, 7, 9, 5], [3, 8, 9, 7], [4, 6, 9, 8]]
    assert irr2.shape == (0, 3)
    # refining one of two neighbours leaves a hanging node on the shared edge,
    # listed with the edge's sorted node pair
    c = np.array([[0.0, 0.0], [8.0, 0.0], [8.0, 8.0], [0.0, 8.0], [16.0, 0.0], [16.0, 8.0]])
    el = np.array([[1, 2, 3, 4], [2, 5, 6, 3]])
    c2, el2, irr2 = q.QrefineR(c, el, np.zeros((0, 3), np.int64), [1])
    # new nodes 7, 8, 9, 10 on edges 1-2, 1-4, 2-3, 3-4 (in edge order)
    assert irr2.tolist() == [[2, 3, 9]]
    assert np.allclose(c2[6:10], [[4, 0], [0, 4], [8, 4], [4, 8]])
