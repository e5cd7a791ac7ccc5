"""Pins of the Subpb2Quadtree oracle (oracle/fem_oracle.py), CPU only.

  * a square element without hanging nodes: the element matrices equal the
    closed-form bilinear mass h^2/36 [4 2 1 2; ...] and stiffness
    1/6 [4 -1 -2 -1; ...] matrices (2 x 2 Gauss integrates both exactly);
  * with hanging nodes: stiffness rows sum to zero and the mass sums to the
    element area (partition of unity of the enriched shape functions), and
    both matrices are symmetric positive (semi)definite;
  * the patch test: U linear, F its gradient, Dirichlet nodes on the
    boundary -- the solve returns U at every node, for several beta, mu,
    with hanging nodes.  A transposed F index (F12 for F21) or a sign slip
    in a derivative fails it;
  * the reference's bookkeeping: nodes in no element get 0, Dirichlet nodes
    keep U.
"""
import fem_oracle as fo
import numpy as np
import pytest

from _cases import quadtree_mesh


def test_square_element_closed_forms():
    h = 3.0
    pts = np.zeros((8, 2))
    pts[:4] = [[0, 0], [h, 0], [h, h], [0, h]]
    M, K, _, _ = fo.element_matrices(pts, np.zeros(4))
    M4 = h * h / 36 * np.array([[4, 2, 1, 2], [2, 4, 2, 1], [1, 2, 4, 2], [2, 1, 2, 4]])
    K4 = 1 / 6 * np.array([[4, -1, -2, -1], [-1, 4, -1, -2], [-2, -1, 4, -1], [-1, -2, -1, 4]])
    assert np.allclose(M[:4, :4], M4, atol=1e-12)
    assert np.allclose(K[:4, :4], K4, atol=1e-12)
    assert np.all(M[4:] == 0) and np.all(K[4:] == 0)


@pytest.mark.parametrize("delta", [(1, 0, 0, 0), (0, 1, 1, 0), (1, 1, 1, 1), (0, 0, 1, 1)])
@pytest.mark.parametrize("order", [2, 3])
def test_hanging_node_element_properties(delta, order):
    h = 4.0
    pts = np.array([[0, 0], [h, 0], [h, h], [0, h], [h, h / 2], [h / 2, h], [0, h / 2],
                    [h / 2, 0]], dtype=float)
    d = np.array(delta, dtype=float)
    pts[4:][d == 0] = 0.0
    M, K, Mfull, Kfull = fo.element_matrices(pts, d, order)
    live = np.concatenate([np.ones(4, bool), d > 0])
    Ml, Kl = M[np.ix_(live, live)], K[np.ix_(live, live)]
    assert np.allclose(Kl.sum(axis=1), 0, atol=1e-12)
    assert abs(Ml.sum() - h * h) < 1e-10
    assert np.allclose(Ml, Ml.T) and np.allclose(Kl, Kl.T)
    # semidefinite: with 2 x 2 points and several hanging nodes the
    # quadrature has fewer points than shape functions
    assert np.linalg.eigvalsh(Ml).min() > -1e-12
    assert np.linalg.eigvalsh(Kl).min() > -1e-12
    # the x and y blocks of the 16 x 16 matrices are the scalar matrix twice
    assert np.array_equal(Kfull[0::2, 1::2], np.zeros((8, 8)))
    assert np.array_equal(Mfull[1::2, 1::2], Mfull[0::2, 0::2])


@pytest.mark.parametrize("beta,mu", [(1e-3 * 64 * 1e-3, 1e-3), (1.0, 1e-3), (1e-2, 1.0)])
def test_patch_test_with_hanging_nodes(beta, mu):
    coords, el, bnd = quadtree_mesh(6, 5, 4.0, refine=[(2, 2), (3, 2), (1, 3)])
    assert np.any(el[:, 4:] != 0)
    n = len(coords)
    B = np.array([[0.01, -0.02], [0.03, 0.015]])  # [u_x u_y; v_x v_y]
    c = np.array([0.7, -1.3])
    U = np.empty(2 * n)
    U[0::2] = c[0] + coords @ B[0]
    U[1::2] = c[1] + coords @ B[1]
    F = np.tile([B[0, 0], B[1, 0], B[0, 1], B[1, 1]], n)  # [F11 F21 F12 F22]
    Uh = fo.subpb2(coords, el, bnd, None, 2, beta, mu, U, F, 0.0, 8.0)
    assert np.max(np.abs(Uh - U)) < 1e-10
    # a transposed gradient breaks it
    Ft = np.tile([B[0, 0], B[0, 1], B[1, 0], B[1, 1]], n)
    Ub = fo.subpb2(coords, el, bnd, None, 2, beta, mu, U, Ft, 0.0, 8.0)
    if beta > 1e-3:
        assert np.max(np.abs(Ub - U)) > 1e-4


def test_bookkeeping_unused_and_dirichlet_nodes():
    coords, el, bnd = quadtree_mesh(3, 3, 4.0)
    n = len(coords)
    coords = np.vstack([coords, [[200.0, 200.0]]])  # a node in no element
    rng = np.random.default_rng(0)
    U = rng.normal(size=2 * (n + 1))
    F = rng.normal(scale=0.01, size=4 * (n + 1))
    Uh = fo.subpb2(coords, el, bnd[:5], None, 2, 1e-3, 1e-3, U, F, 0.0, 8.0)
    assert Uh[2 * n] == 0 and Uh[2 * n + 1] == 0
    d = bnd[:5] - 1
    assert np.array_equal(Uh[2 * d], U[2 * d]) and np.array_equal(Uh[2 * d + 1], U[2 * d + 1])
