"""GPU checks of the global FE step (Subpb2Quadtree.m) through the C ABI
against the FP64 oracle (oracle/fem_oracle.py, direct sparse solve).

Tolerance: |Uhat - Uhat_oracle| <= 1e-9 (relative to max|U|): the library
assembles the same element integrals in FP64 and solves by conjugate
gradients to a relative residual of 1e-13, the oracle solves directly.
"""
import fem_oracle as fo
import numpy as np
import pytest

from _cases import quadtree_mesh

pytestmark = pytest.mark.gpu


def _fields(coords, seed):
    rng = np.random.default_rng(seed)
    n = len(coords)
    x, y = coords[:, 0], coords[:, 1]
    U = np.empty(2 * n)
    U[0::2] = 0.5 + 0.01 * x + 0.3 * np.sin(y / 9.0) + rng.normal(scale=0.05, size=n)
    U[1::2] = -0.2 + 0.004 * y + rng.normal(scale=0.05, size=n)
    F = rng.normal(scale=0.01, size=4 * n)
    return U, F


@pytest.mark.parametrize("order", [2, 3, 4, 5])
def test_subpb2_matches_oracle(pkg, order):
    coords, el, bnd = quadtree_mesh(7, 6, 4.0, refine=[(2, 2), (3, 2), (1, 4), (5, 1)])
    U, F = _fields(coords, order)
    mu = 1e-3
    for beta in (1e-3 * 64 * mu, 1e-1 * 64 * mu):
        mesh = pkg.FEMesh(coords, el, bnd, None, order)
        got, it = mesh.solve(beta, mu, U, F, 0.0, 8.0, return_iterations=True)
        want = fo.subpb2(coords, el, bnd, None, order, beta, mu, U, F, 0.0, 8.0)
        err = np.max(np.abs(got - want)) / np.max(np.abs(U))
        print(f"order {order} beta {beta:.2e}: {it} CG iterations, max rel err {err:.2e}")
        assert err < 1e-9
        mesh.close()


def test_subpb2_alpha_neumann_and_reproducible(pkg):
    coords, el, bnd = quadtree_mesh(6, 5, 4.0, refine=[(1, 1), (4, 3)])
    U, F = _fields(coords, 7)
    n = len(coords)
    # Neumann edges along the bottom boundary (node pairs), outward normal -y
    bottom = np.flatnonzero(np.isclose(coords[:, 1], coords[:, 1].min()))
    bottom = bottom[np.argsort(coords[bottom, 0])] + 1
    nm = np.array([[bottom[k], bottom[k + 1], 0.0, -1.0] for k in range(len(bottom) - 1)])
    top = np.flatnonzero(np.isclose(coords[:, 1], coords[:, 1].max())) + 1
    mesh = pkg.FEMesh(coords, el, top, nm, 2)
    a = mesh.solve(5e-3, 1e-3, U, F, alpha=2e-3, winstepsize=8.0)
    b = mesh.solve(5e-3, 1e-3, U, F, alpha=2e-3, winstepsize=8.0)
    assert np.array_equal(a, b)
    want = fo.subpb2(coords, el, top, nm, 2, 5e-3, 1e-3, U, F, 2e-3, 8.0)
    assert np.max(np.abs(a - want)) / np.max(np.abs(U)) < 1e-9
    info = mesh.info()
    # every node is in an element and unique(elementsFEM) starts with the 0
    # of a missing midpoint, so only the Dirichlet nodes are not free
    assert info["dirichlet"] == len(top) and info["free"] == n - len(top)
    mesh.close()


def test_subpb2_patch_test_and_driver(pkg):
    coords, el, bnd = quadtree_mesh(8, 7, 4.0, refine=[(3, 3), (4, 3), (2, 5)])
    n = len(coords)
    B = np.array([[0.012, -0.021], [0.033, 0.017]])
    U = np.empty(2 * n)
    U[0::2] = 0.7 + coords @ B[0]
    U[1::2] = -1.3 + coords @ B[1]
    F = np.tile([B[0, 0], B[1, 0], B[0, 1], B[1, 1]], n)
    DICmesh = {"coordinatesFEM": coords, "elementsFEM": el, "dirichlet": bnd, "neumann": []}
    Uh = pkg.Subpb2Quadtree(DICmesh, 2, 0.064e-3, 1e-3, U, F, np.zeros(4 * n), np.zeros(2 * n),
                            0.0, 8.0, 0)
    assert np.max(np.abs(Uh - U)) < 1e-9


def test_subpb2_larger_mesh(pkg):
    refine = [(i, j) for i in range(5, 25) for j in range(8, 20) if (i + j) % 3 == 0]
    coords, el, bnd = quadtree_mesh(30, 24, 4.0, refine=refine)
    U, F = _fields(coords, 3)
    mesh = pkg.FEMesh(coords, el, bnd, None, 2)
    got, it = mesh.solve(0.064e-3, 1e-3, U, F, 0.0, 8.0, return_iterations=True)
    want = fo.subpb2(coords, el, bnd, None, 2, 0.064e-3, 1e-3, U, F, 0.0, 8.0)
    print(f"{len(coords)} nodes, {len(el)} elements, {it} CG iterations")
    assert np.max(np.abs(got - want)) / np.max(np.abs(U)) < 1e-9
    mesh.close()


def test_subpb2_iteration_limit_and_replays(pkg):
    """Iteration limits on a small mesh (the one-workgroup CG): a limit of 3
    stops after exactly 3 iterations, twice with the same result, the
    default limit then gives the converged result of the host entry point,
    and a limit of 5 after it stops at 5."""
    from aldic_amd import api
    coords, el, bnd = quadtree_mesh(9, 7, 4.0, refine=[(2, 2), (5, 3), (6, 5)])
    U, F = _fields(coords, 11)
    n = len(coords)
    mesh = pkg.FEMesh(coords, el, bnd, None, 2)
    ctx = pkg.default_context()
    dU = api.DeviceArray.from_host(ctx, U, np.float64)
    dF = api.DeviceArray.from_host(ctx, F, np.float64)
    dO = api.DeviceArray(ctx, 2 * n, np.float64)
    it3 = mesh.solve_dev(ctx, 1e-3, 1e-3, dU.ptr, dF.ptr, dO.ptr, 0.0, 8.0, max_iter=3)
    a3 = dO.to_host()
    assert it3 == 3
    assert mesh.solve_dev(ctx, 1e-3, 1e-3, dU.ptr, dF.ptr, dO.ptr, 0.0, 8.0, max_iter=3) == 3
    assert np.array_equal(dO.to_host(), a3)
    itc = mesh.solve_dev(ctx, 1e-3, 1e-3, dU.ptr, dF.ptr, dO.ptr, 0.0, 8.0)
    conv = dO.to_host()
    want, itw = mesh.solve(1e-3, 1e-3, U, F, 0.0, 8.0, return_iterations=True)
    assert itc == itw > 3 and np.array_equal(conv, want)
    assert not np.allclose(a3, conv, rtol=0, atol=1e-9 * np.abs(U).max())
    assert mesh.solve_dev(ctx, 1e-3, 1e-3, dU.ptr, dF.ptr, dO.ptr, 0.0, 8.0, max_iter=5) == 5
    ref = fo.subpb2(coords, el, bnd, None, 2, 1e-3, 1e-3, U, F, 0.0, 8.0)
    assert np.max(np.abs(conv - ref)) / np.max(np.abs(U)) < 1e-9
    mesh.close()


def test_subpb2_async_graph_replays(pkg):
    """The solve without host synchronisation: eager it gives the
    synchronous solve's result bit for bit and the same iteration count (in
    a device int); recorded as a graph and replayed on new U (written
    between replays, with other device work in between) every replay equals
    the synchronous solve of that U; a limit below the count stops there."""
    from aldic_amd import api
    coords, el, bnd = quadtree_mesh(8, 6, 4.0, refine=[(2, 2), (3, 2), (5, 4)])
    n = len(coords)
    ctx = pkg.default_context()
    mesh = pkg.FEMesh(coords, el, bnd, None, 2)
    DA = api.DeviceArray
    U, F = _fields(coords, 21)
    dU = DA.from_host(ctx, U, np.float64)
    dF = DA.from_host(ctx, F, np.float64)
    dS = DA(ctx, 2 * n, np.float64)
    dA = DA(ctx, 2 * n, np.float64)
    dIt = DA(ctx, 1, np.int32)
    args = (2e-3, 1e-3)
    it = mesh.solve_dev(ctx, *args, dU.ptr, dF.ptr, dS.ptr, 0.0, 8.0)
    mesh.solve_dev_async(ctx, *args, dU.ptr, dF.ptr, dA.ptr, 0.0, 8.0, max_iter=64,
                         d_iterations=dIt.ptr)
    assert 3 < it < 64 and dIt.to_host()[0] == it
    assert np.array_equal(dA.to_host(), dS.to_host())
    dG = DA(ctx, 2 * n, np.float64)
    g = ctx.capture(lambda: mesh.solve_dev_async(ctx, *args, dU.ptr, dF.ptr, dG.ptr, 0.0, 8.0,
                                                 max_iter=64, d_iterations=dIt.ptr))
    other = DA(ctx, 2 * n, np.float64)
    rng = np.random.default_rng(4)
    for rep in range(3):
        Ur = U + rng.normal(scale=0.05, size=2 * n)
        pkg.lib().aldic_memcpy_htod(ctx.ptr, dU.ptr, Ur.ctypes.data, Ur.nbytes)
        g.launch(ctx)
        ctx.synchronize()
        want, itw = mesh.solve(*args, Ur, F, 0.0, 8.0, return_iterations=True)
        assert np.array_equal(dG.to_host(), want), rep
        assert dIt.to_host()[0] == itw
        # eager work on the same mesh between replays
        mesh.solve_dev(ctx, *args, dU.ptr, dF.ptr, other.ptr, 0.0, 8.0)
    g.close()
    mesh.solve_dev_async(ctx, *args, dU.ptr, dF.ptr, dA.ptr, 0.0, 8.0, max_iter=5,
                         d_iterations=dIt.ptr)
    assert dIt.to_host()[0] == 5
    with pytest.raises(pkg.AldicError):
        mesh.solve_dev_async(ctx, *args, dU.ptr, dF.ptr, dA.ptr, 0.0, 8.0, max_iter=0)
    mesh.close()


def test_subpb2_chunked_path_on_a_large_mesh(pkg):
    """Meshes of more than 8192 nodes take the chunked multi-block CG (the
    smaller ones above run the one-workgroup kernel): the solve matches the
    oracle to 1e-9, the asynchronous solve equals the synchronous one bit
    for bit with the same count, and an iteration limit stops it there."""
    from aldic_amd import api
    refine = [(i, j) for i in range(10, 90, 7) for j in range(5, 80, 6)]
    coords, el, bnd = quadtree_mesh(100, 85, 4.0, refine=refine)
    n = len(coords)
    assert n > 8192
    U, F = _fields(coords, 31)
    ctx = pkg.default_context()
    mesh = pkg.FEMesh(coords, el, bnd, None, 2)
    DA = api.DeviceArray
    dU = DA.from_host(ctx, U, np.float64)
    dF = DA.from_host(ctx, F, np.float64)
    dS = DA(ctx, 2 * n, np.float64)
    dA = DA(ctx, 2 * n, np.float64)
    dIt = DA(ctx, 1, np.int32)
    args = (1e-3, 1e-3)
    it = mesh.solve_dev(ctx, *args, dU.ptr, dF.ptr, dS.ptr, 0.0, 8.0)
    want = fo.subpb2(coords, el, bnd, None, 2, *args, U, F, 0.0, 8.0)
    print(f"{n} nodes, {it} CG iterations")
    assert np.max(np.abs(dS.to_host() - want)) / np.max(np.abs(U)) < 1e-9
    mesh.solve_dev_async(ctx, *args, dU.ptr, dF.ptr, dA.ptr, 0.0, 8.0, max_iter=512,
                         d_iterations=dIt.ptr)
    assert dIt.to_host()[0] == it and np.array_equal(dA.to_host(), dS.to_host())
    assert mesh.solve_dev(ctx, *args, dU.ptr, dF.ptr, dA.ptr, 0.0, 8.0, max_iter=5) == 5
    mesh.close()
