"""Edge cases of the local matching path through the C ABI: empty node
sets and search grids, argument errors, non-finite starts, and the largest
subsets (the workgroup kernel's 8192-pixel window, 90 x 90 px).

Tolerances as tests/test_gpu_parity.py: |dU| <= 1e-3 px against the FP64
oracle on the subsets both converge."""
import numpy as np
import pytest

from _cases import small_pair

pytestmark = pytest.mark.gpu


def test_empty_node_sets_and_grids(pkg):
    f, g, _, _ = small_pair()
    im = pkg.Images(f, g)
    c0 = np.zeros((0, 2))
    U, F, H, conv, hole, z = pkg.local_icgn_raw(im, c0, np.zeros(0), 30, 1e-3)
    assert U.shape == (0,) and F.shape == (0,) and H.shape == (0, 21) and conv.shape == (0,)
    U1, conv1 = pkg.subpb1_raw(im, c0, np.zeros(0), np.zeros(0), np.zeros(0), np.zeros(0),
                               np.zeros((0, 2), np.int32), 1e-3, 1e-2)
    assert U1.shape == (0,) and conv1.shape == (0,)
    Df = pkg.funImgGradient(f, g)
    Ud, Fd, Hd, _, convd, badnum, holed = pkg.LocalICGNQuadtree(
        np.zeros(0), c0, Df, f, g, {"winsize": 30, "winstepsize": 16}, "GaussNewton", 1e-3, 1)
    assert Ud.shape == (0,) and badnum == 0 and holed.size == 0
    # a ROI smaller than one template: no search points
    M, N = f.shape
    x, y, u, v, cc = pkg.funIntegerSearch(f, g, [5, 5], [1, 40], [1, N], 30, 16)
    assert u.size == 0 and x.size == 0
    # interpolation at no points
    assert pkg.ba_interp2_spline(f, np.zeros(0), np.zeros(0)).size == 0
    im.close()


def test_argument_errors(pkg):
    f, g, _, _ = small_pair()
    im = pkg.Images(f, g)
    coords = np.array([[60.0, 60.0], [80.0, 70.0]])
    with pytest.raises(pkg.AldicError):  # odd winsize (validateDICpara.m:22-28)
        pkg.local_icgn_raw(im, coords, np.zeros(4), 29, 1e-3)
    with pytest.raises(pkg.AldicError):
        pkg.funIntegerSearch(f, g, [5, 5], [1, 160], [1, 144], 29, 16)
    with pytest.raises(ValueError):  # winsize neither scalar nor per node
        pkg.local_icgn_raw(im, coords, np.zeros(4), [30, 30, 30], 1e-3)
    im.close()


def test_nonfinite_start_is_a_throw(pkg, orc):
    """A NaN start makes the reference's interpolation throw inside its try
    (LocalICGNQuadtree.m:99-104): convIt -1 and NaN U.  An infinite one
    fails the image-bounds test of the first evaluation instead
    (funICGNQuadtree.m:154-157): status 101 with P left at the start.  Both
    as the oracle."""
    f, g, _, _ = small_pair(shift=(0.4, -0.3))
    im = pkg.Images(f, g)
    coords = np.array([[60.0, 60.0], [80.0, 70.0], [100.0, 80.0]])
    U0 = np.array([0.0, 0.0, np.nan, 0.0, 0.0, np.inf])
    U, F, H, conv, hole, z = pkg.local_icgn_raw(im, coords, U0, 30, 1e-3)
    Uo, Fo, Ho, convo, holeo, zo = orc.local_icgn(orc.Images(f, g), coords, U0,
                                                  np.full(3, 30, np.int32), 1e-3)
    assert np.array_equal(conv, convo)
    assert conv[0] >= 1 and conv[1] == -1 and conv[2] == 101
    assert np.all(np.isnan(U[2:4])) and U[4] == 0.0 and np.isinf(U[5])
    assert abs(U[0] - Uo[0]) <= 1e-3 and abs(U[1] - Uo[1]) <= 1e-3
    im.close()


def test_largest_subsets(pkg, orc):
    """Windows up to 89 x 89 px run on the workgroup kernel and match the
    oracle; a 97 x 97 window exceeds its 8192-pixel LDS window and is
    reported as a throw (convIt -1, NaN U) instead of a wrong answer."""
    f, g, uf, vf = small_pair(M=220, N=210, seed=9, shift=(0.7, -0.45), strain=0.001)
    im = pkg.Images(f, g)
    oim = orc.Images(f, g)
    coords = np.array([[110.0, 105.0], [112.0, 100.0], [108.0, 110.0]])
    ws = np.array([70, 88, 96], np.int32)
    U0 = np.zeros(6)
    U0[0::2], U0[1::2] = 1.0, 0.0
    U, F, H, conv, hole, z = pkg.local_icgn_raw(im, coords, U0, ws, 1e-3)
    Uo, Fo, Ho, convo, holeo, zo = orc.local_icgn(oim, coords, U0, ws, 1e-3)
    ok = (conv[:2] >= 1) & (conv[:2] <= 100) & (convo[:2] >= 1) & (convo[:2] <= 100)
    assert ok.all()
    assert np.abs(U[:4] - Uo[:4]).max() <= 1e-3
    assert conv[2] == -1 and np.all(np.isnan(U[4:]))
    assert convo[2] >= 1  # the FP64 oracle has no such limit
    im.close()


def _one_dimensional_band(f, g, cols=60, shift=0.4):
    """Columns below `cols` replaced by a texture that varies along dim 1
    only (g shifted by `shift` px): subsets there have no y gradient."""
    M = f.shape[0]
    x = np.arange(1, M + 1, dtype=np.float64)

    def prof(t):
        return np.sin(0.7 * t) + 0.5 * np.sin(0.23 * t + 1.0)

    f = np.array(f)
    g = np.array(g)
    f[:, :cols] = prof(x)[:, None]
    g[:, :cols] = prof(x - shift)[:, None]
    return np.asfortranarray(f), np.asfortranarray(g)


def test_singular_hessian_subsets(pkg, orc):
    """Subsets whose texture varies along one axis only have a numerically
    singular H: its y entries are the gradient filter's rounding residue
    (about 1e-29).  The FP64 oracle and the FP32 fast kernels both factor it,
    take a step that leaves the image and report status 101; a textured
    subset in the same launch matches the oracle.  (An H whose Cholesky
    fails in FP32 goes to the workgroup kernel's LU; no natural input here
    reaches that branch.)"""
    f0, g0, _, _ = small_pair(shift=(0.4, -0.3))
    f, g = _one_dimensional_band(f0, g0)
    im = pkg.Images(f, g)
    oim = orc.Images(f, g)
    coords = np.array([[60.0, 30.0], [80.0, 28.0], [80.0, 100.0]])
    U0 = np.zeros(6)
    for ws in (30, 16):  # the 33 px and 17 px fast-kernel classes
        U, F, H, conv, hole, z = pkg.local_icgn_raw(im, coords, U0, ws, 1e-3)
        Uo, Fo, Ho, convo, holeo, zo = orc.local_icgn(oim, coords, U0, np.full(3, ws, np.int32),
                                                      1e-3)
        failed = lambda c: (c < 1) | (c > 100)
        assert failed(conv[:2]).all() and failed(convo[:2]).all(), (ws, conv, convo)
        assert (conv[:2] == 101).all() and (convo[:2] == 101).all(), (ws, conv, convo)
        assert 1 <= conv[2] <= 100 and 1 <= convo[2] <= 100
        assert np.abs(U[4:6] - Uo[4:6]).max() <= 1e-3
    im.close()
