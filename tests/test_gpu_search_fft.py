"""The integer search with FFT lag blocks (aldic_ctx_set_search_method
ALDIC_SEARCH_FFT) against the FP64 oracle and against the direct lag tiles.

Bar: integer ZNCC peaks bit-exact against the oracle (funIntegerSearch.m:
166-169), sub-pixel u, v within 1e-3 + 1e-9 (findpeak rounds to 1e-3 and the
FP32 coefficients can land on the other side of a rounding boundary), peak
coefficients within 1e-4.  Between the two device methods the integer peaks
must agree; the sub-pixel fit is the same direct code in both, so u, v, x, y
are then identical, and the peak value within 1e-5 (a peak on the lag
window's edge, which is not fitted, reports its tile's own coefficient)."""
import numpy as np
import pytest

from _cases import c1_pair, small_pair

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctxs(pkg):
    api = pkg.api
    return api.Context(0).set_search_method("fft"), api.Context(0).set_search_method("direct")


def _search(pkg, ctx, f, g, region, gx, gy, ws, step):
    return pkg.funIntegerSearch(f, g, region, gx, gy, ws, step, ctx=ctx, return_peaks=True)


def _check(pkg, orc, ctxs, f, g, region, gx, gy, ws, step, label):
    fctx, dctx = ctxs
    x, y, u, v, cc, pr, pc = _search(pkg, fctx, f, g, region, gx, gy, ws, step)
    xd, yd, ud, vd, ccd, prd, pcd = _search(pkg, dctx, f, g, region, gx, gy, ws, step)
    r = np.concatenate([region, region])
    s = orc.integer_search(f, g, gx, gy, ws, step, r, nthreads=8)
    print(f"{label}: {u.size} points, FFT peaks differ from the oracle at "
          f"{int(np.sum((pr != s['ipeak_row']) | (pc != s['ipeak_col'])))}, from direct at "
          f"{int(np.sum((pr != prd) | (pc != pcd)))}")
    assert np.array_equal(x, s["x"]) and np.array_equal(y, s["y"])
    assert np.array_equal(pr, s["ipeak_row"]) and np.array_equal(pc, s["ipeak_col"])
    assert np.nanmax(np.abs(u - s["u"])) <= 1e-3 + 1e-9
    assert np.nanmax(np.abs(v - s["v"])) <= 1e-3 + 1e-9
    assert np.nanmax(np.abs(cc["max"] - s["ccmax"])) < 1e-4
    assert np.array_equal(pr, prd) and np.array_equal(pc, pcd)
    for a, b in ((u, ud), (v, vd)):
        assert np.array_equal(a, b, equal_nan=True)
    # a peak without a sub-pixel fit reports its block's own coefficient
    assert np.nanmax(np.abs(cc["max"] - ccd["max"])) < 1e-5


def test_fft_search_c1(pkg, orc, ctxs):
    """512 x 512, 31 px templates, radius 10 (one block per point)."""
    f, g, _, _ = c1_pair()
    _check(pkg, orc, ctxs, f, g, [10, 10], [1, 512], [1, 512], 30, 16, "C1")


def test_fft_search_edges_and_radius(pkg, orc, ctxs):
    """Clamped windows at the borders, unequal radii, several template
    sides (the FFT blocks take all of them)."""
    f, g, _, _ = small_pair(M=200, N=190, seed=3, shift=(3.4, -2.2))
    for region, ws, step in (([8, 8], 20, 12), ([5, 9], 16, 10), ([12, 12], 32, 16),
                             ([20, 7], 8, 9)):
        _check(pkg, orc, ctxs, f, g, region, [1, 200], [1, 190], ws, step, f"ws {ws} r {region}")


def test_fft_search_large_radius(pkg, orc, synth, ctxs):
    """65 px templates at radius 50 (103 x 103 lags: four 64 x 64 blocks),
    a 25 px disparity with strain, the grid reaching the borders."""
    M, N = 720, 600
    uf = lambda X, Y: 25.0 + 0.004 * (X - M / 2)  # noqa: E731
    vf = lambda X, Y: -3.0 + 0.002 * (Y - N / 2)  # noqa: E731
    f8, g8, _ = synth.make_pair(M, N, uf, vf, density=0.01, seed=5, bits=12)
    f = np.asfortranarray((f8 - f8.mean()) / f8.std(ddof=1))
    g = np.asfortranarray((g8 - g8.mean()) / g8.std(ddof=1))
    _check(pkg, orc, ctxs, f, g, [50, 50], [1, M], [1, N], 64, 48, "65 px, r 50")


def test_fft_search_larger_template_falls_back_to_direct(pkg, orc, ctxs):
    """A 81 px template does not fit the 128-point blocks: the direct tiles
    run (same results as the direct context)."""
    f, g, _, _ = small_pair(M=260, N=250, seed=4, shift=(2.0, 1.0))
    _check(pkg, orc, ctxs, f, g, [6, 6], [1, 260], [1, 250], 80, 40, "81 px")


def test_fft_search_node_list(pkg, ctxs):
    """The node-centred search (learned-start fallback) with FFT blocks
    equals the direct one."""
    api = pkg.api
    fctx, dctx = ctxs
    f, g, uf, vf = small_pair(M=300, N=280, seed=6, shift=(4.0, -3.0), strain=0.002)
    M, N = f.shape
    rng = np.random.default_rng(3)
    n = 150
    coords = np.asfortranarray(np.stack([rng.integers(20, M - 20, n),
                                         rng.integers(20, N - 20, n)], 1).astype(float))
    out = []
    for ctx in (fctx, dctx):
        im = api.Images(f, g, ctx=ctx)
        DA = api.DeviceArray
        d_c = DA.from_host(ctx, coords, np.float64)
        d_list = DA.from_host(ctx, np.arange(n, dtype=np.int32), np.int32)
        d_cnt = DA.from_host(ctx, np.array([n], np.int32), np.int32)
        d_U0 = DA.from_host(ctx, np.zeros(2 * n), np.float32)
        api.search_nodes_dev(im, d_c.ptr, n, d_list.ptr, d_cnt.ptr, n, 32, [12, 12, 12, 12],
                             d_U0.ptr, ctx=ctx)
        out.append(d_U0.to_host())
        im.close()
    U = out[0]
    ok = np.isfinite(U)
    print(f"node list: {n} nodes, |U0 fft - U0 direct| max {np.abs(U[ok] - out[1][ok]).max():.2e}")
    assert np.array_equal(out[0], out[1], equal_nan=True)
    err = np.abs(U[0::2] - 4.0 - 0.002 * (coords[:, 0] - M / 2))
    assert np.median(err[np.isfinite(err)]) < 0.1


def test_auto_is_the_default(pkg, ctxs):
    """A new context searches with ALDIC_SEARCH_AUTO: FFT blocks for a grid
    of 1600 points with 103 x 103 lag windows, direct sums for 23 x 23 ones
    and for a grid of fewer than 1024 points (bit-identical outputs to the
    explicit methods)."""
    fctx, dctx = ctxs
    actx = pkg.api.Context(0)
    f, g, _, _ = c1_pair()
    for region, step, want in (([50, 50], 12, fctx), ([10, 10], 12, dctx), ([50, 50], 16, dctx)):
        a = _search(pkg, actx, f, g, region, [1, 512], [1, 512], 16, step)
        b = _search(pkg, want, f, g, region, [1, 512], [1, 512], 16, step)
        print(f"auto, r {region[0]}, step {step}: {a[2].size} points")
        for x, y in zip(a[:4] + a[5:], b[:4] + b[5:]):
            assert np.array_equal(x, y, equal_nan=True)
        assert np.array_equal(a[4]["max"], b[4]["max"], equal_nan=True)
    with pytest.raises(ValueError):
        actx.set_search_method("spatial")
