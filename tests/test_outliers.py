"""CPU tests of the integer search's host post-processing: inpaint_nans
(methods 3 and 4), medFilt2, funRemoveOutliers (func/inpaint_nans.m,
func/funRemoveOutliers.m) and the host rules of IntegerSearchQuadtree
(radius choice, mask -> NaN; func_quadtree/IntegerSearchQuadtree.m:52-160).

The dense restatement in oracle/inpaint_oracle.py is pinned by closed forms
(below); the product's sparse version must agree with it to 1e-9 on random
NaN patterns, small grids and border holes included."""
import numpy as np
import pytest

import inpaint_oracle as io


@pytest.fixture(scope="module")
def out(pkg):
    from aldic_amd import outliers
    return outliers


def _both(out, A, method):
    return out.inpaint_nans(A, method), io.inpaint_nans(A, method)


def test_del4_template_is_the_laplacian_squared():
    from scipy.signal import convolve2d
    assert np.array_equal(convolve2d(io.DEL2, io.DEL2), io.DEL4)


def test_springs_closed_forms(out):
    rng = np.random.default_rng(0)
    C = rng.normal(size=(7, 8))
    C[3, 4] = np.nan  # a lone NaN: the mean of its four neighbours
    want = (C[2, 4] + C[4, 4] + C[3, 3] + C[3, 5]) / 4
    for got in _both(out, C, 4):
        assert got[3, 4] == pytest.approx(want, abs=1e-12)
        assert np.array_equal(np.delete(got.ravel(), 3 * 8 + 4), np.delete(C.ravel(), 3 * 8 + 4))
    r = np.array([[1.0, np.nan, np.nan, np.nan, 5.0, np.nan]])  # a run, and an open end
    for got in _both(out, r, 4):
        assert np.allclose(got, [[1, 2, 3, 4, 5, 5]], atol=1e-12)


def test_del4_closed_forms(out):
    i, j = np.meshgrid(np.arange(12.0), np.arange(10.0), indexing="ij")
    lin = 5 + 0.3 * i - 0.7 * j
    B = lin.copy()
    B[4:7, 3:6] = np.nan
    B[0, 0] = B[11, 4] = B[6, 9] = np.nan  # border holes too
    for got in _both(out, B, 3):
        assert np.abs(got - lin).max() < 1e-10
    i, j = np.meshgrid(np.arange(14.0), np.arange(13.0), indexing="ij")
    quad = 0.1 * i ** 2 + 0.2 * j + 0.05 * i * j - 0.03 * j ** 2
    Bq = quad.copy()
    Bq[5:9, 5:8] = np.nan  # every equation touching the hole is del^4
    for got in _both(out, Bq, 3):
        assert np.abs(got - quad).max() < 1e-9
    # the springs do not reproduce a quadratic: the two methods differ
    assert np.abs(out.inpaint_nans(Bq, 4) - quad).max() > 1e-2


@pytest.mark.parametrize("shape", [(9, 11), (3, 4), (1, 9), (6, 1), (20, 17)])
def test_inpaint_matches_the_dense_restatement(out, shape):
    rng = np.random.default_rng(sum(shape))
    for trial in range(4):
        A = rng.normal(size=shape)
        nanmask = rng.random(shape) < (0.15 + 0.1 * trial)
        if nanmask.all():
            nanmask.flat[0] = False
        A[nanmask] = np.nan
        for method in (3, 4):
            try:
                want = io.inpaint_nans(A, method)
            except np.linalg.LinAlgError:
                continue
            # rank-deficient cases (a NaN no equation reaches) are not compared
            if not np.all(np.isfinite(want)):
                continue
            got = out.inpaint_nans(A, method)
            assert got.shape == A.shape
            assert np.abs(got - want).max() < 1e-9, (shape, trial, method)


def test_medfilt2_against_loops(out):
    rng = np.random.default_rng(3)
    for shape in ((5, 7), (1, 6), (6, 1), (2, 2)):
        V = rng.normal(size=shape)
        assert np.array_equal(out.medFilt2(V), io.medFilt2(V))


def test_remove_outliers(out):
    """A smooth field with two spikes and a hole: the spikes exceed the
    threshold and are refilled by the springs, the hole stays NaN, the rest
    is untouched; without spikes the grids come back unchanged."""
    i, j = np.meshgrid(np.arange(14.0), np.arange(12.0), indexing="ij")
    u = 1.0 + 0.02 * i - 0.01 * j
    v = -0.5 + 0.01 * i
    u[5, 6] += 40.0
    v[9, 2] -= 30.0
    u[2, 2] = v[2, 2] = np.nan
    u2, v2, cc, rm = out.funRemoveOutliers(u, v, {"max": None}, 0, 100)
    wu, wv, wout = io.funRemoveOutliers(u, v, 100.0)
    assert set(rm) == set(np.flatnonzero(wout.reshape(-1, order="F")))
    assert {5 + 6 * 14, 9 + 2 * 14} <= set(rm)
    assert np.isnan(u2[2, 2]) and np.isnan(v2[2, 2])
    fin = np.isfinite(wu)
    assert np.abs(u2[fin] - wu[fin]).max() < 1e-9 and np.abs(v2[fin] - wv[fin]).max() < 1e-9
    assert abs(u2[5, 6] - (1.0 + 0.1 - 0.06)) < 0.05
    u0 = 1.0 + 0.02 * i
    v0 = -0.5 + 0.01 * i
    u3, v3, _, rm3 = out.funRemoveOutliers(u0, v0, {}, 0, 100)
    assert rm3.size == 0 and np.array_equal(u3, u0) and np.array_equal(v3, v0)


def test_integer_search_quadtree_host_rules(pkg, monkeypatch):
    """The radius each ImgSeqNum takes, the mask -> NaN rule and its catch,
    with the device search replaced by a recorder (host logic only)."""
    api = pkg.api
    seen = {}
    nY, nX = 6, 5

    def fake_search(f, g, r, gridx, gridy, ws, step, tempNoOfInitPt=0, ctx=None):
        seen["r"] = np.asarray(r).copy()
        xs = 1 + ws / 2 + step * np.arange(nX)
        ys = 1 + ws / 2 + step * np.arange(nY)
        x, y = np.meshgrid(xs, ys)
        u = 0.5 + 0.0 * x
        v = -0.25 + 0.0 * x
        return x, y, u, v, {"max": np.ones_like(x), "A": []}

    monkeypatch.setattr(api, "funIntegerSearch", fake_search)
    base = {"gridxyROIRange": {"gridx": [1, 200], "gridy": [1, 220]}, "winsize": 20,
            "winstepsize": 16, "NewFFTSearchDistance": 7}
    cases = [(-1, {"StereoSearchDistance": 30}, [30, 30]),
             (-1, {"StereoSearchDistance": [40, 12]}, [40, 12]),
             (-1, {"StereoSearchDistance": []}, [7, 7]),          # empty: the fallback
             (2, {"TemporalSearchDistance": 9}, [9, 9]),
             (3, {}, [7, 7]),
             (-1, {"StereoSearchDistance": 0}, [0, 0])]           # 0 is a preset, not empty
    for seq, extra, want in cases:
        P, x0, y0, u, v, cc = api.IntegerSearchQuadtree(None, None, None, {**base, **extra}, seq)
        assert list(seen["r"]) == want, (seq, extra)
        assert list(P["SizeOfFFTSearchRegion"]) == want and P["InitFFTSearchMethod"] == 1
        assert cc["ccThreshold"] == 1.25
        if seq >= 2:
            assert P["fixSearchDistanceOrNot"] == 0
    for seq in (0, 1):  # no preset for the first frame: the reference prompts
        with pytest.raises(ValueError):
            api.IntegerSearchQuadtree(None, None, None,
                                      {k: v for k, v in base.items() if k != "NewFFTSearchDistance"},
                                      seq)
    # mask: the point whose rounded centre is masked becomes NaN
    mask = np.ones((200, 220))
    x0c, y0c = 1 + 10 + 16 * 2, 1 + 10 + 16 * 3  # grid point (ci 2, cj 3)
    mask[x0c - 1, y0c - 1] = 0
    P, x0, y0, u, v, cc = api.IntegerSearchQuadtree(
        None, None, None, {**base, "ImgRefMask": mask, "ImgSize": (200, 220)}, -1)
    assert np.isnan(u[3, 2]) and np.isnan(v[3, 2])
    assert np.isfinite(u).sum() == nX * nY - 1
    # no ImgSize, or a centre outside it: the catch leaves u, v unmasked
    for extra in ({"ImgRefMask": mask}, {"ImgRefMask": mask, "ImgSize": (30, 30)}):
        P, x0, y0, u, v, cc = api.IntegerSearchQuadtree(None, None, None, {**base, **extra}, -1)
        assert np.isfinite(u).all()
