"""CPU tests that pin the FP64 oracle to what the mathematics and the
reference fix (the reference itself ships no golden vectors for this path
and its MEX needs MATLAB's mex.h, so it cannot be run here)."""
import numpy as np
import pytest

from _cases import c1_grid, c1_pair, hole_mask, small_pair

# func/BicubicBsplineBuild.m:9-10 (commented definition of B and C)
B = np.array([[-1, 3, -3, 1], [3, -6, 3, 0], [-3, 0, 3, 0], [1, 4, 1, 0]]) / 6.0
Cm = np.array([[71, -19, 5, -1], [-19, 95, -25, 5], [5, -25, 95, -19], [-1, 5, -19, 71]]) / 56.0
BC = B @ Cm


def closed_form(Q, dx, dy):
    """Value of A = rot90(BC*Q*BC',2) polynomial, A[i][j] dx^i dy^j."""
    A = np.rot90(BC @ Q @ BC.T, 2)
    return sum(A[i, j] * dx ** i * dy ** j for i in range(4) for j in range(4))


def test_bc_matches_reference_table():
    # the numeric BC the reference hard-codes (func/BicubicBsplineBuild.m:13-16)
    ref = np.array([[-0.428571428571429, 1.14285714285714, -1.14285714285714, 0.428571428571429],
                    [1.01785714285714, -2.08928571428571, 1.33928571428571, -0.267857142857143],
                    [-0.589285714285714, -0.0535714285714286, 0.803571428571429, -0.160714285714286],
                    [0, 1, 0, 0]])
    assert np.allclose(BC, ref, atol=1e-13)


def test_interp_matches_closed_form(orc):
    rng = np.random.default_rng(0)
    F = np.asfortranarray(rng.normal(size=(40, 37)))
    n = 500
    X = rng.uniform(3, 35, n)  # dim-2 coordinate
    Y = rng.uniform(3, 38, n)  # dim-1 coordinate
    Z = orc.interp_spline(F, X, Y)
    for k in range(n):
        xb, yb = int(np.floor(X[k])) - 1, int(np.floor(Y[k])) - 1  # 0-based base col/row
        Q = F[yb - 1:yb + 3, xb - 1:xb + 3]  # Q[r][c], r along dim 1
        ref = closed_form(Q, Y[k] - np.floor(Y[k]), X[k] - np.floor(X[k]))
        # the MEX's float-rounded literal scales differ from 1/3136 by < 5e-8
        assert abs(Z[k] - ref) < 1e-7 * (1 + np.abs(Q).sum())


def test_interp_float_literals_are_used(orc):
    # with the MEX's single-precision scales the double result is NOT the
    # exact rational value: a sample where only a33 (scale 1.0f/49) matters
    F = np.zeros((12, 12))
    F[5, 5] = 1.0
    Z = orc.interp_spline(F, np.array([6.5]), np.array([6.5]))[0]
    exact = closed_form(F[4:8, 4:8], 0.5, 0.5)
    assert Z != exact
    assert abs(Z - exact) < 1e-7


def test_interp_integer_nodes_and_constant(orc):
    rng = np.random.default_rng(1)
    F = np.asfortranarray(rng.normal(size=(30, 30)))
    j, i = np.meshgrid(np.arange(1, 31), np.arange(1, 31))
    Z = orc.interp_spline(F, j.astype(float), i.astype(float))
    assert np.array_equal(Z, F)  # a00 = Q11 exactly
    c = np.full((20, 20), 3.25)
    Zc = orc.interp_spline(c, rng.uniform(1, 20, 50), rng.uniform(1, 20, 50))
    assert np.allclose(Zc, 3.25, atol=1e-12)  # partition of unity


def test_interp_ramp_bias(orc):
    i = np.arange(1, 41, dtype=float)
    F = np.asfortranarray(np.repeat(i[:, None], 40, axis=1))  # F = dim-1 coordinate
    t = np.linspace(0.0, 0.95, 20)
    Z = orc.interp_spline(F, np.full(20, 20.0), 20.0 + t)
    e = t * (2 * t - 1) * (t - 1) / 14.0
    assert np.allclose(Z - (20.0 + t), e, atol=1e-7)


def test_interp_clamps_like_edge_padding(orc):
    rng = np.random.default_rng(2)
    F = np.asfortranarray(rng.normal(size=(16, 18)))
    X = rng.uniform(-3, 22, 300)
    Y = rng.uniform(-3, 20, 300)
    Z = orc.interp_spline(F, X, Y)
    pad = np.pad(F, 8, mode="edge")
    Zp = orc.interp_spline(pad, X + 8, Y + 8)
    assert np.allclose(Z, Zp, atol=1e-12)


def test_interp_single(orc):
    rng = np.random.default_rng(3)
    F = np.asfortranarray(rng.normal(size=(30, 30)))
    X = rng.uniform(3, 27, 200)
    Y = rng.uniform(3, 27, 200)
    Zd = orc.interp_spline(F, X, Y)
    Zs = orc.interp_spline(F.astype(np.float32), X.astype(np.float32), Y.astype(np.float32),
                           single=True)
    assert Zs.dtype == np.float32
    assert np.allclose(Zs, Zd, atol=2e-5)


def test_gradient_exact_on_polynomials(orc):
    M, N = 30, 26
    x = np.arange(1, M + 1, dtype=float)[:, None]
    y = np.arange(1, N + 1, dtype=float)[None, :]
    img = np.asfortranarray(0.001 * x ** 5 + 0.01 * (y ** 3) * x - 0.5 * y ** 2)
    gx, gy = orc.img_gradient(img)
    X, Y = x[3:-3], y[:, 3:-3]
    assert np.allclose(gx, 0.005 * X ** 4 + 0.01 * Y ** 3, rtol=1e-10, atol=1e-9)
    assert np.allclose(gy, 0.03 * (Y ** 2) * X - Y, rtol=1e-10, atol=1e-9)
    mask = np.ones((M, N))
    mask[10, 12] = 0
    gxm, _ = orc.img_gradient(img, mask)
    assert gxm[7, 9] == 0 and gxm[7, 10] == gx[7, 10]


def _brute_ncc(T, A):
    s1, s2 = T.shape
    W1, W2 = A.shape
    Ap = np.zeros((W1 + 2, W2 + 2))
    Ap[1:-1, 1:-1] = A
    out = np.zeros((W1 - s1 + 3, W2 - s2 + 3))
    Tc = T - T.mean()
    for p in range(out.shape[0]):
        for q in range(out.shape[1]):
            w = Ap[p:p + s1, q:q + s2]
            wc = w - w.mean()
            d = np.sqrt((Tc ** 2).sum() * (wc ** 2).sum())
            out[p, q] = (Tc * w).sum() / d if d > 0 else 0.0
    return out


def test_normxcorr2_crop_bruteforce(orc):
    rng = np.random.default_rng(4)
    T = rng.normal(size=(7, 7))
    A = rng.normal(size=(15, 13))
    c = orc.normxcorr2_crop(T, A)
    assert np.allclose(c, _brute_ncc(T, A), atol=1e-12)
    # invariance to A -> a*A + b (Pearson correlation)
    c2 = orc.normxcorr2_crop(T, 3.0 * A + 7.0)
    assert np.allclose(c2[1:-1, 1:-1], c[1:-1, 1:-1], atol=1e-12)
    with pytest.raises(ValueError):
        orc.normxcorr2_crop(np.ones((5, 5)), A)


def test_findpeak_quadratic(orc):
    x0, y0 = 0.3137, -0.2249
    r = np.arange(-5, 6, dtype=float)
    Y, X = np.meshgrid(r, r, indexing="ij")  # rows = y, columns = x
    f = 2.0 - 0.7 * (X - x0) ** 2 - 0.4 * (Y - y0) ** 2 + 0.1 * (X - x0) * (Y - y0)
    xp, yp, mf, pr, pc = orc.findpeak(f)
    assert (pr, pc) == (6, 6)
    assert xp == pytest.approx(6 + round(1000 * x0) / 1000, abs=1e-12)
    assert yp == pytest.approx(6 + round(1000 * y0) / 1000, abs=1e-12)
    # on the edge: integer peak
    g = np.zeros((5, 5))
    g[0, 2] = 1.0
    xp, yp, mf, pr, pc = orc.findpeak(g)
    assert (xp, yp, mf) == (3.0, 1.0, 1.0)
    # ties: first in column-major order
    h = np.zeros((5, 5))
    h[3, 1] = h[1, 3] = 1.0
    assert orc.findpeak(h)[3:] == (4, 2)


def test_integer_search_recovers_integer_shift(orc):
    rng = np.random.default_rng(5)
    base = rng.normal(size=(140, 130))
    from scipy.ndimage import gaussian_filter
    base = gaussian_filter(base, 1.5)
    f = np.asfortranarray(base[10:-10, 10:-10])
    g = np.asfortranarray(np.roll(np.roll(base, -3, axis=0), 2, axis=1)[10:-10, 10:-10])
    out = orc.integer_search(f, g, [1, 120], [1, 110], 20, 16, [6, 6, 6, 6])
    # g(i, j) = f(i - 3, j + 2): the match of f at i sits at i - 3 in g
    # (u = -3) and at j + 2 (v = +2).  The first grid row has ii = 1, its
    # left search reach is clamped to 0 (funIntegerSearch.m:140-142) and it
    # can only report the edge lag u = -1.
    inner = out["x"] > 11.0
    assert np.all(out["u"][~inner] == -1.0)
    assert np.all(np.round(out["u"][inner]) == -3.0)
    assert np.all(np.round(out["v"]) == 2.0)
    assert np.all(np.abs(out["u"][inner] + 3.0) < 0.1) and np.all(np.abs(out["v"] - 2.0) < 0.1)
    # the integer peak sits at lag index u + 2 + L (1-based), i.e. 5 for L = 6
    assert np.all(out["ipeak_row"][inner] == 5)
    assert out["x"][0, 0] == 11.0 and out["y"][0, 0] == 11.0


def test_icgn_zero_motion_and_translation(orc):
    f, _, _, _ = small_pair()
    coords = np.asfortranarray(np.array([[60.0, 60.0], [80.0, 70.0], [100.0, 90.0]]))
    im = orc.Images(f, f)
    U, F, H, conv, hole, z = orc.local_icgn(im, coords, np.zeros(6), 20, 1e-3)
    assert np.all(conv == 1) and np.all(U == 0) and np.all(F == 0)
    assert np.allclose(z, 1.0)
    # integer translation: g(x) = f(x - t) exactly, start from the truth
    g = np.asfortranarray(np.roll(np.roll(f, 2, axis=0), -1, axis=1))
    im2 = orc.Images(f, g)
    U, F, H, conv, hole, z = orc.local_icgn(im2, coords, np.tile([2.0, -1.0], 3), 20, 1e-3)
    assert np.all(conv == 1) and np.allclose(U, np.tile([2.0, -1.0], 3), atol=1e-12)
    # and from a sub-pixel-off start it converges back to it
    U, F, H, conv, hole, z = orc.local_icgn(im2, coords, np.tile([2.4, -0.6], 3), 20, 1e-3)
    assert np.all((conv > 1) & (conv < 101))
    assert np.allclose(U, np.tile([2.0, -1.0], 3), atol=2e-3)
    # H symmetric positive definite
    for k in range(3):
        Hm = np.zeros((6, 6))
        t = 0
        for c in range(6):
            for r in range(c, 6):
                Hm[r, c] = Hm[c, r] = H[k, t]
                t += 1
        assert np.all(np.linalg.eigvalsh(Hm) > 0)


def test_icgn_hessian_is_the_packed_steepest_descent_gram(orc):
    """H per node is sum over the subset of sd sd^T with sd = (dx fx, dx fy,
    dy fx, dy fy, fx, fy) (dx, dy: the pixel's offsets from the node; fx, fy:
    the gradient planes, which start 3 px into the image), packed as the
    reference's HGlobal: the lower triangle column by column
    (funICGNQuadtree.m:91-103, :451).  Built here with numpy from the
    oracle's own gradient planes; a reordered sd or a shifted gradient index
    changes it."""
    f, g, _, _ = small_pair(shift=(0.4, -0.3))
    im = orc.Images(f, g)
    for (x0, y0, ws) in ((70, 64, 20), (90, 80, 30)):
        U, F, H, conv, hole, z = orc.local_icgn(im, np.array([[float(x0), float(y0)]]),
                                                np.zeros(2), np.array([ws], np.int32), 1e-3)
        h = ws // 2
        ii = np.arange(x0 - h, x0 + h + 1)
        jj = np.arange(y0 - h, y0 + h + 1)
        DX, DY = np.meshgrid(ii - x0, jj - y0, indexing="ij")
        fx = im.DfDx[np.ix_(ii - 4, jj - 4)]
        fy = im.DfDy[np.ix_(ii - 4, jj - 4)]
        sd = np.stack([DX * fx, DX * fy, DY * fx, DY * fy, fx, fy]).reshape(6, -1)
        Hf = sd @ sd.T
        packed = np.concatenate([Hf[c:, c] for c in range(6)])
        assert np.allclose(H[0], packed, rtol=1e-10, atol=0), np.abs(H[0] - packed).max()


def test_icgn_zncc_is_the_correlation_at_the_returned_warp(orc):
    """The ZNCC returned per node is the zero-normalised correlation of the
    reference subset with the deformed image interpolated at the returned
    warp, x' = x0 + (1 + F1) dx + F3 dy + u, y' = y0 + F2 dx + (1 + F4) dy + v
    (F = [du/dx dv/dx du/dy dv/dy], funICGNQuadtree.m:149-152), i.e. at the
    last evaluation, which the convergence test ends without an update.
    Computed here with numpy and ba_interp2_spline (whose first coordinate
    runs along dim 2); a reordered F or an update after the test changes it."""
    f, g, _, _ = small_pair(shift=(0.4, -0.3), strain=0.002)
    im = orc.Images(f, g)
    for (x0, y0, ws) in ((70, 64, 20), (90, 80, 30)):
        U, F, H, conv, hole, z = orc.local_icgn(im, np.array([[float(x0), float(y0)]]),
                                                np.zeros(2), np.array([ws], np.int32), 1e-3)
        assert 1 < conv[0] <= 100
        h = ws // 2
        ii = np.arange(x0 - h, x0 + h + 1)
        jj = np.arange(y0 - h, y0 + h + 1)
        DX, DY = np.meshgrid((ii - x0).astype(float), (jj - y0).astype(float), indexing="ij")
        fs = f[np.ix_(ii - 1, jj - 1)]
        X = x0 + DX + F[0] * DX + F[2] * DY + U[0]
        Y = y0 + DY + F[1] * DX + F[3] * DY + U[1]
        gs = orc.interp_spline(g, Y, X)
        fc, gc = fs - fs.mean(), gs - gs.mean()
        zz = (fc * gc).sum() / np.sqrt((fc ** 2).sum() * (gc ** 2).sum())
        assert abs(zz - z[0]) <= 1e-12, (zz, z[0])


def test_icgn_affine_intensity_invariance(orc):
    f, g, _, _ = small_pair(shift=(1.3, -0.7))
    coords = np.asfortranarray(np.array([[60.0, 60.0], [90.0, 80.0]]))
    U0 = np.tile([1.0, -1.0], 2)
    a = orc.local_icgn(orc.Images(f, g), coords, U0, 20, 1e-3)
    b = orc.local_icgn(orc.Images(f, np.asfortranarray(2.5 * g + 0.75)), coords, U0, 20, 1e-3)
    assert np.array_equal(a[3], b[3])
    assert np.allclose(a[0], b[0], atol=1e-9)


def test_icgn_status_codes(orc):
    f, g, _, _ = small_pair(shift=(1.0, 0.5))
    M, N = f.shape
    mask = np.ones((M, N))
    mask[40:80, 40:80] = 0  # a masked block
    im = orc.Images(f, g, mask)
    coords = np.asfortranarray(np.array([
        [50.0, 50.0],    # inside the masked block: >= 40% masked -> 102
        [10.0, 60.0],    # x0 - 10 < 4 -> Df index error -> -1
        [120.0, 100.0],  # fine
        [100.0, 60.0],   # U0 far out of the image -> 101
    ]))
    U0 = np.array([1.0, 0.5, 1.0, 0.5, 1.0, 0.5, 500.0, 0.0])
    U, F, H, conv, hole, z = orc.local_icgn(im, coords, U0, 20, 1e-3)
    assert conv[0] == 102 and np.array_equal(U[0:2], U0[0:2]) and np.all(F[0:4] == 0)
    assert np.all(H[0] == 0) and hole[0] == 1
    assert conv[1] == -1 and np.all(np.isnan(U[2:4]))
    assert 1 <= conv[2] <= 20 and hole[2] == 0
    assert conv[3] == 101 and np.array_equal(U[6:8], U0[6:8])
    # NaN initial guess -> -1 (logical(NaN) throws)
    U, F, H, conv, hole, z = orc.local_icgn(im, coords[2:3], np.array([np.nan, 0.0]), 20, 1e-3)
    assert conv[0] == -1


def test_icgn_partly_masked_subset_runs_mask_branch(orc):
    f, g, _, _ = small_pair(shift=(0.6, -0.4))
    M, N = f.shape
    mask = np.ones((M, N))
    mask[:, 78:] = 0  # node at y0 = 70 with winsize 20: columns 60..80 partly masked
    im = orc.Images(f, g, mask)
    U, F, H, conv, hole, z = orc.local_icgn(im, np.asfortranarray([[80.0, 70.0]]),
                                            np.array([1.0, 0.0]), 20, 1e-3)
    # window grows, masked pixels are excluded; the subset still converges
    assert conv[0] == -1 or (1 <= conv[0] <= 100)


def test_bwselect_bruteforce(orc):
    from scipy.ndimage import label
    rng = np.random.default_rng(6)
    for trial in range(30):
        BW = rng.random((11, 9)) < 0.55
        r, c = rng.integers(1, 12), rng.integers(1, 10)
        for conn, st in ((4, None), (8, np.ones((3, 3)))):
            lab, _ = label(BW, structure=st)
            want = (lab == lab[r - 1, c - 1]) & BW[r - 1, c - 1]
            got = orc.bwselect(BW, r, c, conn)
            assert np.array_equal(got, want)


def test_subpb1_limits(orc):
    f, g, _, _ = small_pair(shift=(1.2, -0.3))
    coords = np.asfortranarray(np.array([[70.0, 70.0], [90.0, 80.0]]))
    UOld = np.tile([1.0, 0.0], 2)
    FOld = np.zeros(8)
    vdual = np.tile([0.05, -0.02], 2)
    ws = np.asfortranarray(np.full((2, 2), 20, dtype=np.int32))
    im = orc.Images(f, g)
    # tiny mu: the data term wins, U goes to the true translation
    U0, conv = orc.subpb1(im, coords, UOld, FOld, np.zeros(8), vdual, ws, 1e-12, 1e-6)
    assert np.allclose(U0, np.tile([1.2, -0.3], 2), atol=0.02)
    # Fixed point with a penalty: the step solves (Hd + mu I) dP = -tempb,
    # tempb = 2 b / bottomf^2 + mu (P - UOld - vdual), and the inverse-
    # compositional update moves P by -dP (funICGN_Subpb1quadtree.m:381-440).
    # Near the data optimum P0, 2 b / bottomf^2 ~ -Hd (P - P0) with
    # Hd = 2 H / bottomf^2, so the iteration stops where
    # (Hd - mu I) P = Hd P0 - mu T: the penalty, with the reference's sign,
    # pushes P away from T = UOld + vdual.  Hd from the subset's own
    # gradients and moments, computed independently here.
    gx, gy = orc.img_gradient(f)
    for k in range(2):
        x0, y0 = int(coords[k, 0]), int(coords[k, 1])
        sl = (slice(x0 - 11, x0 + 10), slice(y0 - 11, y0 + 10))
        dsl = (slice(x0 - 14, x0 + 7), slice(y0 - 14, y0 + 7))
        fx, fy = gx[dsl], gy[dsl]
        fs = f[sl]
        bf2 = ((fs - fs.mean()) ** 2).sum()
        Hd = 2.0 / bf2 * np.array([[(fx * fx).sum(), (fx * fy).sum()],
                                   [(fx * fy).sum(), (fy * fy).sum()]])
        mu = 0.05 * np.trace(Hd) / 2
        T = UOld[2 * k:2 * k + 2] + vdual[2 * k:2 * k + 2]
        Uk, ck = orc.subpb1(im, coords[k:k + 1], UOld[2 * k:2 * k + 2], FOld[:4], np.zeros(4),
                            vdual[2 * k:2 * k + 2], ws[k:k + 1], mu, 1e-8)
        P0 = U0[2 * k:2 * k + 2]
        want = np.linalg.solve(Hd - mu * np.eye(2), Hd @ P0 - mu * T)
        assert np.allclose(Uk, want, atol=2e-3)
        # a dropped penalty or a flipped sign would land elsewhere by
        assert np.linalg.norm(want - P0) > 5e-3


def test_c1_reference_configuration(orc):
    """C1: 900 subsets; the first grid row/column throws (DfDx index < 1)."""
    f, g, uf, vf = c1_pair()
    coords, nX, nY = c1_grid()
    assert (nX, nY) == (30, 30)
    s = orc.integer_search(f, g, [1, 512], [1, 512], 30, 16, [10, 10, 10, 10], nthreads=4)
    assert np.nanmax(np.abs(s["u"] - uf(s["x"], s["y"]))) < 0.6
    U0 = np.zeros(2 * 900)
    U0[0::2] = s["u"].T.reshape(-1)
    U0[1::2] = s["v"].T.reshape(-1)
    U, F, H, conv, hole, z = orc.local_icgn(orc.Images(f, g), coords, U0, 30, 1e-3, nthreads=4)
    assert np.sum(conv == -1) == 59
    ok = (conv > 0) & (conv <= 100)
    assert ok.sum() == 841
    eu = U[0::2][ok] - uf(coords[ok, 0], coords[ok, 1])
    assert np.sqrt(np.mean(eu ** 2)) < 0.03  # first-order bias of the sine term


def test_bad_point_rules(pkg):
    conv = np.array([3, 4, 5, 4, 3, 101, -1, 102, 20, 4])
    bad, num = pkg.icgn_bad_points(conv)
    good = np.array([3, 4, 5, 4, 3, 20, 4], dtype=float)
    thr = max(good.mean() + good.std(ddof=1), 6)
    assert num == 3 - 1
    assert set(bad) == {5, 6, 7} | set(np.flatnonzero(conv > thr))
    bad1, num1 = pkg.subpb1_bad_points(conv)
    thr1 = max(good.mean() + 0.25 * good.std(ddof=1), 10)
    assert set(bad1) == {5, 6, 7} | set(np.flatnonzero(conv > thr1))


def _b_at(orc, im, f, g, x0, y0, ws, P):
    """b = bottomf * sum SD^T r at the warp P = [ux vx uy vy u v] of an
    unmasked subset (funICGNQuadtree.m:149-152, :365-374), from the image
    planes with numpy and ba_interp2_spline."""
    h = ws // 2
    ii = np.arange(x0 - h, x0 + h + 1)
    jj = np.arange(y0 - h, y0 + h + 1)
    DX, DY = np.meshgrid((ii - x0).astype(float), (jj - y0).astype(float), indexing="ij")
    fs = f[np.ix_(ii - 1, jj - 1)]
    fx = im.DfDx[np.ix_(ii - 4, jj - 4)]
    fy = im.DfDy[np.ix_(ii - 4, jj - 4)]
    X = x0 + (1 + P[0]) * DX + P[2] * DY + P[4]
    Y = y0 + P[1] * DX + (1 + P[3]) * DY + P[5]
    gs = orc.interp_spline(g, Y, X)
    bf = np.sqrt(((fs - fs.mean()) ** 2).sum())
    bg = np.sqrt(((gs - gs.mean()) ** 2).sum())
    r = (fs - fs.mean()) / bf - (gs - gs.mean()) / bg
    sd = np.stack([DX * fx, DX * fy, DY * fx, DY * fy, fx, fy]).reshape(6, -1)
    return bf * (sd @ r.reshape(-1))


def test_icgn_stopping_rule_is_relative_to_the_first_b(orc):
    """The returned convIt is the first evaluation k with ||b_k|| / ||b_1||
    < tol or ||b_k|| < tol, and the returned warp is the one that evaluation
    saw, not updated (funICGNQuadtree.m:376-389).  The warps P_j after j
    updates come from evaluation-capped runs (a cap of j - 1 leaves P_j);
    ||b|| at each is computed here with numpy from the image planes.  A
    rule relative to the previous evaluation's ||b||, a step-size rule
    ||dP|| < tol, or an update after the test would pick other counts on
    some of these nodes."""
    f, g, _, _ = small_pair(shift=(0.45, -0.3), strain=0.003)
    im = orc.Images(f, g)
    tol = 1e-3
    nodes = [(70, 64, 20), (90, 80, 30), (60, 90, 16), (100, 70, 24), (80, 100, 20),
             (110, 60, 30)]
    prev_rule_differs = dp_rule_differs = 0
    for (x0, y0, ws) in nodes:
        c = np.array([[float(x0), float(y0)]])
        wsa = np.array([ws], np.int32)
        U0 = np.array([0.0, 0.0])
        U, F, H, conv, hole, z = orc.local_icgn(im, c, U0, wsa, tol)
        s = int(conv[0])
        assert 2 < s <= 100, s
        Ps = [np.array([0, 0, 0, 0, 0.0, 0.0])]
        for j in range(1, s):
            Uj, Fj, _, cj, _, _ = orc.local_icgn(im, c, U0, wsa, tol, max_iter=j - 1)
            assert cj[0] == 101, (j, cj[0])  # not converged within j evaluations
            Ps.append(np.concatenate([Fj, Uj]))
        nb = np.array([np.linalg.norm(_b_at(orc, im, f, g, x0, y0, ws, P)) for P in Ps])
        rel = nb / nb[0]
        stop = [k for k in range(len(nb)) if rel[k] < tol or nb[k] < tol]
        assert stop and stop[0] == s - 1, (rel, s)  # evaluation s is index s - 1
        # the returned warp is the last evaluated one (no update after the test)
        assert np.allclose(np.concatenate([F, U]), Ps[s - 1], atol=1e-12)
        # what the alternatives would have returned
        prev = [k for k in range(1, len(nb)) if nb[k] / nb[k - 1] < tol]
        prev_rule_differs += (not prev) or prev[0] != s - 1
        dps = [k for k in range(1, len(Ps)) if np.linalg.norm(Ps[k] - Ps[k - 1]) < tol]
        dp_rule_differs += (not dps) or dps[0] + 1 != s
    assert prev_rule_differs > 0 and dp_rule_differs > 0


def test_icgn_mask_branch_window_growth(orc):
    """A partly masked subset: the window grows to 2 ceil(sqrt(S^2 / (S^2 -
    c0)) ws / 2) with c0 masked pixels in the nominal S x S window, and f,
    DfDx, DfDy keep only the 4-connected mask component of the centre
    (funICGNQuadtree.m:66-88).  With ImgDef = ImgRef and a zero start the
    first evaluation converges, so H (HtempPar) is returned; it must equal
    the Gram matrix of the steepest-descent images over that component of
    the grown window, built here with numpy and scipy's labelling.  A zero
    stripe 3 px wide cuts the window, so the pixels beyond it are masked
    out by the component rule although the mask keeps them."""
    from scipy.ndimage import label
    f, _, _, _ = small_pair(seed=21)
    M, N = f.shape
    x0, y0, ws = 80, 70, 20
    S = ws + 1
    mask = np.ones((M, N))
    mask[:, y0 + 5 - 1:y0 + 7] = 0  # MATLAB columns y0+5 .. y0+7
    im = orc.Images(f, f, mask)
    U, F, H, conv, hole, z = orc.local_icgn(im, np.array([[float(x0), float(y0)]]),
                                            np.zeros(2), np.array([ws], np.int32), 1e-3)
    assert conv[0] == 1, conv
    c0 = 3 * S
    wn = int(2 * np.ceil(np.sqrt(S * S / (S * S - c0)) * ws / 2))
    assert wn == 22  # the grown window differs from the nominal one
    h = wn // 2
    ii = np.arange(x0 - h, x0 + h + 1)
    jj = np.arange(y0 - h, y0 + h + 1)
    m = (f * mask)[np.ix_(ii - 1, jj - 1)] != 0
    lab, _ = label(m)  # 4-connectivity
    seed = (wn + 1) // 2  # bwselect(..., floor((ws+1)/2), floor((ws+1)/2), 4)
    comp = lab == lab[seed - 1, seed - 1]
    assert comp.sum() < m.sum()  # the stripe cuts some of the mask off
    DX, DY = np.meshgrid((ii - x0).astype(float), (jj - y0).astype(float), indexing="ij")

    def packed_gram(keep, lo, hi):
        """HGlobal packing of sum sd sd^T over rows/cols lo:hi of the grown
        window, gradients kept where `keep` holds."""
        gx = (im.DfDx[np.ix_(ii - 4, jj - 4)] * keep)[lo:hi, lo:hi]
        gy = (im.DfDy[np.ix_(ii - 4, jj - 4)] * keep)[lo:hi, lo:hi]
        dx, dy = DX[lo:hi, lo:hi], DY[lo:hi, lo:hi]
        a = np.stack([dx * gx, dx * gy, dy * gx, dy * gy, gx, gy]).reshape(6, -1)
        G = a @ a.T
        return np.concatenate([G[c:, c] for c in range(6)])

    packed = packed_gram(comp, 0, 2 * h + 1)
    assert np.allclose(H[0], packed, rtol=1e-10, atol=0), np.abs(H[0] - packed).max()
    # the nominal 21 x 21 window, or the whole mask instead of the component,
    # give other Hessians
    for other in (packed_gram(comp, 1, 2 * h), packed_gram(m, 0, 2 * h + 1)):
        assert np.abs(other - packed).max() > 1e-3 * np.abs(packed).max()


def _ncc_definition(T, A):
    """normxcorr2 by its definition on the zero-padded image: Pearson-type
    coefficient of the mean-removed template with every template-sized
    window of A padded by s - 1 zeros on each side (n = template size in
    the window's moments)."""
    s1, s2 = T.shape
    W1, W2 = A.shape
    P = np.zeros((W1 + 2 * (s1 - 1), W2 + 2 * (s2 - 1)))
    P[s1 - 1:s1 - 1 + W1, s2 - 1:s2 - 1 + W2] = A
    Tc = T - T.mean()
    n = T.size
    out = np.zeros((W1 + s1 - 1, W2 + s2 - 1))
    for p in range(out.shape[0]):
        for q in range(out.shape[1]):
            w = P[p:p + s1, q:q + s2]
            den = np.sqrt((Tc ** 2).sum() * max((w ** 2).sum() - w.sum() ** 2 / n, 0.0))
            out[p, q] = (Tc * w).sum() / den if den > 1e-12 else 0.0
    return out


def test_normxcorr2_full_map_and_its_crop(orc):
    rng = np.random.default_rng(41)
    T = rng.normal(size=(7, 7))
    A = rng.normal(size=(15, 12))
    full = orc.normxcorr2_full(T, A)
    want = _ncc_definition(T, A)
    assert full.shape == (21, 18)
    assert np.abs(full - want).max() < 1e-12
    # the crop funIntegerSearch.m:166 takes: (ws : end - ws + 1), ws = s - 1
    ws = 6
    assert np.array_equal(orc.normxcorr2_crop(T, A), full[ws - 1:full.shape[0] - ws + 1,
                                                          ws - 1:full.shape[1] - ws + 1])


def test_qfactor_closed_forms(orc):
    """PCE = peak^2 / mean((cc - min)^2); PPE = 1 / entropy of the 30-bin
    histogram of cc - min, with histcounts' nice bin width: for a peak of 1
    the bins are 0.034 wide (ll = 1/30, ul = 1/29: the slack 0.00115 puts the
    rounding at 1e-3), not 1/30."""
    e = orc.hist_edges(0.0, 1.0, 30)
    assert np.allclose(e[:-1], 0.034 * np.arange(30)) and e[-1] == pytest.approx(1.02)
    e = orc.hist_edges(0.0, 0.7346, 30)
    assert e[1] == pytest.approx(0.0245) and e[-1] == pytest.approx(0.735)
    # values at the 0.034-bin centres with known counts, plus 0 and 1
    counts = {0: 5, 29: 3}
    vals = [0.0] * 5 + [1.0] * 3
    for k in range(1, 29):
        vals += [(k + 0.5) * 0.034] * k
        counts[k] = k
    cc = np.array(vals) - 0.25  # the minimum is removed first
    pce, ppe = orc.qfactor(cc)
    c = np.array(list(counts.values()), dtype=float)
    p = c / c.sum()
    assert ppe == pytest.approx(1.0 / np.sum(p * np.log(1 / p)), rel=1e-12)
    cm = cc - cc.min()
    assert pce == pytest.approx(1.0 / np.mean(cm ** 2), rel=1e-12)
    # with 1/30-wide bins some of those centres would share a bin: another
    # entropy
    k_alt = np.minimum(np.floor(np.array(vals) * 30), 29)
    pa = np.unique(k_alt, return_counts=True)[1] / len(vals)
    assert abs(1.0 / np.sum(pa * np.log(1 / pa)) - ppe) > 1e-3
    # a two-valued map: entropy of the two bins
    two = np.array([0.2] * 7 + [0.9] * 3)
    assert orc.qfactor(two)[1] == pytest.approx(1.0 / -(0.7 * np.log(0.7) + 0.3 * np.log(0.3)))


def test_search_qfactors_order_and_normalisation(orc):
    """cc.qfactors rows follow the reference's point order (x index fastest),
    each point's pair is compute_qFactor of its full normxcorr2 map (window
    clamps of funIntegerSearch.m:139-152), and each column is shifted and
    scaled to [0, 1] over the grid (:184-187)."""
    f, g, _, _ = small_pair(shift=(1.3, -0.7))
    M, N = f.shape
    ws, step, r = 20, 16, 6
    s = orc.integer_search(f, g, [1, M], [1, N], ws, step, [r] * 4, qfactors=True)
    nY, nX = s["u"].shape
    q = s["qfactors"]
    assert q.shape == (nX * nY, 2)
    raw = np.zeros((nX * nY, 2))
    for t in range(nX * nY):
        ci, cj = t % nX, t // nX
        ii, jj = 1 + ci * step, 1 + cj * step
        L = min(r, ii - 1)
        U = min(r, jj - 1)
        R = r if ii + ws + r <= M else max(M - ii - ws - 1, 0)
        D = r if jj + ws + r <= N else max(N - jj - ws - 1, 0)
        C = f[ii - 1:ii + ws, jj - 1:jj + ws]
        Dw = g[ii - L - 1:ii + ws + R, jj - U - 1:jj + ws + D]
        raw[t] = orc.qfactor(_ncc_definition(C, Dw))
    want = (raw - raw.min(axis=0)) / (raw - raw.min(axis=0)).max(axis=0)
    assert np.abs(q - want).max() < 1e-9
    assert np.allclose(q.min(axis=0), 0) and np.allclose(q.max(axis=0), 1)


def _warp_matrix(P):
    """x' = W(P) [dx dy 1]^T of funICGNQuadtree.m:149-152, P = [ux vx uy vy u v]."""
    return np.array([[1 + P[0], P[2], P[4]], [P[1], 1 + P[3], P[5]], [0.0, 0.0, 1.0]])


def test_icgn_step_size_rule(orc):
    """Stop rule 1 (the throughput rule, no counterpart in the reference):
    the returned count is the first evaluation whose applied update dP moves
    the subset by less than tol_dp px, sqrt(du^2 + dv^2 + (du/dx^2 +
    dv/dx^2) h^2 + (du/dy^2 + dv/dy^2) h^2) with h the half width, and the
    returned warp includes that update.  dP is recovered here from the warps
    of capped runs through the inverse-compositional relation W(P_j) =
    W(P_{j-1}) W(dP_j)^-1 (funICGNQuadtree.m:396-414)."""
    f, g, _, _ = small_pair(shift=(0.45, -0.3), strain=0.003)
    im = orc.Images(f, g)
    tol_dp = 1e-3
    for (x0, y0, ws) in ((70, 64, 20), (90, 80, 30), (100, 70, 24)):
        c = np.array([[float(x0), float(y0)]])
        wsa = np.array([ws], np.int32)
        U0 = np.array([0.0, 0.0])
        U, F, H, conv, hole, z = orc.local_icgn(im, c, U0, wsa, 1e-3, stop_rule=1, tol_dp=tol_dp)
        s = int(conv[0])
        assert 2 < s <= 100, s
        Ps = [np.zeros(6)]
        for j in range(1, s):
            Uj, Fj, _, cj, _, _ = orc.local_icgn(im, c, U0, wsa, 1e-3, max_iter=j - 1,
                                                 stop_rule=1, tol_dp=tol_dp)
            assert cj[0] == 101
            Ps.append(np.concatenate([Fj, Uj]))
        Ps.append(np.concatenate([F, U]))
        h = ws // 2
        moves = []
        for j in range(1, s + 1):
            Wd = np.linalg.solve(_warp_matrix(Ps[j]), _warp_matrix(Ps[j - 1]))
            d = np.array([Wd[0, 0] - 1, Wd[1, 0], Wd[0, 1], Wd[1, 1] - 1, Wd[0, 2], Wd[1, 2]])
            moves.append(np.sqrt(d[4] ** 2 + d[5] ** 2 + (d[0] ** 2 + d[1] ** 2) * h * h +
                                 (d[2] ** 2 + d[3] ** 2) * h * h))
        assert moves[-1] < tol_dp and all(mv >= tol_dp for mv in moves[:-1]), moves
        # the reference rule on the same node stops elsewhere or at another warp
        Ur, Fr, _, cr, _, _ = orc.local_icgn(im, c, U0, wsa, 1e-3)
        assert cr[0] != s or np.abs(Ur - U).max() > 1e-9


def _update_measures(orc, im, c, U0, wsa, rule, tol_dp, s, F, U):
    """The measures m_1..m_s of the updates of a rule-`rule` run that
    stopped after s evaluations with warp [F U]: the warps after j updates
    come from capped runs, and dP_j from W(P_j) = W(P_{j-1}) W(dP_j)^-1."""
    Ps = [np.zeros(6)]
    for j in range(1, s):
        Uj, Fj, _, cj, _, _ = orc.local_icgn(im, c, U0, wsa, 1e-3, max_iter=j - 1,
                                             stop_rule=rule, tol_dp=tol_dp)
        assert cj[0] == 101
        Ps.append(np.concatenate([Fj, Uj]))
    Ps.append(np.concatenate([F, U]))
    h = int(wsa[0]) // 2
    moves = []
    for j in range(1, s + 1):
        Wd = np.linalg.solve(_warp_matrix(Ps[j]), _warp_matrix(Ps[j - 1]))
        d = np.array([Wd[0, 0] - 1, Wd[1, 0], Wd[0, 1], Wd[1, 1] - 1, Wd[0, 2], Wd[1, 2]])
        moves.append(np.sqrt(d[4] ** 2 + d[5] ** 2 + (d[0] ** 2 + d[1] ** 2) * h * h +
                             (d[2] ** 2 + d[3] ** 2) * h * h))
    return moves


def test_icgn_a_posteriori_rule(orc):
    """Stop rule 2 (no counterpart in the reference): converged at the first
    evaluation k whose update measure m_k is below tol_dp, or whose last two
    measures contract (m_k <= m_{k-1} / 2) with a predicted distance to the
    fixed point m_k q / (1 - q) below tol_dp, q = m_k / m_{k-1}.  The
    measures are recovered from capped runs as for rule 1; the rule never
    takes more evaluations than rule 1, takes fewer on some subset, and its
    warp stays within tol_dp (centre) of a tightly converged one."""
    f, g, _, _ = small_pair(shift=(0.45, -0.3), strain=0.003)
    im = orc.Images(f, g)
    tol_dp = 1e-3
    fewer = 0
    for (x0, y0, ws) in ((70, 64, 20), (90, 80, 30), (100, 70, 24), (80, 90, 28)):
        c = np.array([[float(x0), float(y0)]])
        wsa = np.array([ws], np.int32)
        U0 = np.array([0.0, 0.0])
        U, F, H, conv, hole, z = orc.local_icgn(im, c, U0, wsa, 1e-3, stop_rule=2, tol_dp=tol_dp)
        s = int(conv[0])
        assert 1 < s <= 100, s
        m = _update_measures(orc, im, c, U0, wsa, 2, tol_dp, s, F, U)

        def done(k):  # 0-based evaluation index
            if m[k] < tol_dp:
                return True
            return k >= 1 and 2 * m[k] <= m[k - 1] and m[k] ** 2 < tol_dp * (m[k - 1] - m[k])
        assert done(s - 1) and not any(done(k) for k in range(s - 1)), m
        _, _, _, c1, _, _ = orc.local_icgn(im, c, U0, wsa, 1e-3, stop_rule=1, tol_dp=tol_dp)
        assert s <= c1[0]
        fewer += int(s < c1[0])
        Ut, _, _, ct, _, _ = orc.local_icgn(im, c, U0, wsa, 1e-3, stop_rule=1, tol_dp=1e-9)
        assert ct[0] < 101 and np.abs(Ut - U).max() < tol_dp
    assert fewer >= 1


def test_normalize_roi_moments_and_the_host_version(orc, pkg):
    """funNormalizeImg.m:22-40: (Img - mean) / std over the ROI (clamped to
    the image), applied to the whole image.  Pinned by the ROI's moments
    (mean 0, sample std 1), by invariance to an affine change of intensity,
    and by an ROI that only the clamp keeps inside the image; the host
    funNormalizeImg must agree with it."""
    rng = np.random.default_rng(11)
    img = rng.uniform(0, 4095, size=(37, 29))
    for gx, gy in (([5, 30], [3, 20]), ([-4, 60], [0, 29])):
        out = orc.normalize(img, gx, gy)
        a0, a1 = max(gx[0], 1) - 1, min(gx[1], 37)
        b0, b1 = max(gy[0], 1) - 1, min(gy[1], 29)
        roi = out[a0:a1, b0:b1]
        assert abs(roi.mean()) < 1e-12 and abs(roi.std(ddof=1) - 1.0) < 1e-12
        assert np.allclose(orc.normalize(3.5 * img - 7.0, gx, gy), out, atol=1e-12)
        host, rng_out = pkg.funNormalizeImg([img, img], {"gridx": gx, "gridy": gy})
        assert np.allclose(host[0], out, atol=1e-12) and np.allclose(host[1], out, atol=1e-12)
        assert rng_out["gridx"] == [a0 + 1, a1] and rng_out["gridy"] == [b0 + 1, b1]


def test_bspline_prefilter_and_spline_against_scipy(orc):
    """The optional interpolant (no counterpart on the reference's path):
    coefficients against scipy.ndimage.spline_filter (order 3, mirror), the
    spline against scipy.ndimage.map_coordinates inside the image (where the
    MEX's cell never reaches the clamp), samples reproduced at nodes and
    linear ramps reproduced everywhere inside; the reference's local kernel
    misses a ramp by t(2t-1)(t-1)/14."""
    from scipy import ndimage
    rng = np.random.default_rng(21)
    for shape in ((37, 29), (5, 7), (2, 9), (64, 3)):
        A = rng.normal(size=shape)
        c = orc.bspline_prefilter(A)
        assert np.abs(c - ndimage.spline_filter(A, order=3, mode="mirror")).max() < 1e-12
        # the defining relation, (c[k-1] + 4 c[k] + c[k+1]) / 6 = s[k] along
        # both axes, in the interior; the two axes commute
        if min(shape) >= 3:
            r = (c[:-2] + 4 * c[1:-1] + c[2:]) / 6
            r = (r[:, :-2] + 4 * r[:, 1:-1] + r[:, 2:]) / 6
            assert np.abs(r - A[1:-1, 1:-1]).max() < 1e-12
            assert np.allclose(orc.bspline_prefilter(A.T).T, c, atol=1e-12)
    A = rng.normal(size=(37, 29))
    c = orc.bspline_prefilter(A)
    X = rng.uniform(2, 28 - 1e-9, 500)
    Y = rng.uniform(2, 36 - 1e-9, 500)
    z = orc.interp_bspline(c, X, Y)
    assert np.abs(z - ndimage.map_coordinates(A, [Y - 1, X - 1], order=3, mode="mirror")).max() \
        < 1e-12
    I, J = np.meshgrid(np.arange(2, 37), np.arange(2, 29), indexing="ij")
    assert np.abs(orc.interp_bspline(c, J.ravel(), I.ravel()) - A[I - 1, J - 1].ravel()).max() \
        < 1e-12
    # a ramp: its mirror extension kinks at the border, whose effect on the
    # coefficients decays as |sqrt(3) - 2|^d = 0.268^d with the distance d
    ramp = np.asfortranarray(0.37 * np.arange(1, 65)[:, None] - 0.11 * np.arange(1, 61)[None, :])
    cr = orc.bspline_prefilter(ramp)
    X = rng.uniform(16, 45, 300)
    Y = rng.uniform(16, 49, 300)
    assert np.abs(orc.interp_bspline(cr, X, Y) - (0.37 * Y - 0.11 * X)).max() < 1e-8
    t = Y - np.floor(Y)
    local = orc.interp_spline(ramp, X, Y) - (0.37 * Y - 0.11 * X)
    bias = 0.37 * t * (2 * t - 1) * (t - 1) / 14 - 0.11 * (X - np.floor(X)) * \
        (2 * (X - np.floor(X)) - 1) * ((X - np.floor(X)) - 1) / 14
    assert np.abs(local - bias).max() < 1e-6 and np.abs(local).max() > 1e-3


def test_icgn_bspline_mode_removes_the_kernel_bias(orc, synth):
    """IC-GN on a noise-free, analytically rendered sub-pixel translation:
    with the reference's kernel the displacement carries the kernel's bias
    (several 1e-3 px); with the interpolating spline it is an order of
    magnitude smaller.  The same oracle IC-GN runs both; only the images'
    interpolant differs."""
    M = N = 160
    u0, v0 = 0.3, -0.2
    f8, g8, _ = synth.make_pair(M, N, lambda X, Y: u0 + 0 * X, lambda X, Y: v0 + 0 * X,
                                density=0.02, seed=4, noise=0.0, bits=12)
    f = np.asfortranarray((f8 - f8.mean()) / f8.std())
    g = np.asfortranarray((g8 - g8.mean()) / g8.std())
    xs = np.arange(40.0, 121.0, 20.0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.asfortranarray(np.stack([X.ravel(), Y.ravel()], 1))
    U0 = np.zeros(2 * len(coords))
    err = {}
    for mode in ("local", "bspline"):
        U, F, H, conv, *_ = orc.local_icgn(orc.Images(f, g, interp=mode), coords, U0, 40, 1e-6)
        assert np.all((conv >= 1) & (conv <= 100))
        err[mode] = np.abs(np.stack([U[0::2] - u0, U[1::2] - v0])).max()
    assert err["local"] > 1e-3 and err["bspline"] < err["local"] / 8, err


def test_bspline_zero_rule(orc):
    """Pixels covered by a 4 x 4 window of all-zero samples get a zero
    coefficient, so the spline is exactly zero on the cells whose 16 samples
    are zero (as the reference's kernel is there) and the other coefficients
    are the plain spline's.  An isolated zero sample changes nothing."""
    from numpy.lib.stride_tricks import sliding_window_view
    from scipy import ndimage
    rng = np.random.default_rng(22)
    M, N = 48, 40
    A = rng.normal(size=(M, N))
    I, J = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
    A[(I - 20) ** 2 + (J - 18) ** 2 < 64] = 0.0  # a hole
    A[40, 5] = 0.0  # a lone zero
    A[0:5, 30:36] = 0.0  # a zero block on the border
    A = np.asfortranarray(A)
    c = orc.bspline_prefilter(A)
    plain = ndimage.spline_filter(A, order=3, mode="mirror")
    win0 = np.all(sliding_window_view(A == 0.0, (4, 4)), axis=(2, 3))  # window starts
    kill = np.zeros((M, N), bool)
    for a, b in zip(*np.nonzero(win0)):
        kill[a:a + 4, b:b + 4] = True
    assert kill.sum() > 100 and not kill[40, 5]
    assert np.all(c[kill] == 0.0)
    assert np.abs(c[~kill] - plain[~kill]).max() < 1e-12
    # the spline on cells (1-based X, Y in [2, M-1)): exactly zero iff the
    # cell's 16 samples are zero
    X = rng.uniform(2, N - 1 - 1e-9, 4000)
    Y = rng.uniform(2, M - 1 - 1e-9, 4000)
    z = orc.interp_bspline(c, X, Y)
    a = np.floor(Y).astype(int) - 2  # 0-based first tap row
    b = np.floor(X).astype(int) - 2
    cell0 = win0[a, b]
    assert cell0.sum() > 50
    assert np.all(z[cell0] == 0.0) and np.all(z[~cell0] != 0.0)
    assert np.array_equal(orc.interp_spline(A, X, Y) == 0.0, cell0)
