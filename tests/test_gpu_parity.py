"""GPU parity: the HIP library through its C ABI against the FP64 oracle on
the same seeded inputs.

Tolerances (floating point, the device computes in FP32 except where noted):
  * ba_interp2_spline double path: 1e-12 relative (FP64 on the device);
    single path: 2e-5 absolute on unit-variance data;
  * funImgGradient: 1e-12 (FP64 stencil on the device);
  * IC-GN / Subpb1 displacements: |dU| <= 1e-3 px on subsets that converge in
    both; status codes equal except that the converged iteration count may
    differ by one on <= 3% of the subsets (5% on the mixed-size and masked
    cases: the relative stopping rule ||b_k|| / ||b_1|| < tol sits on the
    FP32 round-off of b near convergence); deformation gradients
    |dF| <= 1e-5 where both stop after the same number of evaluations and
    <= 1e-3 (one converged Gauss-Newton step) where they stop one apart;
    the Hessian H within 1e-5 of the node's largest entry on every subset
    both solve (FP32 sums, about 1e-7 measured), and zero in both on the
    -1 and 102 statuses; the last evaluation's ZNCC within 1e-5 at equal
    iteration counts;
  * the drivers' NaN fill (host code): 1e-9 against the oracle's from the
    same pre-fill U;
  * integer search: integer ZNCC peaks bit-exact; sub-pixel u, v within
    1e-3 + 1e-9 (the reference rounds the fit to 1e-3, so an FP32 ZNCC can
    land on the other side of a rounding boundary).
"""
import natural_oracle
import numpy as np
import pytest

from _cases import c1_grid, c1_pair, hole_mask, small_pair

pytestmark = pytest.mark.gpu


def _cmp_icgn(got, want, conv_tol_frac=0.03, du_tol=1e-3, df_tol=1e-5, ws=None):
    U, F, H, conv, hole, z = got
    Uo, Fo, Ho, convo, holeo, zo = want
    if ws is not None:
        dUa = np.abs(U - Uo).reshape(-1, 2).max(axis=1)
        dFa = np.abs(F - Fo).reshape(-1, 4).max(axis=1)
        for k in np.argsort(-np.nan_to_num(dFa))[:6]:
            print(f"  worst: node {k} ws {ws[k]} conv {conv[k]}/{convo[k]} dU {dUa[k]:.2e} "
                  f"dF {dFa[k]:.2e} zncc {z[k]:.5f}/{zo[k]:.5f}")
    assert np.array_equal(hole, holeo)
    okb = (convo >= 1) & (convo <= 100) & (conv >= 1) & (conv <= 100)
    # the same failure statuses
    for code in (-1, 102, 103):
        assert np.array_equal(conv == code, convo == code), code
    # converged in one and not in the other only through the +-1 rule
    diff = conv != convo
    assert np.all(np.abs(conv[diff].astype(int) - convo[diff].astype(int)) <= 1) or \
        np.sum(diff) == 0 or np.all(np.isin(conv[diff], [100, 101]) | np.isin(convo[diff], [100, 101]))
    assert np.mean(diff) <= conv_tol_frac, f"{np.sum(diff)} iteration counts differ"
    du = np.abs(U - Uo).reshape(-1, 2)[okb]
    assert du.max(initial=0) <= du_tol, du.max()
    dF = np.abs(F - Fo).reshape(-1, 4)
    print(f"  subsets {len(conv)}: both converged {okb.sum()}, iteration counts differ "
          f"{diff.sum()}, max|dU| {du.max(initial=0):.2e} px, max|dF| same-count "
          f"{dF[okb & ~diff].max(initial=0):.2e}")
    # F: 1e-5 where both stop after the same number of evaluations; where
    # they stop one evaluation apart, F differs by the last Gauss-Newton
    # step, which the relative rule bounds only loosely (1e-3)
    same = okb & ~diff
    dF = np.abs(F - Fo).reshape(-1, 4)
    assert dF[same].max(initial=0) <= df_tol, dF[same].max()
    # H (the Gauss-Newton Hessian the reference returns per node, built once
    # before the iterations): FP32 sums over the subset's pixels against FP64,
    # relative to the node's largest entry, on every subset both solved
    both = (conv >= 1) & (conv <= 101) & (convo >= 1) & (convo <= 101)
    Hg, Hw = np.asarray(H).reshape(len(conv), 21), np.asarray(Ho).reshape(len(conv), 21)
    scale = np.maximum(np.abs(Hw[both]).max(axis=1, initial=0), 1e-300)[:, None]
    dH = np.abs(Hg[both] - Hw[both]) / scale
    print(f"  max relative |dH| {dH.max(initial=0):.2e} on {both.sum()} subsets")
    assert dH.max(initial=0) <= 1e-5, dH.max()
    # ZNCC of the last evaluation (1 - sum r^2 / 2) where both stop after the
    # same number of evaluations
    dz = np.abs(np.asarray(z) - np.asarray(zo))[same]
    print(f"  max |dZNCC| {dz.max(initial=0):.2e} on {same.sum()} subsets")
    assert dz.max(initial=0) <= 1e-5, dz.max()
    # the failure statuses carry a zero H in both
    for code in (-1, 102):
        sel = conv == code
        assert np.all(Hg[sel] == 0) and np.all(Hw[sel] == 0), code
    assert dF[okb & diff].max(initial=0) <= 1e-3, dF[okb & diff].max()
    # failed subsets report NaN exactly where the oracle does
    assert np.array_equal(np.isnan(U), np.isnan(Uo))
    return du.max(initial=0), dF[okb].max(initial=0)


def test_interp_f64(pkg, orc):
    rng = np.random.default_rng(10)
    F = np.asfortranarray(rng.normal(size=(50, 47)))
    X = rng.uniform(-4, 52, 4000)
    Y = rng.uniform(-4, 55, 4000)
    Z = pkg.ba_interp2_spline(F, X, Y)
    Zo = orc.interp_spline(F, X, Y)
    assert Z.shape == X.shape
    assert np.max(np.abs(Z - Zo)) <= 1e-12 * (1 + np.abs(Zo).max())
    # multi-channel F (P = 3) and matrix-shaped X, Y
    F3 = np.asfortranarray(rng.normal(size=(20, 22, 3)))
    X2 = rng.uniform(1, 22, (7, 9))
    Y2 = rng.uniform(1, 20, (7, 9))
    Z3 = pkg.ba_interp2_spline(F3, X2, Y2)
    assert Z3.shape == (7, 9, 3)
    for ch in range(3):
        Zc = orc.interp_spline(np.asfortranarray(F3[:, :, ch]), X2, Y2)
        assert np.max(np.abs(Z3[:, :, ch] - Zc)) <= 1e-12 * 10


def test_interp_f32_and_class_check(pkg, orc):
    rng = np.random.default_rng(11)
    F = np.asfortranarray(rng.normal(size=(40, 40)).astype(np.float32))
    X = rng.uniform(1, 40, 2000).astype(np.float32)
    Y = rng.uniform(1, 40, 2000).astype(np.float32)
    Z = pkg.ba_interp2_spline(F, X, Y)
    assert Z.dtype == np.float32
    Zo = orc.interp_spline(F, X, Y, single=True)
    assert np.max(np.abs(Z - Zo)) <= 2e-5
    with pytest.raises(TypeError):
        pkg.ba_interp2_spline(F, X.astype(np.float64), Y)


def test_img_gradient(pkg, orc):
    f, _, _, _ = small_pair()
    M, N = f.shape
    mask = hole_mask(M, N, (80, 70), 20)
    Df = pkg.funImgGradient(f, f, mask)
    gx, gy = orc.img_gradient(f, mask)
    assert Df["DfDx"].shape == (M - 6, N - 6) and Df["DfCropWidth"] == 3
    assert np.max(np.abs(Df["DfDx"] - gx)) <= 1e-12
    assert np.max(np.abs(Df["DfDy"] - gy)) <= 1e-12


@pytest.fixture(scope="module")
def c1_setup(orc):
    f, g, uf, vf = c1_pair()
    coords, nX, nY = c1_grid()
    s = orc.integer_search(f, g, [1, 512], [1, 512], 30, 16, [10, 10, 10, 10], nthreads=8)
    U0 = np.zeros(2 * coords.shape[0])
    U0[0::2] = s["u"].T.reshape(-1)
    U0[1::2] = s["v"].T.reshape(-1)
    return f, g, uf, vf, coords, U0, s


def test_local_icgn_c1(pkg, orc, c1_setup):
    f, g, uf, vf, coords, U0, _ = c1_setup
    im = pkg.Images(f, g)
    got = pkg.local_icgn_raw(im, coords, U0, 30, 1e-3)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, 30, 1e-3, nthreads=8)
    du, dF = _cmp_icgn(got, want)
    print(f"C1 LocalICGN: max|dU| {du:.2e} px, max|dF| {dF:.2e}")
    conv = got[3]
    assert np.sum(conv == -1) == 59 and np.sum((conv >= 1) & (conv <= 100)) == 841
    # Hessians (HtempPar) agree to FP32 accumulation accuracy
    ok = (want[3] >= 1) & (want[3] <= 100)
    rel = np.abs(got[2][ok] - want[2][ok]) / (np.abs(want[2][ok]).max(axis=1, keepdims=True))
    assert rel.max() < 1e-4
    assert np.nanmax(np.abs(got[5][ok] - want[5][ok])) < 1e-4  # ZNCC


def test_local_icgn_driver_outputs(pkg, c1_setup):
    f, g, uf, vf, coords, U0, _ = c1_setup
    Df = pkg.funImgGradient(f, g)
    U, F, H, t, conv, badnum, hole = pkg.LocalICGNQuadtree(
        U0, coords, Df, f, g, {"winsize": 30, "winstepsize": 16, "ClusterNo": 1}, "GaussNewton",
        1e-3, 1)
    assert U.shape == (1800,) and F.shape == (3600,) and H.shape == (900, 21)
    bad, num = pkg.icgn_bad_points(conv)
    assert badnum == num and len(bad) > 0
    good = np.setdiff1d(np.arange(900), bad)
    eu = U[2 * good] - uf(coords[good, 0], coords[good, 1])
    assert np.sqrt(np.mean(eu ** 2)) < 0.03
    # the NaN fill (LocalICGNQuadtree.m:200-325) against the oracle's, from
    # the same pre-fill U
    U1, F1, *_ = pkg.LocalICGNQuadtree(
        U0, coords, Df, f, g, {"winsize": 30, "winstepsize": 16, "ClusterNo": 1}, "GaussNewton",
        1e-3, 1, fill_nans=False)
    assert np.all(np.isnan(U1[2 * bad])) and np.array_equal(U1[2 * good], U[2 * good])
    assert np.all(np.isfinite(U))
    want = natural_oracle.nan_fill(coords, U1, F1, Df["ImgRefMask"])
    assert np.max(np.abs(U - want)) < 1e-9
    # filled values follow the field as well as their neighbours do
    ef = U[2 * bad] - uf(coords[bad, 0], coords[bad, 1])
    inner = (coords[bad, 0] > 16) & (coords[bad, 1] > 16)
    assert np.max(np.abs(ef[inner])) < 0.5


def test_local_icgn_status_codes(pkg, orc):
    f, g, _, _ = small_pair(shift=(1.0, 0.5))
    M, N = f.shape
    mask = np.ones((M, N))
    mask[40:80, 40:80] = 0
    coords = np.asfortranarray(np.array([
        [50.0, 50.0], [10.0, 60.0], [120.0, 100.0], [100.0, 60.0], [79.0, 100.0], [60.0, 110.0],
    ]))
    U0 = np.array([1.0, 0.5, 1.0, 0.5, 1.0, 0.5, 500.0, 0.0, 1.2, 0.4, np.nan, 0.0])
    Df = {"DfDx": orc.img_gradient(f, mask)[0], "DfDy": orc.img_gradient(f, mask)[1],
          "ImgRefMask": mask}
    im = pkg.Images(f, g, mask, Df)
    got = pkg.local_icgn_raw(im, coords, U0, 20, 1e-3)
    want = orc.local_icgn(orc.Images(f, g, mask), coords, U0, 20, 1e-3)
    assert list(got[3]) == list(want[3])
    _cmp_icgn(got, want, conv_tol_frac=0.0)


def test_local_icgn_mask_branch(pkg, orc):
    """Partly masked subsets: window growth, bwselect, g masking (generic
    kernel) against the oracle."""
    f, g, _, _ = small_pair(M=200, N=180, seed=7, shift=(0.8, -0.6))
    M, N = f.shape
    mask = hole_mask(M, N, (100, 90), 28)
    fm = np.asfortranarray(f * mask)
    gm = np.asfortranarray(g * hole_mask(M, N, (100.8, 89.4), 28))
    xs = np.arange(40, 161, 8, dtype=float)
    ys = np.arange(40, 141, 8, dtype=float)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    coords = np.asfortranarray(np.stack([X.reshape(-1), Y.reshape(-1)], 1))
    n = coords.shape[0]
    U0 = np.tile([1.0, -1.0], n)
    im = pkg.Images(fm, gm, mask)
    got = pkg.local_icgn_raw(im, coords, U0, 16, 1e-3)
    want = orc.local_icgn(orc.Images(fm, gm, mask), coords, U0, 16, 1e-3)
    # a good number of subsets take the mask branch
    assert np.sum((want[3] == 102)) > 5 and np.sum(want[3] == -1) >= 0
    _cmp_icgn(got, want, conv_tol_frac=0.05)


def test_local_icgn_mixed_sizes(pkg, orc):
    """Per-node winsize 8..64 and 80 (every kernel class incl. the generic
    one for > 65 px) against the oracle."""
    f, g, uf, vf = small_pair(M=256, N=240, seed=9, shift=(0.4, 0.3), strain=0.002)
    rng = np.random.default_rng(12)
    n = 120
    ws = rng.choice([8, 10, 16, 20, 30, 32, 48, 64, 80], size=n).astype(np.int32)
    cx = rng.integers(60, 197, n).astype(float)
    cy = rng.integers(60, 181, n).astype(float)
    coords = np.asfortranarray(np.stack([cx, cy], 1))
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(cx, cy))
    U0[1::2] = np.round(vf(cx, cy))
    im = pkg.Images(f, g)
    got = pkg.local_icgn_raw(im, coords, U0, ws, 1e-3)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3)
    _cmp_icgn(got, want, conv_tol_frac=0.05, ws=ws)


@pytest.mark.parametrize("ws", [10, 16, 32, 64])
def test_local_icgn_full_class(pkg, orc, ws):
    """Every subset of a class at the class's full side (11, 17, 33, 65 px):
    the compile-time-size wave kernels against the oracle."""
    f, g, uf, vf = small_pair(M=256, N=240, seed=10, shift=(0.6, -0.2), strain=0.002)
    rng = np.random.default_rng(ws)
    n = 80
    cx = rng.integers(40, 217, n).astype(float)
    cy = rng.integers(40, 201, n).astype(float)
    coords = np.asfortranarray(np.stack([cx, cy], 1))
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(cx, cy))
    U0[1::2] = np.round(vf(cx, cy))
    im = pkg.Images(f, g)
    got = pkg.local_icgn_raw(im, coords, U0, ws, 1e-3)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, np.full(n, ws, np.int32), 1e-3)
    _cmp_icgn(got, want, conv_tol_frac=0.05)


def test_subpb1(pkg, orc, c1_setup):
    f, g, uf, vf, coords, U0, _ = c1_setup
    want0 = orc.local_icgn(orc.Images(f, g), coords, U0, 30, 1e-3, nthreads=8)
    UOld = np.nan_to_num(want0[0])
    FOld = np.nan_to_num(want0[1])
    n = coords.shape[0]
    rng = np.random.default_rng(13)
    vdual = rng.normal(scale=0.02, size=2 * n)
    udual = np.zeros(4 * n)
    wl = np.full((n, 2), 30, dtype=np.int32)
    U, conv = pkg.subpb1_raw(pkg.Images(f, g), coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2)
    Uo, convo = orc.subpb1(orc.Images(f, g), coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2,
                           nthreads=8)
    ok = (convo >= 1) & (convo <= 100) & (conv >= 1) & (conv <= 100)
    assert np.array_equal(convo == 0, conv == 0)
    assert np.mean(conv != convo) <= 0.03
    assert np.abs(U - Uo).reshape(-1, 2)[ok].max() <= 1e-3
    # the driver: bad-point rules and NaN fill (Subpb1Quadtree.m:167-254)
    mask = np.ones(f.shape)
    mask[200:230, :] = 0  # two regions
    Df = pkg.funImgGradient(f, g)
    DICpara = {"winsize": 30, "winsize_List": wl, "ImgRefMask": mask}
    args = (UOld, FOld, udual, vdual, coords, Df, f, g, 1e-3, 0.0, None, 2, DICpara, "GaussNewton",
            1e-2)
    Ud, _, _, convd, badnum = pkg.Subpb1Quadtree(*args)
    Upre, _, _, convp, _ = pkg.Subpb1Quadtree(*args, fill_nans=False)
    assert np.array_equal(convd, convp)
    bad, num = pkg.subpb1_bad_points(convd)
    assert badnum == num
    # NaN: the bad nodes, and the nodes whose solve threw (convIt 0, U NaN)
    nan_nodes = np.flatnonzero(np.isnan(Upre[0::2]))
    assert np.all(np.isin(bad, nan_nodes)) and nan_nodes.size > 0
    want = natural_oracle.nan_fill(coords, Upre, None, mask)
    assert np.all(np.isfinite(Ud)) and np.max(np.abs(Ud - want)) < 1e-9


def test_integer_search_c1(pkg, orc, c1_setup):
    f, g, uf, vf, coords, U0, s = c1_setup
    x, y, u, v, cc, pr, pc = pkg.funIntegerSearch(f, g, [10, 10], [1, 512], [1, 512], 30, 16,
                                                  return_peaks=True)
    assert u.shape == (30, 30)
    assert np.array_equal(x, s["x"]) and np.array_equal(y, s["y"])
    assert np.array_equal(pr, s["ipeak_row"]) and np.array_equal(pc, s["ipeak_col"])
    assert np.max(np.abs(u - s["u"])) <= 1e-3 + 1e-9
    assert np.max(np.abs(v - s["v"])) <= 1e-3 + 1e-9
    assert np.mean(u != s["u"]) < 0.02
    assert np.max(np.abs(cc["max"] - s["ccmax"])) < 1e-4


def test_integer_search_edges_and_radius(pkg, orc):
    f, g, _, _ = small_pair(M=200, N=190, seed=3, shift=(3.4, -2.2))
    for region, ws, step in (([8, 8], 20, 12), ([5, 9], 16, 10), ([12, 12], 32, 16)):
        x, y, u, v, cc, pr, pc = pkg.funIntegerSearch(f, g, region, [1, 200], [1, 190], ws, step,
                                                      return_peaks=True)
        r = np.concatenate([region, region])
        s = orc.integer_search(f, g, [1, 200], [1, 190], ws, step, r)
        assert np.array_equal(pr, s["ipeak_row"]) and np.array_equal(pc, s["ipeak_col"])
        assert np.max(np.abs(u - s["u"])) <= 1e-3 + 1e-9
        assert np.max(np.abs(v - s["v"])) <= 1e-3 + 1e-9


def test_integer_search_qfactors(pkg, orc):
    """cc.qfactors (PCE, PPE of each point's full normxcorr2 map, normalised
    over the grid) against the oracle: PCE within 1e-4 and PPE within 2e-3
    after normalisation (the device's FP32 coefficients can put a value
    within 1e-6 of a histogram edge in the other bin).  A point with a
    constant template (normxcorr2 throws) is NaN and turns every other
    point's q-factors to 0, in both."""
    f, g, uf, vf = c1_pair()
    cases = [(f, g, [1, 512], [1, 512], 30, 16, [10, 10]),
             (f, g, [1, 300], [1, 280], 20, 12, [5, 9])]
    fc = np.array(f)
    fc[100:140, 100:140] = 0.5  # one grid point's template is constant
    cases.append((np.asfortranarray(fc), g, [1, 300], [1, 280], 20, 12, [6, 6]))
    for k, (ff, gg, gx, gy, ws, st, r) in enumerate(cases):
        x, y, u, v, cc = pkg.funIntegerSearch(ff, gg, r, gx, gy, ws, st)
        s = orc.integer_search(ff, gg, gx, gy, ws, st, np.concatenate([r, r]), nthreads=8,
                               qfactors=True)
        q, qo = cc["qfactors"], s["qfactors"]
        assert q.shape == qo.shape == (u.size, 2)
        assert np.array_equal(np.isnan(q), np.isnan(qo))
        fin = ~np.isnan(qo[:, 0])
        d0 = np.abs(q[fin, 0] - qo[fin, 0]).max()
        d1 = np.abs(q[fin, 1] - qo[fin, 1]).max()
        print(f"case {k}: {u.size} points, NaN {np.isnan(qo[:, 0]).sum()}, "
              f"max |dPCE| {d0:.2e}, max |dPPE| {d1:.2e}")
        assert d0 <= 1e-4 and d1 <= 2e-3
        if k == 2:
            assert np.isnan(qo[:, 0]).sum() >= 1 and np.all(q[fin] == 0)


def test_integer_search_quadtree_driver(pkg, orc):
    """IntegerSearchQuadtree end to end (stereo radius, mask -> NaN, the
    median test and the springs fill of funRemoveOutliers) against the
    oracle's search followed by the dense restatement of the outlier pass."""
    import inpaint_oracle
    f, g, uf, vf = c1_pair()
    M, N = f.shape
    mask = hole_mask(M, N, (256, 256), 60)
    DICpara = {"gridxyROIRange": {"gridx": [1, M], "gridy": [1, N]}, "winsize": 30,
               "winstepsize": 16, "StereoSearchDistance": 10, "ImgRefMask": mask,
               "ImgSize": (M, N), "showImgOrNot": 0}
    P, x0, y0, u, v, cc = pkg.IntegerSearchQuadtree(f, g, None, DICpara, -1)
    s = orc.integer_search(f, g, [1, M], [1, N], 30, 16, [10, 10, 10, 10], nthreads=8)
    hole = mask[np.round(s["x"]).astype(int) - 1, np.round(s["y"]).astype(int) - 1] == 0
    uo = np.where(hole, np.nan, s["u"])
    vo = np.where(hole, np.nan, s["v"])
    uo, vo, out = inpaint_oracle.funRemoveOutliers(uo, vo, 100.0)
    assert hole.sum() > 20 and np.array_equal(np.isnan(u), np.isnan(uo))
    fin = ~np.isnan(uo)
    # sub-pixel peaks within the 1e-3 rounding of findpeak, as the search test
    assert np.abs(u[fin] - uo[fin]).max() <= 1e-3 + 1e-9
    assert np.abs(v[fin] - vo[fin]).max() <= 1e-3 + 1e-9
    assert list(P["SizeOfFFTSearchRegion"]) == [10, 10] and cc["ccThreshold"] == 1.25
