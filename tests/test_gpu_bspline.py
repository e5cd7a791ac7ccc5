"""The interpolating cubic B-spline option (aldic_images_set_interp,
aldic_bspline_prefilter_dev) against the FP64 oracle.  No counterpart on the
reference's path: its MEX evaluates a local kernel on raw pixels
(ba_interp2_spline.cpp:310-325); this option evaluates the exact spline whose
coefficients that kernel approximates cell by cell.

Tolerances: the FP32 coefficients within 2e-5 of the FP64 oracle's on
unit-variance data (the device recursion starts 16 samples outside each tile
from a zero state, 0.268^16 = 7e-10, so FP32 rounding dominates); IC-GN and
Subpb1 in this mode to the same bar as the reference kernel's parity tests
(|dU| <= 1e-3 px where both converge, equal failure statuses, iteration
counts one apart on a few percent of subsets)."""

import numpy as np
import pytest

from _cases import c1_grid, c1_pair, hole_mask, small_pair
from test_gpu_parity import _cmp_icgn

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape", [(300, 257), (64, 64), (70, 5), (1, 40)])
def test_prefilter_matches_the_oracle(pkg, orc, shape):
    api = pkg.api
    ctx = api.Context(0)
    rng = np.random.default_rng(sum(shape))
    A = rng.normal(size=shape)
    M, N = shape
    if M > 60 and N > 60:
        # a zeroed hole and a zero block on the border: the zero rule
        I, J = np.meshgrid(np.arange(M), np.arange(N), indexing="ij")
        A[(I - M // 2) ** 2 + (J - N // 3) ** 2 < 400] = 0.0
        A[:6, N - 9:] = 0.0
        A[M // 3, 2] = 0.0  # a lone zero: no effect
    A = np.asfortranarray(A)
    d_in = api.DeviceArray.from_host(ctx, A, np.float32)
    d_out = api.DeviceArray(ctx, M * N, np.float32)
    pkg._capi.check(pkg.lib().aldic_bspline_prefilter_dev(ctx.ptr, d_in.ptr, d_out.ptr, M, N))
    got = d_out.to_host().reshape((M, N), order="F")
    want = orc.bspline_prefilter(A.astype(np.float32).astype(np.float64))
    d = np.abs(got - want).max()
    print(f"{shape}: max |coef - oracle| {d:.2e}, zero coefficients {(want == 0).sum()}")
    assert d <= 2e-5
    assert np.array_equal(got == 0.0, want == 0.0)
    # overlapping input and output are refused (the tiles read aprons)
    with pytest.raises(pkg.AldicError):
        pkg._capi.check(pkg.lib().aldic_bspline_prefilter_dev(ctx.ptr, d_in.ptr, d_in.ptr, M, N))


def test_images_interp_mode_switch(pkg, orc):
    """IC-GN on the same images in both modes: 'local' is the reference's
    kernel (bit-identical to images never switched), 'bspline' matches the
    oracle's spline IC-GN, and switching back restores the first result."""
    f, g, uf, vf = small_pair(M=200, N=180, seed=3, shift=(0.35, -0.45), strain=0.002)
    rng = np.random.default_rng(5)
    n = 60
    cx = rng.integers(40, 161, n).astype(float)
    cy = rng.integers(40, 141, n).astype(float)
    coords = np.asfortranarray(np.stack([cx, cy], 1))
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(cx, cy))
    U0[1::2] = np.round(vf(cx, cy))
    im = pkg.Images(f, g)
    base = pkg.local_icgn_raw(im, coords, U0, 24, 1e-3)
    im.set_interp("bspline")
    got = pkg.local_icgn_raw(im, coords, U0, 24, 1e-3)
    want = orc.local_icgn(orc.Images(f, g, interp="bspline"), coords, U0,
                          np.full(n, 24, np.int32), 1e-3)
    _cmp_icgn(got, want, conv_tol_frac=0.05)
    # the two interpolants give different displacements (by about the
    # local kernel's bias)
    ok = (base[3] >= 1) & (base[3] <= 100) & (got[3] >= 1) & (got[3] <= 100)
    assert np.abs(base[0] - got[0]).reshape(-1, 2)[ok].max() > 1e-4
    im.set_interp("local")
    again = pkg.local_icgn_raw(im, coords, U0, 24, 1e-3)
    for a, b in zip(base, again):
        assert np.array_equal(a, b, equal_nan=True)
    with pytest.raises(ValueError):
        im.set_interp("cubic")
    with pytest.raises(pkg.AldicError):
        pkg._capi.check(pkg.lib().aldic_images_set_interp(im.ptr, 7))


@pytest.mark.parametrize("ws", [16, 32, 64])
def test_local_icgn_bspline_full_class(pkg, orc, ws):
    """The compile-time-size wave kernels in spline mode."""
    f, g, uf, vf = small_pair(M=256, N=240, seed=10, shift=(0.6, -0.2), strain=0.002)
    rng = np.random.default_rng(ws + 1)
    n = 80
    cx = rng.integers(40, 217, n).astype(float)
    cy = rng.integers(40, 201, n).astype(float)
    coords = np.asfortranarray(np.stack([cx, cy], 1))
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(cx, cy))
    U0[1::2] = np.round(vf(cx, cy))
    im = pkg.Images(f, g).set_interp("bspline")
    got = pkg.local_icgn_raw(im, coords, U0, ws, 1e-3)
    want = orc.local_icgn(orc.Images(f, g, interp="bspline"), coords, U0,
                          np.full(n, ws, np.int32), 1e-3)
    _cmp_icgn(got, want, conv_tol_frac=0.05)


def test_local_icgn_bspline_mask_branch(pkg, orc):
    """Partly masked subsets (the workgroup kernel) in spline mode."""
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
    im = pkg.Images(fm, gm, mask).set_interp("bspline")
    got = pkg.local_icgn_raw(im, coords, U0, 16, 1e-3)
    want = orc.local_icgn(orc.Images(fm, gm, mask, interp="bspline"), coords, U0, 16, 1e-3)
    assert np.sum(want[3] == 102) > 5
    _cmp_icgn(got, want, conv_tol_frac=0.05)


def test_subpb1_bspline(pkg, orc):
    f, g, uf, vf = c1_pair()
    coords, nX, nY = c1_grid()
    n = coords.shape[0]
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(coords[:, 0], coords[:, 1]))
    U0[1::2] = np.round(vf(coords[:, 0], coords[:, 1]))
    oim = orc.Images(f, g, interp="bspline")
    want0 = orc.local_icgn(oim, coords, U0, 30, 1e-3, nthreads=8)
    UOld = np.nan_to_num(want0[0])
    FOld = np.nan_to_num(want0[1])
    rng = np.random.default_rng(13)
    vdual = rng.normal(scale=0.02, size=2 * n)
    udual = np.zeros(4 * n)
    wl = np.full((n, 2), 30, dtype=np.int32)
    im = pkg.Images(f, g).set_interp("bspline")
    U, conv = pkg.subpb1_raw(im, coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2)
    Uo, convo = orc.subpb1(oim, coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2, nthreads=8)
    ok = (convo >= 1) & (convo <= 100) & (conv >= 1) & (conv <= 100)
    assert np.array_equal(convo == 0, conv == 0)
    assert np.mean(conv != convo) <= 0.03
    d = np.abs(U - Uo).reshape(-1, 2)[ok].max()
    print(f"Subpb1 spline mode: {ok.sum()} nodes, max |dU| {d:.2e} px")
    assert d <= 1e-3


def test_driver_interp_method_and_accuracy(pkg, synth):
    """DICpara["InterpMethod"] = "bspline" through LocalICGNQuadtree, on a
    noise-free, analytically rendered sub-pixel translation: the spline's
    error is far below the reference kernel's bias."""
    M = N = 256
    u0, v0 = 0.3, -0.2
    f8, g8, _ = synth.make_pair(M, N, lambda X, Y: u0 + 0 * X, lambda X, Y: v0 + 0 * X,
                                density=0.02, seed=4, noise=0.0, bits=12)
    f = np.asfortranarray((f8 - f8.mean()) / f8.std())
    g = np.asfortranarray((g8 - g8.mean()) / g8.std())
    xs = np.arange(48.0, 209.0, 16.0)
    X, Y = np.meshgrid(xs, xs, indexing="ij")
    coords = np.asfortranarray(np.stack([X.ravel(), Y.ravel()], 1))
    U0 = np.zeros(2 * len(coords))
    Df = pkg.funImgGradient(f, g)
    err = {}
    for mode in ("local", "bspline"):
        U, F, H, t, conv, bad, hole = pkg.LocalICGNQuadtree(
            U0, coords, Df, f, g, {"winsize": 40, "InterpMethod": mode}, "GaussNewton", 1e-4, 1)
        err[mode] = np.abs(np.stack([U[0::2] - u0, U[1::2] - v0])).max()
    print(f"max |U - truth|: local {err['local']:.2e} px, spline {err['bspline']:.2e} px")
    assert err["local"] > 1e-3 and err["bspline"] < err["local"] / 8
