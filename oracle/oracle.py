"""ctypes view of the FP64 oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.

Arrays follow the reference's MATLAB layout: an image is a numpy array of
shape (M, N) in Fortran order whose first axis is image x
(gui/loadStereoImages.m:79 loads images transposed), so ``img[i-1, j-1]`` is
MATLAB's ``img(i, j)``.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="F_CONTIGUOUS")
_i64 = C.c_int64


class _Images(C.Structure):
    _fields_ = [
        ("ImgRef", C.c_void_p),
        ("ImgDef", C.c_void_p),
        ("mask", C.c_void_p),
        ("DfDx", C.c_void_p),
        ("DfDy", C.c_void_p),
        ("M", C.c_int64),
        ("N", C.c_int64),
        ("ImgDefCoef", C.c_void_p),
    ]


def build() -> str:
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.orc_interp_spline_d.argtypes = [_dp, _i64, _i64, _i64, _dp, _dp, _i64, _dp]
        L.orc_interp_spline_f.argtypes = [_fp, _i64, _i64, _i64, _fp, _fp, _i64, _fp]
        L.orc_img_gradient.argtypes = [_dp, C.c_void_p, _i64, _i64, _dp, _dp]
        L.orc_normalize.argtypes = [_dp, _i64, _i64, _i64, _i64, _i64, _i64, _dp]
        L.orc_bspline_prefilter.argtypes = [_dp, _i64, _i64, _dp]
        L.orc_interp_bspline_d.argtypes = [_dp, _i64, _i64, _dp, _dp, _i64, _dp]
        L.orc_icgn_subset.argtypes = [C.POINTER(_Images), _i64, _i64, C.c_int, C.c_double,
                                      _dp, _dp, _dp, _dp, C.POINTER(C.c_double)]
        L.orc_icgn_subset.restype = C.c_int
        L.orc_subpb1_subset.argtypes = [C.POINTER(_Images), _i64, _i64, C.c_int, C.c_int,
                                        C.c_double, _dp, _dp, _dp, _dp, C.c_double, _dp]
        L.orc_subpb1_subset.restype = C.c_int
        L.orc_local_icgn.argtypes = [C.POINTER(_Images), _dp, _dp, _i64, _i32, C.c_int,
                                     C.c_double, C.c_int, _dp, _dp, _dp, _i32, _u8, _dp]
        L.orc_local_icgn_ex.argtypes = [C.POINTER(_Images), _dp, _dp, C.c_void_p, _i64, _i32,
                                        C.c_int, C.c_double, C.c_int, C.c_int, C.c_double,
                                        C.c_int, _dp, _dp, _dp, _i32, _u8, _dp]
        L.orc_subpb1.argtypes = [C.POINTER(_Images), _dp, _dp, _dp, _dp, _dp, _i64, _i32,
                                 C.c_double, C.c_double, C.c_int, _dp, _i32]
        L.orc_normxcorr2_crop.argtypes = [_dp, _i64, _i64, _dp, _i64, _i64, _dp]
        L.orc_normxcorr2_crop.restype = C.c_int
        L.orc_findpeak.argtypes = [_dp, _i64, _i64, C.POINTER(C.c_double), C.POINTER(C.c_double),
                                   C.POINTER(C.c_double), C.POINTER(C.c_int64),
                                   C.POINTER(C.c_int64)]
        L.orc_integer_search.argtypes = [_dp, _dp, _i64, _i64, C.POINTER(C.c_int64),
                                         C.POINTER(C.c_int64), C.c_int, C.c_int,
                                         C.POINTER(C.c_int64), C.c_int, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_integer_search.restype = C.c_int
        L.orc_integer_search_q.argtypes = (L.orc_integer_search.argtypes[:17] + [C.c_void_p] +
                                           L.orc_integer_search.argtypes[17:])
        L.orc_integer_search_q.restype = C.c_int
        L.orc_normxcorr2_full.argtypes = [_dp, _i64, _i64, _dp, _i64, _i64, _dp]
        L.orc_normxcorr2_full.restype = C.c_int
        L.orc_qfactor.argtypes = [_dp, _i64, _dp]
        L.orc_hist_edges.argtypes = [C.c_double, C.c_double, C.c_int, _dp]
        L.orc_bwselect.argtypes = [_u8, _i64, _i64, _i64, _i64, C.c_int, _u8]
        _lib = L
    return _lib


def _F(a, dt=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dt))


def interp_spline(F, X, Y, single=False):
    """ba_interp2_spline(F, X, Y): output has X's shape (+ F's channels)."""
    dt = np.float32 if single else np.float64
    F = _F(F, dt)
    X = _F(X, dt)
    Y = _F(Y, dt)
    assert X.shape == Y.shape
    M, N = F.shape[0], F.shape[1]
    P = int(np.prod(F.shape[2:])) if F.ndim > 2 else 1
    ND = X.size
    Z = np.zeros((ND, P), dtype=dt, order="F")
    fn = lib().orc_interp_spline_f if single else lib().orc_interp_spline_d
    fn(F, M, N, P, X.reshape(-1, order="F"), Y.reshape(-1, order="F"), ND, Z)
    out_shape = X.shape + (F.shape[2:] if F.ndim > 2 else ())
    return Z.reshape(out_shape, order="F")


def img_gradient(img, mask=None):
    img = _F(img)
    M, N = img.shape
    gx = np.zeros((M - 6, N - 6), order="F")
    gy = np.zeros((M - 6, N - 6), order="F")
    m = None
    if mask is not None:
        mask = _F(mask)
        m = mask.ctypes.data
    lib().orc_img_gradient(img, m, M, N, gx, gy)
    return gx, gy


def normalize(img, gridx, gridy):
    img = _F(img)
    out = np.zeros_like(img, order="F")
    lib().orc_normalize(img, img.shape[0], img.shape[1], int(gridx[0]), int(gridx[1]),
                        int(gridy[0]), int(gridy[1]), out)
    return out


def bspline_prefilter(img):
    """Interpolating cubic B-spline coefficients (mirror boundaries)."""
    img = _F(img)
    out = np.zeros_like(img, order="F")
    lib().orc_bspline_prefilter(img, img.shape[0], img.shape[1], out)
    return out


def interp_bspline(coef, X, Y):
    """The spline of coefficients coef at (X, Y): X along dim 2, Y along
    dim 1, 1-based (the addressing of interp_spline)."""
    coef = _F(coef)
    X = np.ascontiguousarray(X, dtype=np.float64).ravel()
    Y = np.ascontiguousarray(Y, dtype=np.float64).ravel()
    Z = np.zeros(X.size)
    lib().orc_interp_bspline_d(coef, coef.shape[0], coef.shape[1], X, Y, X.size, Z)
    return Z


class Images:
    """Holds the arrays an orc_images struct points to (keeps them alive)."""

    def __init__(self, ImgRef, ImgDef, mask=None, DfDx=None, DfDy=None, interp="local"):
        """interp "local": the reference's kernel; "bspline": the
        interpolating cubic B-spline of ImgDef (not the reference's)."""
        self.ImgRef = _F(ImgRef)
        self.ImgDef = _F(ImgDef)
        M, N = self.ImgRef.shape
        self.mask = _F(np.ones((M, N)) if mask is None else mask)
        if DfDx is None or DfDy is None:
            DfDx, DfDy = img_gradient(self.ImgRef, self.mask)
        self.DfDx = _F(DfDx)
        self.DfDy = _F(DfDy)
        self.M, self.N = M, N
        if interp not in ("local", "bspline"):
            raise ValueError(f"interp {interp!r}: 'local' or 'bspline'")
        self.coef = bspline_prefilter(self.ImgDef) if interp == "bspline" else None
        self.s = _Images(self.ImgRef.ctypes.data, self.ImgDef.ctypes.data, self.mask.ctypes.data,
                         self.DfDx.ctypes.data, self.DfDy.ctypes.data, M, N,
                         None if self.coef is None else self.coef.ctypes.data)


def icgn_subset(im: Images, x0, y0, winsize, tol, U0):
    U = np.zeros(2, order="F")
    F = np.zeros(4, order="F")
    H = np.zeros(21, order="F")
    z = C.c_double(0)
    st = lib().orc_icgn_subset(C.byref(im.s), int(x0), int(y0), int(winsize), float(tol),
                               _F(U0), U, F, H, C.byref(z))
    return U, F, st, H, z.value


def subpb1_subset(im: Images, x0, y0, wx, wy, mu, udual, vdual, UOld, FOld, tol):
    U = np.zeros(2, order="F")
    st = lib().orc_subpb1_subset(C.byref(im.s), int(x0), int(y0), int(wx), int(wy), float(mu),
                                 _F(udual), _F(vdual), _F(UOld), _F(FOld), float(tol), U)
    return U, st


def local_icgn(im: Images, coords, U0, winsize, tol, nthreads=1, F0=None, max_iter=None,
               stop_rule=0, tol_dp=0.0):
    """LocalICGNQuadtree's per-node loop (no bad-point rule, no NaN fill).

    F0 (4n, optional) starts each node at P0 = [F0 U0]; max_iter (>= 0)
    runs the loop while k <= max_iter (None: the reference's 100), i.e. at
    most max_iter + 1 evaluations; stop_rule 1 replaces the reference's
    relative ||b|| rule by the step-size rule ||dP|| < tol_dp px, stop_rule 2
    by its a-posteriori form (also converged when the last two updates'
    measures satisfy m_k <= m_{k-1} / 2 and m_k^2 < tol_dp (m_{k-1} - m_k)).
    Returns U (2n), F (4n), H (n x 21), convIt (n), markHole (n), zncc (n)."""
    coords = _F(coords)
    n = coords.shape[0]
    ws = np.asarray(winsize, dtype=np.int32).reshape(-1)
    per = 1 if ws.size == n and n != 1 else 0
    ws = _F(ws, np.int32)
    U = np.zeros(2 * n, order="F")
    Fo = np.zeros(4 * n, order="F")
    H = np.zeros((n, 21), order="F")
    conv = np.zeros(n, dtype=np.int32, order="F")
    hole = np.zeros(n, dtype=np.uint8, order="F")
    z = np.zeros(n, order="F")
    if F0 is None and max_iter is None and stop_rule == 0:
        lib().orc_local_icgn(C.byref(im.s), coords, _F(U0), n, ws, per, float(tol),
                             int(nthreads), U, Fo, H, conv, hole, z)
    else:
        f0 = None if F0 is None else _F(np.asarray(F0).reshape(-1))
        if f0 is not None and f0.size != 4 * n:
            raise ValueError("F0 must hold 4 values per node")
        lib().orc_local_icgn_ex(C.byref(im.s), coords, _F(U0), None if f0 is None else
                                f0.ctypes.data, n, ws, per, float(tol),
                                100 if max_iter is None else int(max_iter), int(stop_rule),
                                float(tol_dp),
                                int(nthreads), U, Fo, H, conv, hole, z)
    return U, Fo, H, conv, hole, z


def subpb1(im: Images, coords, UOld, FOld, udual, vdual, winsize_xy, mu, tol, nthreads=1):
    coords = _F(coords)
    n = coords.shape[0]
    U = np.zeros(2 * n, order="F")
    conv = np.zeros(n, dtype=np.int32, order="F")
    lib().orc_subpb1(C.byref(im.s), coords, _F(UOld), _F(FOld), _F(udual), _F(vdual), n,
                     _F(winsize_xy, np.int32), float(mu), float(tol), int(nthreads), U, conv)
    return U, conv


def normxcorr2_crop(T, A):
    T = _F(T)
    A = _F(A)
    s1, s2 = T.shape
    W1, W2 = A.shape
    out = np.zeros((W1 - s1 + 3, W2 - s2 + 3), order="F")
    rc = lib().orc_normxcorr2_crop(T, s1, s2, A, W1, W2, out)
    if rc != 0:
        raise ValueError("normxcorr2: template values are all the same")
    return out


def normxcorr2_full(T, A):
    """The full (W1 + s1 - 1) x (W2 + s2 - 1) normxcorr2 map."""
    T = _F(T)
    A = _F(A)
    s1, s2 = T.shape
    W1, W2 = A.shape
    out = np.zeros((W1 + s1 - 1, W2 + s2 - 1), order="F")
    if lib().orc_normxcorr2_full(T, s1, s2, A, W1, W2, out) != 0:
        raise ValueError("normxcorr2: template values are all the same")
    return out


def qfactor(cc):
    """(PCE, PPE) of one correlation map (compute_qFactor)."""
    cc = _F(np.asarray(cc).reshape(-1))
    q = np.zeros(2, order="F")
    lib().orc_qfactor(cc, cc.size, q)
    return q


def hist_edges(xmin, xmax, nbins):
    e = np.zeros(nbins + 1, order="F")
    lib().orc_hist_edges(float(xmin), float(xmax), int(nbins), e)
    return e


def findpeak(f):
    f = _F(f)
    xp, yp, mf = C.c_double(), C.c_double(), C.c_double()
    pr, pc = C.c_int64(), C.c_int64()
    lib().orc_findpeak(f, f.shape[0], f.shape[1], C.byref(xp), C.byref(yp), C.byref(mf),
                       C.byref(pr), C.byref(pc))
    return xp.value, yp.value, mf.value, pr.value, pc.value


def integer_search(f, g, gridx, gridy, winsize, winstepsize, region, nthreads=1,
                   qfactors=False):
    """funIntegerSearchWholeField with region = [L R U D] already expanded.

    Returns dict of nY x nX arrays x, y, u, v, ccmax, ipeak_row, ipeak_col,
    and with qfactors=True also "qfactors" (nX*nY x 2, cc.qfactors)."""
    f = _F(f)
    g = _F(g)
    M, N = f.shape
    gx = (C.c_int64 * 2)(*[int(v) for v in gridx])
    gy = (C.c_int64 * 2)(*[int(v) for v in gridy])
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    nX, nY = C.c_int64(), C.c_int64()
    L = lib()
    L.orc_integer_search(f, g, M, N, gx, gy, int(winsize), int(winstepsize), rg, 1, None, None,
                         None, None, None, None, None, C.byref(nX), C.byref(nY))
    shp = (nY.value, nX.value)
    out = {k: np.zeros(shp, order="F") for k in ("x", "y", "u", "v", "ccmax")}
    out["ipeak_row"] = np.zeros(shp, dtype=np.int32, order="F")
    out["ipeak_col"] = np.zeros(shp, dtype=np.int32, order="F")
    q = np.zeros((shp[0] * shp[1], 2), order="F") if qfactors else None
    L.orc_integer_search_q(f, g, M, N, gx, gy, int(winsize), int(winstepsize), rg,
                           int(nthreads), out["x"].ctypes.data, out["y"].ctypes.data,
                           out["u"].ctypes.data, out["v"].ctypes.data, out["ccmax"].ctypes.data,
                           out["ipeak_row"].ctypes.data, out["ipeak_col"].ctypes.data,
                           None if q is None else q.ctypes.data, C.byref(nX), C.byref(nY))
    if qfactors:
        out["qfactors"] = q
    return out


def bwselect(BW, r, c, conn):
    BW = _F(BW, np.uint8)
    out = np.zeros_like(BW, order="F")
    lib().orc_bwselect(BW, BW.shape[0], BW.shape[1], int(r), int(c), int(conn), out)
    return out.astype(bool)
