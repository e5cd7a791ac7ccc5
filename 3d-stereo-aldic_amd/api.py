"""Host-side mirror of the reference's MATLAB interface for the local matching
path, over the C ABI of libaldic.so (HIP on gfx950).

Names, argument order, argument meaning and outputs follow the reference:

* ``ba_interp2_spline(F, X, Y)``                         ba_interp2_spline.cpp:418-435
* ``funImgGradient(ImgRef, ImgDef, ImgRefMask)``           func/funImgGradient.m:1-87
* ``IntegerSearchQuadtree(ImgRef, ImgDef, file_name, DICpara, ImgSeqNum)``
                                                          func_quadtree/IntegerSearchQuadtree.m:1-228
* ``funIntegerSearch(f, g, ...)`` (whole field)            func/funIntegerSearch.m:2-98
* ``LocalICGNQuadtree(U0, coordinatesFEM, Df, ImgRef, ImgDef, DICpara, ICGNmethod, tol,
  shapeFunctionOrder)``                                   func_quadtree/LocalICGNQuadtree.m:1-349
* ``Subpb1Quadtree(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta,
  HPar, ALSolveStep, DICpara, ICGNmethod, tol)``          func_quadtree/Subpb1Quadtree.m:1-276

Arrays use MATLAB's layout: an image is ``(M, N)`` with the first axis along
image x, node coordinates are 1-based, vectors are flat.  ``Df`` and
``DICpara`` are dicts with the reference's field names.

Differences that a caller sees, all stated where they occur:
* the natural-neighbour NaN fill after the bad-point rules
  (LocalICGNQuadtree.m:200-325, Subpb1Quadtree.m:189-254) runs on the host in
  libaldic (:func:`nan_fill`); where sites tie for the nearest one outside
  the hull it takes the first, and degenerate (collinear) sites in the
  global pass raise an error;
* funRemoveOutliers and inpaint_nans run on the host (:mod:`.outliers`).
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from . import _capi
from ._capi import check, lib
from .outliers import funMedRemoveOutliers, funRemoveOutliers, inpaint_nans, medFilt2  # noqa: F401


def _F(a, dt=np.float64):
    return np.asfortranarray(np.asarray(a, dtype=dt))


class Context:
    """One HIP device + stream + reusable staging buffers (aldic_ctx)."""

    def __init__(self, device: int = 0):
        p = C.c_void_p()
        check(lib().aldic_ctx_create(int(device), C.byref(p)))
        self.ptr = p
        self.device = device

    @property
    def stream(self) -> int:
        return lib().aldic_ctx_stream(self.ptr) or 0

    def synchronize(self):
        check(lib().aldic_ctx_synchronize(self.ptr))

    def set_search_method(self, method):
        """Integer-search peak finding on this context: "auto" (default:
        FFT lag blocks for whole grids of at least 1024 points with lag
        windows of at least 32 x 32), "direct" or "fft"
        (aldic_ctx_set_search_method)."""
        modes = {"direct": 0, "fft": 1, "auto": 2}
        if method not in modes:
            raise ValueError(f"search method {method!r}: 'auto', 'direct' or 'fft'")
        check(lib().aldic_ctx_set_search_method(self.ptr, modes[method]))
        self.search_method = method
        return self

    def record(self, mark: int):
        """Mark the point reached by the work submitted so far."""
        check(lib().aldic_ctx_record(self.ptr, int(mark)))

    def wait(self, other: "Context", mark: int):
        """Make later work on this context wait for `mark` of `other`."""
        check(lib().aldic_ctx_wait(self.ptr, other.ptr, int(mark)))

    def capture(self, fn):
        """Record everything fn() submits to this context as a Graph (not
        run)."""
        check(lib().aldic_ctx_capture_begin(self.ptr))
        try:
            fn()
        except BaseException:
            p = C.c_void_p()
            if lib().aldic_ctx_capture_end(self.ptr, C.byref(p)) == 0:
                lib().aldic_graph_destroy(p)
            raise
        p = C.c_void_p()
        check(lib().aldic_ctx_capture_end(self.ptr, C.byref(p)))
        return Graph(p)

    def close(self):
        if self.ptr:
            lib().aldic_ctx_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Graph:
    """A captured sequence of device work (aldic_graph)."""

    def __init__(self, ptr):
        self.ptr = ptr

    def launch(self, ctx: Context):
        check(lib().aldic_graph_launch(self.ptr, ctx.ptr))

    def close(self):
        if getattr(self, "ptr", None):
            lib().aldic_graph_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = None


def default_context() -> Context:
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = Context(0)
    return _default_ctx


# ---------------------------------------------------------------------------
# ba_interp2_spline
# ---------------------------------------------------------------------------
def ba_interp2_spline(F, X, Y, method=None, ctx: Context | None = None):
    """Z = ba_interp2_spline(F, X, Y, [method]).

    F is M x N (x P); X (dim-2 coordinate) and Y (dim-1 coordinate) have the
    same shape; all three double or all three single (else an error, as
    ba_interp2_spline.cpp:427-434).  ``method`` is ignored like the MEX's."""
    ctx = ctx or default_context()
    F = np.asarray(F)
    X = np.asarray(X)
    Y = np.asarray(Y)
    if X.shape != Y.shape:
        raise ValueError("X, Y should have the same size")
    kinds = {F.dtype, X.dtype, Y.dtype}
    if kinds == {np.dtype(np.float64)}:
        dt, fn = np.float64, lib().aldic_interp_spline_f64
    elif kinds == {np.dtype(np.float32)}:
        dt, fn = np.float32, lib().aldic_interp_spline_f32
    else:
        raise TypeError("ba_interp2 takes only double or single arguments for IMG,X,Y, "
                        "and all must have the same class")
    F = _F(F, dt)
    M, N = F.shape[0], F.shape[1]
    P = int(np.prod(F.shape[2:])) if F.ndim > 2 else 1
    ND = X.size
    Z = np.zeros((ND, P), dtype=dt, order="F")
    check(fn(ctx.ptr, F, M, N, P, _F(X, dt).reshape(-1, order="F"),
             _F(Y, dt).reshape(-1, order="F"), ND, Z))
    return Z.reshape(X.shape + (F.shape[2:] if F.ndim > 2 else ()), order="F")


# ---------------------------------------------------------------------------
# funImgGradient / funNormalizeImg
# ---------------------------------------------------------------------------
def funImgGradient(ImgRef, ImgDef=None, ImgRefMask=None, ctx: Context | None = None):
    """Df = funImgGradient(ImgRef, ImgDef, ImgRefMask) (finite-difference branch)."""
    ctx = ctx or default_context()
    ImgRef = _F(ImgRef)
    M, N = ImgRef.shape
    gx = np.zeros((M - 6, N - 6), order="F")
    gy = np.zeros((M - 6, N - 6), order="F")
    mask = None if ImgRefMask is None else _F(ImgRefMask)
    check(lib().aldic_img_gradient(ctx.ptr, ImgRef, None if mask is None else mask.ctypes.data,
                                   M, N, gx, gy))
    m = np.ones((M, N)) if mask is None else mask
    return {"DfDx": gx, "DfDy": gy, "DfCropWidth": 3, "imgSize": (M, N),
            "DfAxis": [4, M - 3, 4, N - 3], "ImgRefMask": m}


def funNormalizeImg(Img, gridxyROIRange):
    """[ImgNormalized, gridxyROIRange] = funNormalizeImg(Img, gridxyROIRange)
    (func/funNormalizeImg.m:22-40): (Img - mean_ROI) / std_ROI per image."""
    gx = list(gridxyROIRange["gridx"])
    gy = list(gridxyROIRange["gridy"])
    M, N = np.asarray(Img[0]).shape
    gx[0] = max(gx[0], 1)
    gx[1] = min(gx[1], M)
    gy[0] = max(gy[0], 1)
    gy[1] = min(gy[1], N)
    out = []
    for im in Img:
        im = np.asarray(im, dtype=np.float64)
        roi = im[gx[0] - 1:gx[1], gy[0] - 1:gy[1]]
        out.append(_F((im - roi.mean()) / roi.std(ddof=1)))
    return out, {"gridx": gx, "gridy": gy}


# ---------------------------------------------------------------------------
# resident images
# ---------------------------------------------------------------------------
class Images:
    """The (Df, ImgRef, ImgDef) state of the solvers, resident on the GPU."""

    def __init__(self, ImgRef, ImgDef, ImgRefMask=None, Df=None, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        ImgRef = _F(ImgRef)
        ImgDef = _F(ImgDef)
        M, N = ImgRef.shape
        if ImgDef.shape != (M, N):
            raise ValueError("ImgRef and ImgDef must have the same size")
        mask = None if ImgRefMask is None else _F(ImgRefMask)
        gx = gy = None
        if Df is not None:
            gx = _F(Df["DfDx"])
            gy = _F(Df["DfDy"])
            if gx.shape != (M - 6, N - 6):
                raise ValueError("Df.DfDx must be (M-6) x (N-6) (DfCropWidth = 3)")
        self._keep = (ImgRef, ImgDef, mask, gx, gy)
        p = C.c_void_p()
        check(lib().aldic_images_create(
            self.ctx.ptr, ImgRef, ImgDef, None if mask is None else mask.ctypes.data,
            None if gx is None else gx.ctypes.data, None if gy is None else gy.ctypes.data,
            M, N, C.byref(p)))
        self._keep = None
        self.ptr = p
        self.M, self.N = M, N
        self.interp = "local"

    @classmethod
    def from_device(cls, fm_ptr, g_ptr, M, N, mask_ptr=None, fx_ptr=None, fy_ptr=None,
                    ctx: Context | None = None):
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        p = C.c_void_p()
        check(lib().aldic_images_create_dev(self.ctx.ptr, fm_ptr, g_ptr, mask_ptr, fx_ptr, fy_ptr,
                                            M, N, C.byref(p)))
        self.ptr = p
        self.M, self.N = M, N
        self.interp = "local"
        return self

    @classmethod
    def from_raw(cls, ImgRef, ImgDef, ImgRefMask=None, gridxyROIRange=None,
                 ctx: Context | None = None):
        """Images from raw 8- or 16-bit pixel arrays (H x W, row-major as
        decoded from the file): funNormalizeImg over the ROI and the
        gradients run on the device (aldic_images_create_raw).  The MATLAB
        image is the transpose, M = W by N = H.  Sets ``stats`` to
        (mean_f, std_f, mean_g, std_g)."""
        self = cls.__new__(cls)
        self.ctx = ctx or default_context()
        a = np.ascontiguousarray(ImgRef)
        b = np.ascontiguousarray(ImgDef)
        if a.dtype == np.uint8 and b.dtype == np.uint8:
            bits = 8
        elif a.dtype == np.uint16 and b.dtype == np.uint16:
            bits = 16
        else:
            raise TypeError("raw images must both be uint8 or both uint16")
        if a.shape != b.shape or a.ndim != 2:
            raise ValueError("raw images must be two H x W arrays of one size")
        H, W = a.shape
        M, N = W, H
        roi = gridxyROIRange or {"gridx": [1, M], "gridy": [1, N]}
        gx = np.array(roi["gridx"], dtype=np.int64)
        gy = np.array(roi["gridy"], dtype=np.int64)
        mask = None if ImgRefMask is None else _F(ImgRefMask)
        if mask is not None and mask.shape != (M, N):
            raise ValueError("the mask is M x N = W x H")
        stats = np.zeros(4)
        p = C.c_void_p()
        check(lib().aldic_images_create_raw(self.ctx.ptr, a.ctypes.data, b.ctypes.data, bits,
                                            None if mask is None else mask.ctypes.data, M, N,
                                            gx.ctypes.data, gy.ctypes.data, C.byref(p),
                                            stats.ctypes.data))
        self.ptr = p
        self.M, self.N = M, N
        self.interp = "local"
        self.stats = tuple(stats)
        return self

    def set_interp(self, mode):
        """Interpolant of ImgDef for every solver on these images: "local"
        (the reference's kernel, ba_interp2_spline.cpp:310-325; the default
        and the only mode with reference parity) or "bspline" (the
        interpolating cubic B-spline on prefiltered coefficients; no
        reference counterpart).  Calling it again remakes the
        coefficients."""
        modes = {"local": 0, "bspline": 1}
        if mode not in modes:
            raise ValueError(f"interp {mode!r}: 'local' or 'bspline'")
        check(lib().aldic_images_set_interp(self.ptr, modes[mode]))
        self.interp = mode
        return self

    def planes(self):
        """Host copies of the resident planes: {fm, g, mask, fx, fy}, M x N FP32."""
        ptrs = [C.c_void_p() for _ in range(5)]
        check(lib().aldic_images_planes(self.ptr, *[C.byref(q) for q in ptrs]))
        out = {}
        for name, q in zip(("fm", "g", "mask", "fx", "fy"), ptrs):
            a = np.empty((self.N, self.M), dtype=np.float32)
            check(lib().aldic_memcpy_dtoh(self.ctx.ptr, a.ctypes.data, q, a.nbytes))
            out[name] = a.T
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib().aldic_images_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _images_for(Df, ImgRef, ImgDef, ctx, DICpara=None):
    """Resident images for a driver call.  ``DICpara["InterpMethod"]``
    ("local" by default, or "bspline": see Images.set_interp; no counterpart
    in the reference's DICpara) picks the interpolant of images built here;
    an Images argument keeps its own."""
    if isinstance(ImgRef, Images):
        return ImgRef
    mask = None if Df is None else Df.get("ImgRefMask")
    im = Images(ImgRef, ImgDef, mask, Df, ctx=ctx)
    interp = (DICpara or {}).get("InterpMethod", "local")
    if interp != "local":
        im.set_interp(interp)
    return im


# ---------------------------------------------------------------------------
# per-node solvers (raw outputs, no bad-point rules)
# ---------------------------------------------------------------------------
def local_icgn_raw(images: Images, coordinatesFEM, U0, winsize, tol):
    """One funICGNQuadtree per node (LocalICGNQuadtree.m:67-163).

    Returns U (2n), F (4n), HtempPar (n x 21), ConvItPerEle (n),
    markCoordHoleStrainOrNot (n), zncc (n)."""
    coords = _F(coordinatesFEM)[:, :2]
    coords = _F(coords)
    n = coords.shape[0]
    ws = np.asarray(winsize, dtype=np.int32).reshape(-1)
    per = 1 if ws.size == n and n > 1 else 0
    if not per and ws.size != 1:
        raise ValueError("winsize must be a scalar or one value per node")
    U = np.zeros(2 * n, order="F")
    Fo = np.zeros(4 * n, order="F")
    H = np.zeros((n, 21), order="F")
    conv = np.zeros(n, dtype=np.int32, order="F")
    hole = np.zeros(n, dtype=np.uint8, order="F")
    z = np.zeros(n, order="F")
    check(lib().aldic_local_icgn(images.ctx.ptr, images.ptr, coords, _F(U0).reshape(-1), n,
                                 _F(ws, np.int32), per, float(tol), U, Fo, H.ctypes.data, conv,
                                 hole.ctypes.data, z.ctypes.data))
    return U, Fo, H, conv, hole, z


def subpb1_raw(images: Images, coordinatesFEM, UOld, FOld, udual, vdual, winsize_List, mu, tol):
    """One funICGN_Subpb1quadtree per node (Subpb1Quadtree.m:71-165).

    Returns U (2n) and ConvItPerEle (n)."""
    coords = _F(_F(coordinatesFEM)[:, :2])
    n = coords.shape[0]
    wl = _F(np.asarray(winsize_List, dtype=np.int32).reshape(n, 2), np.int32)
    U = np.zeros(2 * n, order="F")
    conv = np.zeros(n, dtype=np.int32, order="F")
    check(lib().aldic_subpb1(images.ctx.ptr, images.ptr, coords, _F(UOld).reshape(-1),
                             _F(FOld).reshape(-1), _F(udual).reshape(-1), _F(vdual).reshape(-1),
                             n, wl, float(mu), float(tol), U, conv))
    return U, conv


# ---------------------------------------------------------------------------
# bad-point rules of the drivers (host statistics)
# ---------------------------------------------------------------------------
def icgn_bad_points(ConvItPerEle, maxIterNum=100):
    """LocalICGNQuadtree.m:167-192.  Returns (bad index array, LocalICGNBadPtNum)."""
    c = np.asarray(ConvItPerEle, dtype=np.float64)
    row12 = np.union1d(np.flatnonzero(c < 0), np.flatnonzero(c > maxIterNum - 1))
    row3 = np.flatnonzero(c == maxIterNum + 2)
    badnum = len(row12) - len(row3)
    good = np.setdiff1d(np.arange(c.size), row12)
    if good.size >= 3:
        mu = c[good].mean()
        sigma = c[good].std(ddof=1)  # fitdist 'Normal' sigma is the unbiased std
    else:
        mu, sigma = np.inf, 0.0
    row4 = np.flatnonzero(c > max(mu + sigma, 6))
    return np.union1d(row12, row4), badnum


def subpb1_bad_points(ConvItPerEle, maxIterNum=100):
    """Subpb1Quadtree.m:167-182.  Returns (bad index array, bad count)."""
    c = np.asarray(ConvItPerEle, dtype=np.float64)
    row12 = np.union1d(np.flatnonzero(c < 0), np.flatnonzero(c > maxIterNum - 1))
    row3 = np.flatnonzero(c == maxIterNum + 2)
    badnum = len(row12) - len(row3)
    good = np.setdiff1d(np.arange(c.size), row12)
    if good.size > 0:
        m = c[good].mean()
        s = c[good].std(ddof=1) if good.size > 1 else 0.0
        thr = np.nanmax([m + 0.25 * s, 10.0])
    else:
        thr = 10.0  # max([NaN, 10]) = 10
    row4 = np.flatnonzero(c > thr)
    return np.union1d(row12, row4), badnum


# ---------------------------------------------------------------------------
# NaN fill of the drivers (host, natural-neighbour interpolation)
# ---------------------------------------------------------------------------
def scattered_natural(x, y, v, xq, yq):
    """scatteredInterpolant(x, y, v, 'natural', 'nearest') evaluated at
    (xq, yq).  v may hold several value columns (n x nv); the result has the
    shape of xq (times nv columns when v is 2-D)."""
    x = _F(x).reshape(-1)
    y = _F(y).reshape(-1)
    v = np.asarray(v, dtype=np.float64)
    cols = v.reshape(x.size, -1)
    xq = np.asarray(xq, dtype=np.float64)
    shape = xq.shape
    xqf = _F(xq).reshape(-1)
    yqf = _F(yq).reshape(-1)
    if xqf.size != yqf.size or cols.shape[0] != x.size or y.size != x.size:
        raise ValueError("scattered_natural: sizes of x, y, v or xq, yq disagree")
    out = np.zeros((xqf.size, cols.shape[1]), order="F")
    check(lib().aldic_natural_interp(x, y, _F(cols), x.size, cols.shape[1], xqf, yqf, xqf.size,
                                     out))
    return out.reshape(shape) if v.ndim == 1 else out


def nan_fill(coordinatesFEM, U, F=None, mask=None):
    """The NaN fill after the bad-point rules: per 8-connected mask region,
    then over every node (LocalICGNQuadtree.m:200-247, :290-325 with F and
    mask = Df.ImgRefMask; Subpb1Quadtree.m:189-254 with F = None and mask =
    DICpara.ImgRefMask).  Returns a filled copy of U."""
    coords = _F(np.asarray(coordinatesFEM, dtype=np.float64)[:, :2])
    n = coords.shape[0]
    Uo = _F(U).reshape(-1).copy()
    if Uo.size != 2 * n:
        raise ValueError("nan_fill: U must hold 2 values per node")
    Fp = None
    if F is not None:
        Ff = _F(F).reshape(-1)
        if Ff.size != 4 * n:
            raise ValueError("nan_fill: F must hold 4 values per node")
        Fp = Ff.ctypes.data
    mp, M, N = None, 0, 0
    if mask is not None:
        m8 = np.asfortranarray(np.asarray(mask) != 0, dtype=np.uint8)
        M, N = m8.shape
        mp = m8.ctypes.data
    check(lib().aldic_nan_fill(coords, n, Uo, Fp, mp, M, N, None))
    return Uo


# ---------------------------------------------------------------------------
# Subpb2Quadtree: the global FE step
# ---------------------------------------------------------------------------
class FEMesh:
    """The FE Q4 quadtree mesh of the global step (aldic_fem): DICmesh's
    coordinatesFEM (n x 2), elementsFEM (ne x 4 or ne x 8, 1-based, 0 = no
    hanging node), dirichlet (1-based nodes) and neumann (n_edges x 4)."""

    def __init__(self, coordinatesFEM, elementsFEM, dirichlet=None, neumann=None,
                 GaussPtOrder=2, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        coords = _F(np.asarray(coordinatesFEM, dtype=np.float64)[:, :2])
        el = np.asarray(elementsFEM, dtype=np.int64)
        if el.ndim != 2 or el.shape[1] not in (4, 8):
            raise ValueError("elementsFEM must be ne x 4 or ne x 8")
        if el.shape[1] == 4:
            el = np.concatenate([el, np.zeros((el.shape[0], 4), np.int64)], axis=1)
        el = _F(el, np.int32)
        d = np.zeros(0, np.int32) if dirichlet is None else np.asarray(dirichlet).reshape(-1)
        d = _F(d, np.int32)
        nm = None if neumann is None or np.size(neumann) == 0 else _F(np.asarray(neumann)
                                                                        .reshape(-1, 4))
        self.n, self.ne = coords.shape[0], el.shape[0]
        p = C.c_void_p()
        check(lib().aldic_fem_create(self.ctx.ptr, coords, self.n, el.ctypes.data, self.ne,
                                     d.ctypes.data, d.size, None if nm is None else nm.ctypes.data,
                                     0 if nm is None else nm.shape[0], int(GaussPtOrder),
                                     C.byref(p)))
        self.ptr = p

    def info(self):
        nnz, nf, nd = C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().aldic_fem_info(self.ptr, C.byref(nnz), C.byref(nf), C.byref(nd)))
        return {"nnz": nnz.value, "free": nf.value, "dirichlet": nd.value}

    def solve(self, beta, mu, U, F, alpha=0.0, winstepsize=1.0, return_iterations=False):
        U = _F(U).reshape(-1)
        F = _F(F).reshape(-1)
        if U.size != 2 * self.n or F.size != 4 * self.n:
            raise ValueError("U must hold 2 and F 4 values per node")
        out = np.zeros(2 * self.n, order="F")
        it = C.c_int32()
        check(lib().aldic_subpb2(self.ctx.ptr, self.ptr, float(beta), float(mu), float(alpha),
                                 float(winstepsize), U, F, out, C.byref(it)))
        return (out, it.value) if return_iterations else out

    def solve_dev(self, ctx: Context, beta, mu, d_U, d_F, d_Uhat, alpha=0.0, winstepsize=1.0,
                  max_iter=0):
        it = C.c_int32()
        check(lib().aldic_subpb2_dev(ctx.ptr, self.ptr, float(beta), float(mu), float(alpha),
                                     float(winstepsize), d_U, d_F, d_Uhat, int(max_iter),
                                     C.byref(it)))
        return it.value

    def solve_dev_async(self, ctx: Context, beta, mu, d_U, d_F, d_Uhat, alpha=0.0,
                        winstepsize=1.0, max_iter=64, d_iterations=None):
        """solve_dev without a host synchronisation (recordable in a graph):
        every chunk of 16 CG iterations up to max_iter is launched, and the
        count lands in the device int d_iterations."""
        check(lib().aldic_subpb2_dev_async(ctx.ptr, self.ptr, float(beta), float(mu), float(alpha),
                                           float(winstepsize), d_U, d_F, d_Uhat, int(max_iter),
                                           d_iterations))

    def close(self):
        if getattr(self, "ptr", None):
            lib().aldic_fem_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def Subpb2Quadtree(DICmesh, GaussPtOrder, beta, mu, U, F, udual, vdual, alpha, winstepsize,
                   waitBarDisplayOrNot=0, ctx: Context | None = None):
    """Uhat = Subpb2Quadtree(DICmesh, GaussPtOrder, beta, mu, U, F, udual, vdual,
    alpha, winstepsize, waitBarDisplayOrNot) (func_quadtree/Subpb2Quadtree.m:
    1-397).  ``DICmesh`` is a dict with coordinatesFEM, elementsFEM,
    dirichlet and neumann, or an :class:`FEMesh` built once for the mesh.
    udual and vdual are accepted and unused: the reference multiplies both
    by zero (:42-44)."""
    mesh = DICmesh
    own = False
    if not isinstance(mesh, FEMesh):
        mesh = FEMesh(DICmesh["coordinatesFEM"], DICmesh["elementsFEM"],
                      DICmesh.get("dirichlet"), DICmesh.get("neumann"), GaussPtOrder, ctx=ctx)
        own = True
    try:
        return mesh.solve(beta, mu, U, F, alpha, winstepsize)
    finally:
        if own:
            mesh.close()


# ---------------------------------------------------------------------------
# MATLAB-level drivers
# ---------------------------------------------------------------------------
def LocalICGNQuadtree(U0, coordinatesFEM, Df, ImgRef, ImgDef, DICpara, ICGNmethod="GaussNewton",
                      tol=1e-3, shapeFunctionOrder=1, ctx: Context | None = None,
                      fill_nans=True):
    """[U,F,HtempPar,LocalTime,ConvItPerEle,LocalICGNBadPtNum,markCoordHoleStrain] =
    LocalICGNQuadtree(U0,coordinatesFEM,Df,ImgRef,ImgDef,DICpara,ICGNmethod,tol,
    shapeFunctionOrder)

    ``DICpara["winsize"]`` may also be an array with one winsize per node.
    markCoordHoleStrain is 0-based (numpy indexing).  fill_nans=False
    returns U with the bad nodes still NaN (before :200)."""
    if shapeFunctionOrder != 1:
        raise NotImplementedError("only first-order IC-GN (shapeFunctionOrder = 1) is built")
    if ICGNmethod != "GaussNewton":
        raise NotImplementedError("only ICGNmethod = 'GaussNewton' is used by the reference")
    images = _images_for(Df, ImgRef, ImgDef, ctx, DICpara)
    t0 = time.perf_counter()
    U, F, H, conv, hole, _z = local_icgn_raw(images, coordinatesFEM, U0, DICpara["winsize"], tol)
    LocalTime = time.perf_counter() - t0
    bad, badnum = icgn_bad_points(conv)
    U = U.copy()
    U[2 * bad] = np.nan
    U[2 * bad + 1] = np.nan
    if fill_nans:
        mask = None if Df is None else Df.get("ImgRefMask")
        if mask is None:
            mask = np.ones((images.M, images.N))
        U = nan_fill(coordinatesFEM, U, F, mask)
    return U, F, H, LocalTime, conv.astype(np.float64), badnum, np.flatnonzero(hole == 1)


def Subpb1Quadtree(UOld, FOld, udual, vdual, coordinatesFEM, Df, ImgRef, ImgDef, mu, beta, HPar,
                   ALSolveStep, DICpara, ICGNmethod="GaussNewton", tol=1e-2,
                   ctx: Context | None = None, fill_nans=True):
    """[U,HPar,ALSub1Time,ConvItPerEle,LocalICGNBadPtNum] = Subpb1Quadtree(UOld,FOld,udual,
    vdual,coordinatesFEM,Df,ImgRef,ImgDef,mu,beta,HPar,ALSolveStep,DICpara,ICGNmethod,tol)

    ``beta`` and ``HPar`` are accepted and unused, as in the reference's
    quadtree variant (the local H is rebuilt inside each subset).  The
    ``ALSolveStep == 1`` branch of the reference reads an undefined variable
    (Subpb1Quadtree.m:101-107) and is rejected.  The region pass of the NaN
    fill reads ``DICpara["ImgRefMask"]`` (all ones when absent);
    fill_nans=False returns U with the bad nodes still NaN."""
    if ALSolveStep == 1:
        raise ValueError("ALSolveStep == 1 is not reachable in the reference (undefined Htemp)")
    images = _images_for(Df, ImgRef, ImgDef, ctx, DICpara)
    n = np.asarray(coordinatesFEM).shape[0]
    wl = DICpara.get("winsize_List")
    if wl is None:
        wl = np.full((n, 2), int(DICpara["winsize"]), dtype=np.int32)
    t0 = time.perf_counter()
    U, conv = subpb1_raw(images, coordinatesFEM, UOld, FOld, udual, vdual, wl, mu, tol)
    ALSub1Time = time.perf_counter() - t0
    bad, badnum = subpb1_bad_points(conv)
    U = U.copy()
    U[2 * bad] = np.nan
    U[2 * bad + 1] = np.nan
    if fill_nans:
        mask = DICpara.get("ImgRefMask")
        if mask is None:
            mask = np.ones((images.M, images.N))
        U = nan_fill(coordinatesFEM, U, None, mask)
    return U, HPar, ALSub1Time, conv.astype(np.float64), badnum


def funIntegerSearch(f, g, tempSizeOfSearchRegion, gridx, gridy, winsize, winstepsize,
                     tempNoOfInitPt=0, ctx: Context | None = None, return_peaks=False):
    """[x,y,u,v,cc] = funIntegerSearch(f,g,tempSizeOfSearchRegion,gridx,gridy,winsize,
    winstepsize,0) -- whole-field branch (func/funIntegerSearch.m:7-46, :100-203).

    Grids are nY x nX.  cc = {"max": ..., "A": [], "qfactors": (nX*nY) x 2}
    with the q-factors in the reference's point order (x index fastest)."""
    if tempNoOfInitPt != 0:
        raise NotImplementedError("only the whole-field search (tempNoOfInitPt = 0) is built")
    ctx = ctx or default_context()
    f = _F(f)
    g = _F(g)
    M, N = f.shape
    r = np.asarray(tempSizeOfSearchRegion, dtype=np.int64).reshape(-1)
    if r.size == 1:
        r = np.array([r[0], r[0]])
    region = np.concatenate([r, r])[:4]  # repmat([sx sy],1,2) -> [L R U D] = [sx sy sx sy]
    ws = int(np.asarray(winsize).reshape(-1)[0])
    st = int(np.asarray(winstepsize).reshape(-1)[0])
    gx = (C.c_int64 * 2)(int(gridx[0]), int(gridx[-1]))
    gy = (C.c_int64 * 2)(int(gridy[0]), int(gridy[-1]))
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    nX, nY = C.c_int64(), C.c_int64()
    L = lib()
    check(L.aldic_integer_search(ctx.ptr, f, g, M, N, gx, gy, ws, st, rg, None, None, None, None,
                                 None, None, None, None, C.byref(nX), C.byref(nY)))
    shp = (nY.value, nX.value)
    out = {k: np.zeros(shp, order="F") for k in ("x", "y", "u", "v", "max")}
    pr = np.zeros(shp, dtype=np.int32, order="F")
    pc = np.zeros(shp, dtype=np.int32, order="F")
    q = np.zeros((shp[0] * shp[1], 2), order="F")
    if shp[0] * shp[1] > 0:
        check(L.aldic_integer_search(ctx.ptr, f, g, M, N, gx, gy, ws, st, rg,
                                     out["x"].ctypes.data, out["y"].ctypes.data,
                                     out["u"].ctypes.data, out["v"].ctypes.data,
                                     out["max"].ctypes.data, pr.ctypes.data, pc.ctypes.data,
                                     q.ctypes.data, C.byref(nX), C.byref(nY)))
    cc = {"max": out["max"], "A": [], "qfactors": q}
    if return_peaks:
        return out["x"], out["y"], out["u"], out["v"], cc, pr, pc
    return out["x"], out["y"], out["u"], out["v"], cc


def params_from_dicpara(DICpara):
    """An aldic_params struct from a DICpara dict: defaults
    (aldic_params_default, setDICparaDefaults.m) overridden by the fields the
    dict holds; empty fields stay empty."""
    from ._capi import AldicParams
    p = AldicParams()
    check(lib().aldic_params_default(C.byref(p)))
    for key in ("winsize", "winstepsize", "winsizeMin", "fixSearchDistanceOrNot",
                "InitFFTSearchMethod", "DICIncOrNot", "UseGlobal", "GaussPtOrder", "ClusterNo"):
        v = DICpara.get(key)
        if v is not None and np.size(v) > 0:
            setattr(p, key, int(np.asarray(v).reshape(-1)[0]))
    for key in ("mu", "beta", "alpha"):
        v = DICpara.get(key)
        if v is not None and np.size(v) > 0:
            setattr(p, key, float(np.asarray(v).reshape(-1)[0]))
    roi = DICpara.get("gridxyROIRange") or {}
    for key in ("gridx", "gridy"):
        v = roi.get(key)
        if v is not None and np.size(v) >= 2:
            v = np.asarray(v).reshape(-1)
            getattr(p, key)[0], getattr(p, key)[1] = int(v[0]), int(v[-1])
    for key in ("StereoSearchDistance", "TemporalSearchDistance", "NewFFTSearchDistance"):
        v = _preset(DICpara, key)
        if v is None:
            continue
        v = np.asarray(v).reshape(-1)
        if v.size not in (1, 2, 4):
            raise ValueError(f"{key} must be empty, a scalar, [x y] or [L R U D]")
        for k in range(v.size):
            getattr(p, key)[k] = int(v[k])
        setattr(p, "n" + key, int(v.size))
    return p


def validate_dicpara(DICpara):
    """validateDICpara.m's rules on the path's fields (aldic_params_validate);
    raises AldicError with the reference's message."""
    check(lib().aldic_params_validate(C.byref(params_from_dicpara(DICpara))))


def search_region(DICpara, ImgSeqNum):
    """[L R U D] of IntegerSearchQuadtree.m:52-90 and funIntegerSearch.m:12
    (aldic_params_search_region)."""
    region = (C.c_int64 * 4)()
    check(lib().aldic_params_search_region(C.byref(params_from_dicpara(DICpara)), int(ImgSeqNum),
                                           region))
    return [int(v) for v in region]


def _preset(DICpara, key):
    """DICpara.(key) when the field exists and is not empty (isfield and
    ~isempty), else None."""
    v = DICpara.get(key)
    if v is None or np.size(v) == 0:
        return None
    return v


def IntegerSearchQuadtree(ImgRef, ImgDef, file_name, DICpara, ImgSeqNum,
                          ctx: Context | None = None):
    """[DICpara,x0,y0,u,v,cc] = IntegerSearchQuadtree(ImgRef,ImgDef,file_name,DICpara,ImgSeqNum)
    with InitFFTSearchMethod = 1 (IntegerSearchQuadtree.m:47-160, :226-228).

    * The search radius (:52-90): stereo (ImgSeqNum == -1) takes
      StereoSearchDistance, else NewFFTSearchDistance; temporal (ImgSeqNum
      >= 2) TemporalSearchDistance, else NewFFTSearchDistance, and sets
      fixSearchDistanceOrNot = 0 when the field is absent.  Where the
      reference would prompt with input() (no preset), this raises.  A
      scalar becomes [s s].
    * funIntegerSearch over gridxyROIRange with winsize/winstepsize (:93).
    * Points whose rounded centre is 0 in ImgRefMask become NaN (:106-115);
      as in the reference's try/catch, a missing ImgSize/ImgRefMask or a
      centre outside ImgSize leaves u, v unmasked.
    * cc.ccThreshold = 1.25 and funRemoveOutliers(u,v,cc,0,100) (:158-160).
    cc also carries q-factors (PCE, PPE per point, normalised over the grid,
    funIntegerSearch.m:162-163, :184-187)."""
    method = DICpara.get("InitFFTSearchMethod", 1)
    if method != 1:
        raise NotImplementedError("only InitFFTSearchMethod = 1 (whole field) is built")
    DICpara = dict(DICpara)
    r = None
    if ImgSeqNum == -1:
        r = _preset(DICpara, "StereoSearchDistance")
        if r is None:
            r = _preset(DICpara, "NewFFTSearchDistance")
    elif ImgSeqNum >= 2:
        r = _preset(DICpara, "TemporalSearchDistance")
        if r is None:
            r = _preset(DICpara, "NewFFTSearchDistance")
        if r is not None and "fixSearchDistanceOrNot" not in DICpara:
            DICpara["fixSearchDistanceOrNot"] = 0
    if r is None:
        raise ValueError("search distance not set (the reference would prompt with input())")
    r = np.asarray(r).reshape(-1)
    if r.size == 1:
        r = np.array([r[0], r[0]])
    roi = DICpara["gridxyROIRange"]
    x0, y0, u, v, cc = funIntegerSearch(ImgRef, ImgDef, r, roi["gridx"], roi["gridy"],
                                        DICpara["winsize"], DICpara["winstepsize"], 0, ctx=ctx)
    mask = DICpara.get("ImgRefMask")
    size = DICpara.get("ImgSize")
    if mask is not None and size is not None:
        mask = np.asarray(mask)
        xi = np.round(x0).astype(np.int64)
        yi = np.round(y0).astype(np.int64)
        M, N = int(size[0]), int(size[1])
        # sub2ind(ImgSize, ...) throws outside ImgSize, and so does an index
        # beyond the mask: the catch leaves u, v as they are
        if (xi.size == 0 or (xi.min() >= 1 and yi.min() >= 1 and xi.max() <= M and
                             yi.max() <= N and M * N <= mask.size)):
            lin = (xi - 1) + (yi - 1) * M
            hole = mask.reshape(-1, order="F")[lin] == 0
            u = np.where(hole, np.nan, u)
            v = np.where(hole, np.nan, v)
    cc["ccThreshold"] = 1.25
    u, v, cc, _ = funRemoveOutliers(u, v, cc, 0, 100)
    DICpara["InitFFTSearchMethod"] = 1
    DICpara["SizeOfFFTSearchRegion"] = r
    return DICpara, x0, y0, u, v, cc


# ---------------------------------------------------------------------------
# device-resident entry points (pointers are device addresses as ints, e.g.
# torch tensor.data_ptr(); torch is only the allocator here)
# ---------------------------------------------------------------------------
class Plan:
    """Size-class node lists + hand-on scratch for repeated device solves."""

    def __init__(self, winsize, n, rect=False, ctx: Context | None = None):
        self.ctx = ctx or default_context()
        ws = np.asarray(winsize, dtype=np.int32)
        per = 1 if ws.size > (2 if rect else 1) else 0
        ws = _F(ws.reshape(-1, order="F"), np.int32)
        p = C.c_void_p()
        check(lib().aldic_plan_create(self.ctx.ptr, ws, int(n), per, 1 if rect else 0,
                                      C.byref(p)))
        self.ptr = p
        self.n = int(n)

    def set_variant(self, solver: int, cls: int, variant: int):
        """Fix the launch shape of one size class (solver 0 IC-GN, 1 Subpb1)."""
        check(lib().aldic_plan_set_variant(self.ptr, int(solver), int(cls), int(variant)))

    def set_stopping(self, rule: int, tol_dp: float = 0.0):
        """IC-GN stopping rule of this plan's solves: 0 the reference's
        relative ||b|| rule, 1 the step-size rule ||dP|| < tol_dp px, 2 its
        a-posteriori form (also converged once the contraction observed
        over the last two updates predicts a distance to the fixed point
        below tol_dp px)."""
        check(lib().aldic_plan_set_stopping(self.ptr, int(rule), float(tol_dp)))

    def set_persistent(self, on: bool = True):
        """Run the four fast size classes as one persistent launch fed by a
        size-sorted work queue (aldic_plan_set_persistent)."""
        check(lib().aldic_plan_set_persistent(self.ptr, 1 if on else 0))
        return self

    def close(self):
        if getattr(self, "ptr", None):
            lib().aldic_plan_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def local_icgn_dev(images: Images, plan: Plan, d_coords, d_U0, d_winsize, tol, d_U, d_F=None,
                   d_H21=None, d_conv=None, d_hole=None, d_zncc=None, ctx: Context | None = None):
    c = ctx or images.ctx
    check(lib().aldic_local_icgn_dev(c.ptr, images.ptr, plan.ptr, d_coords, d_U0, d_winsize,
                                     float(tol), d_U, d_F, d_H21, d_conv, d_hole, d_zncc))


def subpb1_dev(images: Images, plan: Plan, d_coords, d_UOld, d_FOld, d_vdual, d_winsize_xy, mu,
               tol, d_U, d_conv):
    check(lib().aldic_subpb1_dev(images.ctx.ptr, images.ptr, plan.ptr, d_coords, d_UOld, d_FOld,
                                 d_vdual, d_winsize_xy, float(mu), float(tol), d_U, d_conv))


def alm_iteration_dev(images: Images, plan: Plan, fem: "FEMesh", d_coords, d_winsize_xy, mu, beta,
                      tol, tol2, check_update, first, d_F1, d_hole, d_U1, d_U2, d_vdual, d_update,
                      d_state, alpha=0.0, winstepsize=1.0, cg_max_iter=64,
                      ctx: Context | None = None):
    """One augmented-Lagrangian iteration on the device
    (aldic_alm_iteration_dev, TemporalMatch_quadtree_ST1.m:337-387): d_U1,
    d_U2, d_vdual (FP64, 2n) are updated in place; d_update receives
    (UpdateY, UpdateY2), d_state (int32[2]) the done flag and the number of
    bad Subpb1 nodes, which take the global estimate instead of the host's
    natural-neighbour fill."""
    c = ctx or images.ctx
    check(lib().aldic_alm_iteration_dev(c.ptr, images.ptr, plan.ptr, fem.ptr, d_coords,
                                        d_winsize_xy, float(mu), float(beta), float(alpha),
                                        float(winstepsize), float(tol), float(tol2),
                                        1 if check_update else 0, 1 if first else 0, d_F1, d_hole,
                                        d_U1, d_U2, d_vdual, d_update, d_state, int(cg_max_iter)))


def integer_search_dev(ctx: Context, d_f, d_g, M, N, gridx, gridy, winsize, winstepsize, region,
                       d_x, d_y, d_u, d_v, d_ccmax, d_ipr, d_ipc):
    gx = (C.c_int64 * 2)(int(gridx[0]), int(gridx[-1]))
    gy = (C.c_int64 * 2)(int(gridy[0]), int(gridy[-1]))
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    check(lib().aldic_integer_search_dev(ctx.ptr, d_f, d_g, int(M), int(N), gx, gy, int(winsize),
                                         int(winstepsize), rg, d_x, d_y, d_u, d_v, d_ccmax,
                                         d_ipr, d_ipc))


def integer_search_noalloc_dev(ctx: Context, d_f, d_g, M, N, gridx, gridy, winsize,
                               winstepsize, region, d_x, d_y, d_u, d_v, d_ccmax, d_ipr, d_ipc):
    """integer_search_dev with its scratch in the library's fixed per-GPU
    array (no device allocation; calls that overlap on one GPU race)."""
    gx = (C.c_int64 * 2)(int(gridx[0]), int(gridx[-1]))
    gy = (C.c_int64 * 2)(int(gridy[0]), int(gridy[-1]))
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    check(lib().aldic_integer_search_noalloc_dev(ctx.ptr, d_f, d_g, int(M), int(N), gx, gy,
                                                 int(winsize), int(winstepsize), rg, d_x, d_y,
                                                 d_u, d_v, d_ccmax, d_ipr, d_ipc))


def bspline_prefilter_dev(ctx: Context, d_in, d_out, M, N):
    """Interpolating cubic B-spline coefficients of an M x N FP32 device
    plane into d_out (aldic_bspline_prefilter_dev; stream-ordered on ctx; no
    reference counterpart)."""
    check(lib().aldic_bspline_prefilter_dev(ctx.ptr, d_in, d_out, int(M), int(N)))


def field_to_nodes_dev(ctx: Context, d_Ug, d_Fg, nX, nY, x0, y0, step, d_coords, n, d_U0,
                       d_F0=None):
    """A warp field on a regular grid (point ci + cj*nX) -> its bilinear
    interpolant at the nodes (aldic_field_to_nodes_dev)."""
    check(lib().aldic_field_to_nodes_dev(ctx.ptr, d_Ug, d_Fg, int(nX), int(nY), float(x0),
                                         float(y0), float(step), d_coords, int(n), d_U0, d_F0))


def grid_to_nodes_dev(ctx: Context, d_u, d_v, nX, nY, x0, y0, step, d_coords, n, d_U0):
    check(lib().aldic_grid_to_nodes_dev(ctx.ptr, d_u, d_v, int(nX), int(nY), float(x0), float(y0),
                                        float(step), d_coords, int(n), d_U0))


def local_icgn_f0_dev(images: Images, plan: Plan, d_coords, d_U0, d_F0, d_winsize, tol, d_U,
                      d_F=None, d_H21=None, d_conv=None, d_hole=None, d_zncc=None, max_iter=0,
                      ctx: Context | None = None):
    """IC-GN from P0 = [F0 U0] (d_F0 None: zeros, as local_icgn_dev);
    max_iter caps k (0: the reference's 100); runs on ctx (default: the
    images' context)."""
    c = ctx or images.ctx
    check(lib().aldic_local_icgn_f0_dev(c.ptr, images.ptr, plan.ptr, d_coords, d_U0,
                                        d_F0, d_winsize, float(tol), int(max_iter), d_U, d_F,
                                        d_H21, d_conv, d_hole, d_zncc))


def gate_dev(ctx: Context, n, d_conv, d_zncc, max_iter, min_zncc, d_list, d_count):
    """List the nodes whose IC-GN took more than max_iter evaluations, did
    not converge, or ended below min_zncc (count on the device)."""
    check(lib().aldic_gate_dev(ctx.ptr, int(n), d_conv, d_zncc, int(max_iter), float(min_zncc),
                               d_list, d_count))


def search_nodes_dev(images: Images, d_coords, n, d_list, d_count, max_count, winsize, region,
                     d_U0, d_F0=None, ctx: Context | None = None):
    """Integer search centred on each listed node -> U0 (and F0 = 0); runs
    on ctx (default: the images' context)."""
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    c = ctx or images.ctx
    check(lib().aldic_search_nodes_dev(c.ptr, images.ptr, d_coords, int(n), d_list, d_count,
                                       int(max_count), int(winsize), rg, d_U0, d_F0))


def search_grid_nodes_dev(ctx: Context, d_f, d_g, M, N, gridx, gridy, winsize, winstepsize,
                          region, d_coords, n, d_list, d_count, d_U0, d_F0=None):
    """Search-init U0 of the listed nodes: only the grid points their
    bilinear transfer reads are searched."""
    gx = (C.c_int64 * 2)(int(gridx[0]), int(gridx[-1]))
    gy = (C.c_int64 * 2)(int(gridy[0]), int(gridy[-1]))
    rg = (C.c_int64 * 4)(*[int(v) for v in region])
    check(lib().aldic_search_grid_nodes_dev(ctx.ptr, d_f, d_g, int(M), int(N), gx, gy,
                                            int(winsize), int(winstepsize), rg, d_coords, int(n),
                                            d_list, d_count, d_U0, d_F0))


def local_icgn_list_dev(images: Images, plan: Plan, d_coords, d_U0, d_F0, d_winsize, d_list,
                        d_count, tol, d_U, d_F=None, d_H21=None, d_conv=None, d_hole=None,
                        d_zncc=None, ctx: Context | None = None):
    """IC-GN of the listed nodes of a plan only (outputs of the others
    untouched); runs on ctx (default: the images' context)."""
    c = ctx or images.ctx
    check(lib().aldic_local_icgn_list_dev(c.ptr, images.ptr, plan.ptr, d_coords, d_U0, d_F0,
                                          d_winsize, d_list, d_count, float(tol), d_U, d_F, d_H21,
                                          d_conv, d_hole, d_zncc))


STAGES = {0: "search", 1: "grid_to_nodes", 2: "icgn_c0", 3: "icgn_c1", 4: "icgn_c2",
          5: "icgn_c3", 6: "icgn_generic", 7: "icgn_handed_on", 8: "subpb1_c0", 9: "subpb1_c1",
          10: "subpb1_c2", 11: "subpb1_c3", 12: "subpb1_generic", 13: "subpb1_handed_on",
          14: "cnn", 15: "fallback", 16: "subpb2"}
NSTAGES = 17


def set_timing(ctx: Context, enable: bool):
    check(lib().aldic_ctx_set_timing(ctx.ptr, 1 if enable else 0))


def stage_times(ctx: Context, reset=True):
    """{stage name: (total ms, launches)} since the last reset."""
    ms = np.zeros(NSTAGES)
    cnt = np.zeros(NSTAGES, dtype=np.int32)
    check(lib().aldic_ctx_stage_times(ctx.ptr, ms.ctypes.data, cnt.ctypes.data, 1 if reset else 0))
    return {STAGES.get(k, f"stage{k}"): (float(ms[k]), int(cnt[k])) for k in range(NSTAGES)
            if cnt[k] > 0}


class DeviceArray:
    """A flat device buffer owned through the C ABI (aldic_dev_malloc)."""

    def __init__(self, ctx: Context, count, dtype):
        self.ctx = ctx
        self.dtype = np.dtype(dtype)
        self.count = int(count)
        p = C.c_void_p()
        check(lib().aldic_dev_malloc(ctx.ptr, self.count * self.dtype.itemsize, C.byref(p)))
        self.ptr = p.value

    @classmethod
    def from_host(cls, ctx: Context, a, dtype=None):
        a = np.ascontiguousarray(np.asarray(a, dtype=dtype).reshape(-1, order="F"))
        d = cls(ctx, a.size, a.dtype)
        check(lib().aldic_memcpy_htod(ctx.ptr, d.ptr, a.ctypes.data, a.nbytes))
        return d

    def upload(self, a):
        """Copy host data (count elements, column-major flattened) in."""
        a = np.ascontiguousarray(np.asarray(a, dtype=self.dtype).reshape(-1, order="F"))
        if a.size != self.count:
            raise ValueError("upload: size differs from the device array")
        check(lib().aldic_memcpy_htod(self.ctx.ptr, self.ptr, a.ctypes.data, a.nbytes))

    def to_host(self):
        out = np.empty(self.count, dtype=self.dtype)
        check(lib().aldic_memcpy_dtoh(self.ctx.ptr, out.ctypes.data, self.ptr, out.nbytes))
        return out

    def free(self):
        if self.ptr:
            lib().aldic_dev_free(self.ctx.ptr, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def local_icgn_autotune_dev(images: Images, plan: Plan, d_coords, d_U0, d_winsize, tol, d_U, d_F,
                            d_conv, d_zncc, reps=3):
    """Pick the fastest launch shape per size class on these nodes; returns
    the chosen shape index per class."""
    chosen = np.zeros(4, dtype=np.int32)
    check(lib().aldic_local_icgn_autotune_dev(images.ctx.ptr, images.ptr, plan.ptr, d_coords,
                                              d_U0, d_winsize, float(tol), d_U, d_F, d_conv,
                                              d_zncc, int(reps), chosen.ctypes.data))
    return chosen
