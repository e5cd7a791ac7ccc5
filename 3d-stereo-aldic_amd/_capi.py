"""ctypes bindings of libaldic.so (include/aldic.h).

The library must exist: there is no fallback path.  Import raises
:class:`ImportError` naming the build command when it is missing.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libaldic.so")

OK, EINVAL, EDEVICE, ENOMEM = 0, 1, 2, 3

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="F_CONTIGUOUS")
_fp = np.ctypeslib.ndpointer(dtype=np.float32, flags="F_CONTIGUOUS")
_i32 = np.ctypeslib.ndpointer(dtype=np.int32, flags="F_CONTIGUOUS")
_u8 = np.ctypeslib.ndpointer(dtype=np.uint8, flags="F_CONTIGUOUS")
_i64 = C.c_int64
_vp = C.c_void_p
_pp = C.POINTER(C.c_void_p)

class AldicParams(C.Structure):
    """aldic_params: the DICpara fields the local matching path reads."""
    _fields_ = [
        ("winsize", C.c_int32), ("winstepsize", C.c_int32), ("winsizeMin", C.c_int32),
        ("gridx", C.c_int64 * 2), ("gridy", C.c_int64 * 2),
        ("StereoSearchDistance", C.c_int32 * 4), ("nStereoSearchDistance", C.c_int32),
        ("TemporalSearchDistance", C.c_int32 * 4), ("nTemporalSearchDistance", C.c_int32),
        ("NewFFTSearchDistance", C.c_int32 * 4), ("nNewFFTSearchDistance", C.c_int32),
        ("fixSearchDistanceOrNot", C.c_int32), ("InitFFTSearchMethod", C.c_int32),
        ("DICIncOrNot", C.c_int32), ("UseGlobal", C.c_int32), ("GaussPtOrder", C.c_int32),
        ("ClusterNo", C.c_int32),
        ("tol_stereo", C.c_double), ("tol_temporal", C.c_double),
        ("mu", C.c_double), ("beta", C.c_double), ("alpha", C.c_double), ("tol2", C.c_double),
    ]


_prm = C.POINTER(AldicParams)

# Every symbol of include/aldic.h with (restype, argtypes).
SIGNATURES = {
    "aldic_last_error": (C.c_char_p, []),
    "aldic_ctx_create": (C.c_int, [C.c_int, _pp]),
    "aldic_ctx_destroy": (C.c_int, [_vp]),
    "aldic_ctx_stream": (_vp, [_vp]),
    "aldic_ctx_synchronize": (C.c_int, [_vp]),
    "aldic_dev_malloc": (C.c_int, [_vp, _i64, _pp]),
    "aldic_dev_free": (C.c_int, [_vp, _vp]),
    "aldic_memcpy_htod": (C.c_int, [_vp, _vp, _vp, _i64]),
    "aldic_memcpy_dtoh": (C.c_int, [_vp, _vp, _vp, _i64]),
    "aldic_ctx_set_timing": (C.c_int, [_vp, C.c_int32]),
    "aldic_ctx_set_search_method": (C.c_int, [_vp, C.c_int32]),
    "aldic_plan_set_persistent": (C.c_int, [_vp, C.c_int32]),
    "aldic_field_to_nodes_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.c_double, C.c_double,
                                           C.c_double, _vp, _i64, _vp, _vp]),
    "aldic_params_default": (C.c_int, [_prm]),
    "aldic_params_validate": (C.c_int, [_prm]),
    "aldic_params_search_region": (C.c_int, [_prm, C.c_int32, C.POINTER(_i64)]),
    "aldic_ctx_record": (C.c_int, [_vp, C.c_int32]),
    "aldic_ctx_capture_begin": (C.c_int, [_vp]),
    "aldic_ctx_capture_end": (C.c_int, [_vp, _pp]),
    "aldic_graph_launch": (C.c_int, [_vp, _vp]),
    "aldic_graph_destroy": (C.c_int, [_vp]),
    "aldic_ctx_wait": (C.c_int, [_vp, _vp, C.c_int32]),
    "aldic_ctx_stage_times": (C.c_int, [_vp, _vp, _vp, C.c_int32]),
    "aldic_grid_to_nodes_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.c_double, C.c_double,
                                          C.c_double, _vp, _i64, _vp]),
    "aldic_interp_spline_f64": (C.c_int, [_vp, _dp, _i64, _i64, _i64, _dp, _dp, _i64, _dp]),
    "aldic_interp_spline_f32": (C.c_int, [_vp, _fp, _i64, _i64, _i64, _fp, _fp, _i64, _fp]),
    "aldic_img_gradient": (C.c_int, [_vp, _dp, _vp, _i64, _i64, _dp, _dp]),
    "aldic_fem_create": (C.c_int, [_vp, _dp, _i64, _vp, _i64, _vp, _i64, _vp, _i64, C.c_int32,
                                   _pp]),
    "aldic_fem_destroy": (C.c_int, [_vp]),
    "aldic_fem_info": (C.c_int, [_vp, _vp, _vp, _vp]),
    "aldic_subpb2": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, _dp, _dp,
                               _dp, _vp]),
    "aldic_subpb2_dev": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double, _vp,
                                   _vp, _vp, C.c_int32, _vp]),
    "aldic_subpb2_dev_async": (C.c_int, [_vp, _vp, C.c_double, C.c_double, C.c_double, C.c_double,
                                         _vp, _vp, _vp, C.c_int32, _vp]),
    "aldic_images_planes": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "aldic_images_set_interp": (C.c_int, [_vp, C.c_int32]),
    "aldic_bspline_prefilter_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64]),
    "aldic_images_create_raw": (C.c_int, [_vp, _vp, _vp, C.c_int32, _vp, _i64, _i64, _vp, _vp,
                                          _pp, _vp]),
    "aldic_normalize_raw_dev": (C.c_int, [_vp, _vp, C.c_int32, _i64, _i64, _vp, _vp, _vp, _vp]),
    "aldic_natural_interp": (C.c_int, [_dp, _dp, _dp, _i64, C.c_int32, _dp, _dp, _i64, _dp]),
    "aldic_nan_fill": (C.c_int, [_dp, _i64, _dp, _vp, _vp, _i64, _i64, _vp]),
    "aldic_images_create": (C.c_int, [_vp, _dp, _dp, _vp, _vp, _vp, _i64, _i64, _pp]),
    "aldic_images_create_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _i64, _i64, _pp]),
    "aldic_images_destroy": (C.c_int, [_vp]),
    "aldic_local_icgn": (C.c_int, [_vp, _vp, _dp, _dp, _i64, _i32, C.c_int32, C.c_double,
                                   _dp, _dp, _vp, _i32, _vp, _vp]),
    "aldic_subpb1": (C.c_int, [_vp, _vp, _dp, _dp, _dp, _dp, _dp, _i64, _i32, C.c_double,
                               C.c_double, _dp, _i32]),
    "aldic_plan_create": (C.c_int, [_vp, _i32, _i64, C.c_int32, C.c_int32, _pp]),
    "aldic_plan_destroy": (C.c_int, [_vp]),
    "aldic_local_icgn_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp, _vp, _vp,
                                       _vp, _vp, _vp]),
    "aldic_subpb1_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                   C.c_double, _vp, _vp]),
    "aldic_local_icgn_autotune_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, _vp,
                                                _vp, _vp, _vp, C.c_int32, _vp]),
    "aldic_plan_set_variant": (C.c_int, [_vp, C.c_int32, C.c_int32, C.c_int32]),
    "aldic_plan_set_stopping": (C.c_int, [_vp, C.c_int32, C.c_double]),
    "aldic_alm_iteration_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, C.c_double, C.c_double,
                                          C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_int32, C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp,
                                          _vp, C.c_int32]),
    "aldic_integer_search": (C.c_int, [_vp, _dp, _dp, _i64, _i64, C.POINTER(_i64),
                                       C.POINTER(_i64), C.c_int32, C.c_int32, C.POINTER(_i64),
                                       _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, C.POINTER(_i64),
                                       C.POINTER(_i64)]),
    "aldic_integer_search_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.POINTER(_i64),
                                           C.POINTER(_i64), C.c_int32, C.c_int32,
                                           C.POINTER(_i64), _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "aldic_integer_search_noalloc_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.POINTER(_i64),
                                                   C.POINTER(_i64), C.c_int32, C.c_int32,
                                                   C.POINTER(_i64), _vp, _vp, _vp, _vp, _vp, _vp,
                                                   _vp]),
    "aldic_local_icgn_f0_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, C.c_double,
                                          C.c_int32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "aldic_cnn_create": (C.c_int, [_vp, C.c_int32, _pp]),
    "aldic_cnn_destroy": (C.c_int, [_vp]),
    "aldic_cnn_num_params": (C.c_int, [_vp, C.POINTER(_i64)]),
    "aldic_cnn_layout": (C.c_int, [_vp, C.c_int32, _vp]),
    "aldic_cnn_scales": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "aldic_cnn_set_params": (C.c_int, [_vp, _vp, _i64]),
    "aldic_cnn_get_params": (C.c_int, [_vp, _vp, _i64]),
    "aldic_cnn_get_grads": (C.c_int, [_vp, _vp, _i64]),
    "aldic_cnn_set_loss_weights": (C.c_int, [_vp, _vp]),
    "aldic_cnn_forward_dev": (C.c_int, [_vp, _vp, _i64, _vp]),
    "aldic_cnn_train_step_dev": (C.c_int, [_vp, _vp, _vp, _i64, C.c_float, C.c_int32, _vp]),
    "aldic_cnn_patches_dev": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _vp]),
    "aldic_cnn_predict_dev": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _vp]),
    "aldic_speckle_mosaic_dev": (C.c_int, [_vp, _i64, _i64, C.c_int32, _vp, C.c_uint32,
                                           C.c_uint32, C.c_double, C.c_double, C.c_double,
                                           C.c_double, C.c_double, _vp]),
    "aldic_gate_dev": (C.c_int, [_vp, _i64, _vp, _vp, C.c_int32, C.c_double, _vp, _vp]),
    "aldic_search_nodes_dev": (C.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i64, C.c_int32,
                                         C.POINTER(_i64), _vp, _vp]),
    "aldic_search_grid_nodes_dev": (C.c_int, [_vp, _vp, _vp, _i64, _i64, C.POINTER(_i64),
                                              C.POINTER(_i64), C.c_int32, C.c_int32,
                                              C.POINTER(_i64), _vp, _i64, _vp, _vp, _vp, _vp]),
    "aldic_local_icgn_list_dev": (C.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                            C.c_double, _vp, _vp, _vp, _vp, _vp, _vp]),
}

_lib = None


class AldicError(RuntimeError):
    """A non-OK return of the C ABI (bad argument, HIP error, out of memory)."""

    def __init__(self, code: int, msg: str):
        super().__init__(f"[aldic error {code}] {msg}")
        self.code = code


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ImportError(f"{LIBPATH} is missing: build it with "
                              f"`python 3d-stereo-aldic_amd/build.py` (hipcc, gfx950)")
        L = C.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int):
    if rc != OK:
        msg = lib().aldic_last_error()
        raise AldicError(rc, msg.decode() if msg else "")
