"""MI355X (gfx950) implementation of ALDIC's local matching path.

The compute lives in libaldic.so (HIP kernels behind the C ABI of
include/aldic.h); this package is the host-side mirror of the reference's
MATLAB interface for that path (see :mod:`.api`).  Importing it loads the
library and fails loudly if it has not been built.
"""
from . import _capi
from ._capi import AldicError, lib

lib()  # no library, no package: there is no CPU fallback

from .api import (  # noqa: E402
    Context,
    FEMesh,
    Images,
    IntegerSearchQuadtree,
    LocalICGNQuadtree,
    Subpb1Quadtree,
    Subpb2Quadtree,
    ba_interp2_spline,
    default_context,
    funImgGradient,
    funIntegerSearch,
    funNormalizeImg,
    icgn_bad_points,
    local_icgn_raw,
    nan_fill,
    funRemoveOutliers,
    inpaint_nans,
    scattered_natural,
    subpb1_bad_points,
    subpb1_raw,
)

__all__ = [
    "AldicError", "Context", "FEMesh", "Images", "IntegerSearchQuadtree",
    "LocalICGNQuadtree", "Subpb1Quadtree", "Subpb2Quadtree", "ba_interp2_spline", "default_context", "funImgGradient",
    "funIntegerSearch", "funNormalizeImg", "icgn_bad_points", "local_icgn_raw",
    "nan_fill", "funRemoveOutliers", "inpaint_nans", "scattered_natural", "subpb1_bad_points", "subpb1_raw", "lib",
]
