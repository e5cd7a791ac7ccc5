"""Raw 8/16-bit images normalised on the device (funNormalizeImg.m:22-40)
against the host path: double(raw)' (loadStereoImages.m:79), numpy
(Img - mean_ROI) / std_ROI, then the FP64 upload.

Tolerances: mean and std 1e-13 relative (the device forms them from exact
integer sums, numpy by pairwise floating-point sums); FP64 normalised
pixels 1e-13 relative; the FP32 planes agree to one unit in the last place
(an FP64 difference in the last bits can move an FP32 rounding).
"""
import numpy as np
import pytest

from _cases import hole_mask

pytestmark = pytest.mark.gpu


def _raw(H, W, bits, seed):
    rng = np.random.default_rng(seed)
    top = 4095 if bits == 16 else 235
    base = rng.integers(20, top, size=(H, W))
    # some spatial structure so the gradients are not noise only
    yy, xx = np.mgrid[0:H, 0:W]
    img = (base * 0.5 + 0.5 * top * (0.5 + 0.5 * np.sin(xx / 7.0) * np.cos(yy / 5.0)))
    return img.astype(np.uint16 if bits == 16 else np.uint8)


def _host(raw, roi):
    img = raw.T.astype(np.float64)  # double(raw)'
    M, N = img.shape
    gx = [max(roi["gridx"][0], 1), min(roi["gridx"][1], M)]
    gy = [max(roi["gridy"][0], 1), min(roi["gridy"][1], N)]
    t = img[gx[0] - 1:gx[1], gy[0] - 1:gy[1]]
    m, s = t.mean(), t.std(ddof=1)
    return (img - m) / s, m, s


@pytest.mark.parametrize("bits", [16, 8])
def test_images_from_raw_match_host_path(pkg, bits):
    H, W = 180, 200
    rl, rr = _raw(H, W, bits, 1), _raw(H, W, bits, 2)
    roi = {"gridx": [-5, 170], "gridy": [12, 400]}  # clamped like the reference
    mask = hole_mask(W, H, (90.0, 80.0), 20.0)
    fl, ml, sl = _host(rl, roi)
    fr, mr, sr = _host(rr, roi)
    want = pkg.Images(fl, fr, mask).planes()
    imr = pkg.Images.from_raw(rl, rr, mask, roi)
    got = imr.planes()
    st = np.array(imr.stats)
    assert np.allclose(st, [ml, sl, mr, sr], rtol=1e-13, atol=0)
    for k in ("fm", "g", "mask", "fx", "fy"):
        a, b = got[k], want[k]
        tol = np.spacing(np.abs(b).max().astype(np.float32))
        assert np.max(np.abs(a - b)) <= tol, k
        assert np.mean(a != b) < 1e-3, k
    # the library's own normalised FP64 image
    ctx = pkg.default_context()
    from aldic_amd import api
    d_raw = api.DeviceArray.from_host(ctx, np.ascontiguousarray(rl).reshape(-1),
                                      np.uint16 if bits == 16 else np.uint8)
    d_out = api.DeviceArray(ctx, W * H, np.float64)
    gx = np.array(roi["gridx"], np.int64)
    gy = np.array(roi["gridy"], np.int64)
    ms = np.zeros(2)
    pkg.lib().aldic_normalize_raw_dev(ctx.ptr, d_raw.ptr, bits, W, H, gx.ctypes.data,
                                      gy.ctypes.data, d_out.ptr, ms.ctypes.data)
    out = d_out.to_host().reshape(H, W).T
    assert np.max(np.abs(out - fl)) <= 1e-13 * np.abs(fl).max()
    imr.close()


def test_raw_rejects_bad_input(pkg):
    a = np.zeros((20, 20), np.uint16)
    with pytest.raises(TypeError):
        pkg.Images.from_raw(a, a.astype(np.uint8))
    with pytest.raises(pkg.AldicError):
        pkg.Images.from_raw(a, a, None, {"gridx": [30, 40], "gridy": [1, 20]})  # empty ROI
