"""CPU checks of the network oracle (oracle/cnn_oracle.py) and of the
library's fixed parameter layout.

The convolution is pinned against scipy.signal.correlate and the backward
pass against central finite differences of the FP64 forward pass; the
layout the oracle assumes must equal the one the library reports."""
import ctypes as C
import os
import sys

import numpy as np
import pytest
from scipy import signal

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import cnn_oracle as co  # noqa: E402


def test_layout_matches_library(pkg):
    lib = pkg.lib()
    n = C.c_int64()
    assert lib.aldic_cnn_num_params(None, C.byref(n)) == 0
    lay, total = co.layout()
    assert n.value == total
    info = (C.c_int64 * 9)()
    for i, L in enumerate(lay):
        assert lib.aldic_cnn_layout(None, i, info) == 0
        assert list(info) == [L["w_off"], L["b_off"], L["cout"], L["K"], L["Kp"], L["cin"],
                              L["stride"], L["side"] if L["conv"] else 1, int(L["conv"])]
    sf, su = C.c_double(), C.c_double()
    lib.aldic_cnn_scales(C.byref(sf), C.byref(su))
    assert sf.value == pytest.approx(0.05) and su.value == pytest.approx(4.0)


@pytest.mark.parametrize("stride", [1, 2])
def test_conv_matches_scipy_correlate(stride):
    rng = np.random.default_rng(0)
    B, H, Cin, Cout = 2, 7, 3, 4
    x = rng.normal(size=(B, H, H, Cin))
    W = rng.normal(size=(Cout, 9 * Cin))
    b = rng.normal(size=Cout)
    y = co.conv3x3(x, W, b, stride)
    Wk = W.reshape(Cout, 3, 3, Cin)  # column (ky*3 + kx)*Cin + ci
    for n in range(B):
        for co_ in range(Cout):
            ref = np.full((H, H), b[co_])
            for ci in range(Cin):
                ref += signal.correlate(x[n, :, :, ci], Wk[co_, :, :, ci], mode="same")
            np.testing.assert_allclose(y[n, :, :, co_], ref[::stride, ::stride], atol=1e-12)


def test_col2im_is_adjoint_of_im2col():
    rng = np.random.default_rng(1)
    for stride in (1, 2):
        x = rng.normal(size=(2, 6, 6, 3))
        col, _, _ = co.im2col(x, stride)
        d = rng.normal(size=col.shape)
        lhs = np.sum(col * d)
        rhs = np.sum(x * co.col2im(d, x.shape, stride))
        assert lhs == pytest.approx(rhs, rel=1e-12)


def test_bf16_rounding():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 1e-3], dtype=np.float64)
    r = co.bf16(v)
    # 1 + 2^-8 is a tie between 1 and 1 + 2^-7 -> even (1); 1 + 3*2^-8 -> 1 + 2^-6
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0 + 2 ** -6 and r[3] == -2.5
    assert abs(r[4] - 1e-3) < 1e-3 * 2 ** -8


def test_gradients_match_finite_differences():
    rng = np.random.default_rng(2)
    p = co.init_params(3).astype(np.float64)
    lay, n = co.layout()
    # make the last layer large enough that the loss is not flat
    L = lay[-1]
    p[L["w_off"]:L["w_off"] + L["cout"] * L["Kp"]] *= 10.0
    for L in lay:
        p[L["b_off"]:L["b_off"] + L["cout"]] = rng.normal(0, 0.05, L["cout"])
    x = rng.normal(size=(2, 48, 48, 2))
    lab = rng.normal(size=(2, 6))
    loss, g, _ = co.loss_and_grads(p, x, lab)
    picks = []
    for L in lay:
        W = rng.integers(0, L["cout"], 3)
        K = rng.integers(0, L["K"], 3)
        picks += [L["w_off"] + w * L["Kp"] + k for w, k in zip(W, K)]
        picks.append(L["b_off"] + int(rng.integers(0, L["cout"])))
    errs = []
    for i in picks:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (co.loss_and_grads(pp, x, lab)[0] - co.loss_and_grads(pm, x, lab)[0]) / (2 * h)
        errs.append(abs(fd - g[i]) / max(1e-6, abs(fd), abs(g[i])))
    errs = np.array(errs)
    # a ReLU kink inside +-h can spoil one difference; the rest agree to 1e-5
    assert np.sum(errs > 1e-5) <= 1, errs
    assert np.median(errs) < 1e-7


def test_padding_gradients_are_zero():
    rng = np.random.default_rng(4)
    p = co.init_params(5).astype(np.float64)
    x = rng.normal(size=(2, 48, 48, 2))
    _, g, _ = co.loss_and_grads(p, x, rng.normal(size=(2, 6)))
    lay, n = co.layout()
    used = np.zeros(n, dtype=bool)
    for L in lay:
        for r in range(L["cout"]):
            used[L["w_off"] + r * L["Kp"]:L["w_off"] + r * L["Kp"] + L["K"]] = True
        used[L["b_off"]:L["b_off"] + L["cout"]] = True
    assert np.all(g[~used] == 0.0)
