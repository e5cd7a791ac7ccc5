"""FP64 numpy oracle of the learned initial-guess network.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing on the product
path).  The network has no counterpart in the reference; this file states
its mathematics independently of the HIP kernels so their forward pass and
gradients can be checked.

Layout (shared with the library's aldic_cnn_layout): one FP32 vector; per
layer a weight matrix W[cout][Kp] (K = 9*cin for a 3x3 convolution, column
(ky*3 + kx)*cin + ci, zero padding to Kp = K rounded up to 16) at w_off,
then the bias at b_off; offsets rounded up to 64 entries.  Activations are
NHWC; 3x3 convolutions use zero padding 1 and the stated stride.

``quant=True`` rounds to bfloat16 (round to nearest even) where the device
stores bfloat16: weights, the input, every post-ReLU activation, and the
output gradient wherever it feeds a matrix product (bias gradients use the
FP32 value).  Everything else is FP64.

Pinned by tests/test_cnn_oracle.py: the convolution against
scipy.signal.correlate (a separate formula of the same operation), and the
backward pass against central finite differences of the forward pass.
"""
from __future__ import annotations

import numpy as np

PATCH = 48
# (cin, cout, stride, input side); fully connected layers have side 0
LAYERS = [(2, 32, 2, 48), (32, 32, 2, 24), (32, 64, 2, 12), (64, 64, 1, 6),
          (64, 64, 2, 6), (64, 64, 1, 3), (576, 128, 0, 0), (128, 6, 0, 0)]
NCONV = 6


def _up(v, a):
    return (v + a - 1) // a * a


def layout():
    out, off = [], 0
    for cin, cout, stride, side in LAYERS:
        conv = side > 0
        K = 9 * cin if conv else cin
        Kp = _up(K, 16)
        w_off = off
        off = _up(off + cout * Kp, 64)
        b_off = off
        off = _up(off + cout, 64)
        out.append(dict(w_off=w_off, b_off=b_off, cout=cout, K=K, Kp=Kp, cin=cin,
                        stride=stride, side=side, conv=conv,
                        out_side=((side - 1) // stride + 1) if conv else 1))
    return out, off


def bf16(a):
    """Round to the nearest bfloat16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def unpack(params):
    lay, _ = layout()
    Ws, bs = [], []
    for L in lay:
        W = params[L["w_off"]:L["w_off"] + L["cout"] * L["Kp"]].reshape(L["cout"], L["Kp"])
        Ws.append(np.asarray(W[:, :L["K"]], dtype=np.float64))
        bs.append(np.asarray(params[L["b_off"]:L["b_off"] + L["cout"]], dtype=np.float64))
    return Ws, bs


def init_params(seed=0):
    """He-normal weights, zero biases (padding entries zero)."""
    rng = np.random.default_rng(seed)
    lay, n = layout()
    p = np.zeros(n, dtype=np.float32)
    for i, L in enumerate(lay):
        std = np.sqrt(2.0 / L["K"])
        if i == len(lay) - 1:
            std *= 0.1  # small initial outputs
        W = np.zeros((L["cout"], L["Kp"]), dtype=np.float32)
        W[:, :L["K"]] = rng.normal(0.0, std, size=(L["cout"], L["K"]))
        p[L["w_off"]:L["w_off"] + W.size] = W.reshape(-1)
    return p


def im2col(x, stride):
    """x (B, H, W, C) -> (B*Ho*Wo, 9*C) with column (ky*3 + kx)*C + ci."""
    B, H, W, C = x.shape
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    xp = np.zeros((B, H + 2, W + 2, C), dtype=x.dtype)
    xp[:, 1:H + 1, 1:W + 1] = x
    cols = np.empty((B, Ho, Wo, 9, C), dtype=x.dtype)
    for ky in range(3):
        for kx in range(3):
            cols[:, :, :, ky * 3 + kx] = xp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                                            kx:kx + stride * (Wo - 1) + 1:stride]
    return cols.reshape(B * Ho * Wo, 9 * C), Ho, Wo


def col2im(dcol, shape, stride):
    B, H, W, C = shape
    Ho, Wo = (H - 1) // stride + 1, (W - 1) // stride + 1
    d = dcol.reshape(B, Ho, Wo, 9, C)
    dxp = np.zeros((B, H + 2, W + 2, C))
    for ky in range(3):
        for kx in range(3):
            dxp[:, ky:ky + stride * (Ho - 1) + 1:stride,
                kx:kx + stride * (Wo - 1) + 1:stride] += d[:, :, :, ky * 3 + kx]
    return dxp[:, 1:H + 1, 1:W + 1]


def conv3x3(x, W, b, stride):
    col, Ho, Wo = im2col(x, stride)
    y = col @ W.T + b
    return y.reshape(x.shape[0], Ho, Wo, W.shape[0])


def forward(params, x, quant=False):
    """x (B, 48, 48, 2) -> (out (B, 6), cache)."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    Ws, bs = unpack(params)
    Wq = [q(W) for W in Ws]
    if quant:  # the device adds the FP32 master biases
        bs = [np.asarray(b, dtype=np.float32).astype(np.float64) for b in bs]
    a = q(x)
    acts = [a]
    for i in range(NCONV):
        a = q(np.maximum(conv3x3(a, Wq[i], bs[i], LAYERS[i][2]), 0.0))
        acts.append(a)
    B = x.shape[0]
    flat = a.reshape(B, -1)
    h = q(np.maximum(flat @ Wq[NCONV].T + bs[NCONV], 0.0))
    acts.append(h)
    out = h @ Wq[NCONV + 1].T + bs[NCONV + 1]
    return out, (acts, Wq)


def loss_and_grads(params, x, labels, lscale=None, quant=False):
    """Mean over the batch of sum_k s_k (out_k - label_k)^2 and its
    gradient in the parameter layout."""
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    lscale = np.ones(6) if lscale is None else np.asarray(lscale, dtype=np.float64)
    out, (acts, Wq) = forward(params, x, quant)
    B = x.shape[0]
    e = out - labels
    loss = float(np.sum(lscale * e * e) / B)
    dout = 2.0 * lscale * e / B
    lay, n = layout()
    g = np.zeros(n)

    def put(i, dW, db):
        L = lay[i]
        Wfull = np.zeros((L["cout"], L["Kp"]))
        Wfull[:, :L["K"]] = dW
        g[L["w_off"]:L["w_off"] + Wfull.size] = Wfull.reshape(-1)
        g[L["b_off"]:L["b_off"] + L["cout"]] = db

    # fc2
    h = acts[-1]
    put(NCONV + 1, q(dout).T @ h, dout.sum(0))
    dh = (q(dout) @ Wq[NCONV + 1]) * (h > 0)
    # fc1
    a6 = acts[NCONV].reshape(B, -1)
    put(NCONV, q(dh).T @ a6, dh.sum(0))
    da = ((q(dh) @ Wq[NCONV]) * (a6 > 0)).reshape(acts[NCONV].shape)
    # convolutions
    for i in range(NCONV - 1, -1, -1):
        xin = acts[i]
        stride = LAYERS[i][2]
        col, Ho, Wo = im2col(xin, stride)
        dY = da.reshape(-1, LAYERS[i][1])
        put(i, q(dY).T @ col, dY.sum(0))
        if i == 0:
            break
        dcol = q(dY) @ Wq[i]
        dx = col2im(dcol, xin.shape, stride)
        da = dx * (xin > 0)
    return loss, g, out
