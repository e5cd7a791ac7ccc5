"""Learned initial guess for IC-GN (no counterpart in the reference).

A small CNN maps a 48 x 48 reference/deformed patch pair to the six
first-order warp parameters [u_x v_x u_y v_y u v] that IC-GN starts from.
Everything runs in libaldic.so: the network's forward and backward passes
(BF16 matrix-core GEMMs, FP32 accumulation), Adam, the synthetic training
images and the gate that sends a subset back to the integer search.

Training data are IC-GN-converged parameters only: each training image is a
mosaic of square tiles of seeded speckle, each tile of the deformed image
under its own random affine displacement (translation within +-12 px,
gradients within +-0.05); IC-GN (41 px subsets) run from the true warp
gives the label, and only subsets that converge with ZNCC >= 0.9 are kept.

Inference (``LearnedInit``): a sparse integer search supplies a coarse
prior (the network's translation range is +-12 px around it), the network
predicts the warp relative to the rounded prior, IC-GN starts from that,
and every subset that takes more than 30 evaluations or ends with ZNCC
below 0.9 is searched again at its own position and rerun from there.
"""
from __future__ import annotations

import ctypes as C
import time

import numpy as np

from ._capi import check, lib
from .api import (Context, DeviceArray, Images, Plan, default_context, field_to_nodes_dev,
                  gate_dev, grid_to_nodes_dev, integer_search_dev, local_icgn_f0_dev,
                  local_icgn_list_dev, search_grid_nodes_dev)

PATCH = 48


def scales():
    sf, su = C.c_double(), C.c_double()
    check(lib().aldic_cnn_scales(C.byref(sf), C.byref(su)))
    return sf.value, su.value


def num_params():
    n = C.c_int64()
    check(lib().aldic_cnn_num_params(None, C.byref(n)))
    return n.value


def layout():
    """Per layer: dict(w_off, b_off, cout, K, Kp, cin, stride, side, conv)."""
    out = []
    info = (C.c_int64 * 9)()
    for i in range(8):
        check(lib().aldic_cnn_layout(None, i, info))
        v = list(info)
        out.append(dict(w_off=v[0], b_off=v[1], cout=v[2], K=v[3], Kp=v[4], cin=v[5],
                        stride=v[6], side=v[7], conv=bool(v[8])))
    return out


def he_init(seed=0):
    """He-normal weights (the last layer scaled by 0.1), zero biases."""
    rng = np.random.default_rng(seed)
    lay = layout()
    p = np.zeros(num_params(), dtype=np.float32)
    for i, L in enumerate(lay):
        std = np.sqrt(2.0 / L["K"]) * (0.1 if i == len(lay) - 1 else 1.0)
        W = np.zeros((L["cout"], L["Kp"]), dtype=np.float32)
        W[:, :L["K"]] = rng.normal(0.0, std, size=(L["cout"], L["K"]))
        p[L["w_off"]:L["w_off"] + W.size] = W.reshape(-1)
    return p


class Network:
    """The device-resident network (aldic_cnn)."""

    def __init__(self, ctx: Context | None = None, max_batch=256, params=None, seed=0):
        self.ctx = ctx or default_context()
        self.max_batch = int(max_batch)
        p = C.c_void_p()
        check(lib().aldic_cnn_create(self.ctx.ptr, self.max_batch, C.byref(p)))
        self.ptr = p
        self.n = num_params()
        self.set_params(he_init(seed) if params is None else params)

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        check(lib().aldic_cnn_set_params(self.ptr, a.ctypes.data, a.size))

    def get_params(self):
        a = np.empty(self.n, dtype=np.float32)
        check(lib().aldic_cnn_get_params(self.ptr, a.ctypes.data, a.size))
        return a

    def get_grads(self):
        a = np.empty(self.n, dtype=np.float32)
        check(lib().aldic_cnn_get_grads(self.ptr, a.ctypes.data, a.size))
        return a

    def set_loss_weights(self, w):
        a = np.ascontiguousarray(w, dtype=np.float32)
        check(lib().aldic_cnn_set_loss_weights(self.ptr, a.ctypes.data))

    def forward_dev(self, d_x, B, d_out=None):
        check(lib().aldic_cnn_forward_dev(self.ptr, d_x, int(B), d_out))

    def train_step_dev(self, d_x, d_labels, B, lr, update=True, d_loss=None):
        check(lib().aldic_cnn_train_step_dev(self.ptr, d_x, d_labels, int(B), float(lr),
                                             1 if update else 0, d_loss))

    def predict_dev(self, images: Images, d_coords, n, d_prior, d_U0, d_F0):
        check(lib().aldic_cnn_predict_dev(self.ptr, images.ptr, d_coords, int(n), d_prior, d_U0,
                                          d_F0))

    def save(self, path):
        np.savez(path, params=self.get_params())

    @classmethod
    def load(cls, path, ctx=None, max_batch=256):
        with np.load(path, allow_pickle=False) as z:
            return cls(ctx, max_batch, params=z["params"])

    def close(self):
        if getattr(self, "ptr", None):
            lib().aldic_cnn_destroy(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def patches_dev(images: Images, d_coords, n, d_prior, d_out):
    check(lib().aldic_cnn_patches_dev(images.ctx.ptr, images.ptr, d_coords, int(n), d_prior,
                                      d_out))


def speckle_mosaic_dev(ctx: Context, M, N, tile, d_warps, seed, noise_seed, d_img,
                       density=0.015, radius=3.0, bits=12, noise=2.0):
    scale = 16.0 if bits == 12 else 1.0
    check(lib().aldic_speckle_mosaic_dev(ctx.ptr, int(M), int(N), int(tile), d_warps,
                                         int(seed) & 0xFFFFFFFF, int(noise_seed) & 0xFFFFFFFF,
                                         float(density), float(radius), scale, float(noise),
                                         float(2 ** bits - 1), d_img))


# ---------------------------------------------------------------------------
# training data
# ---------------------------------------------------------------------------
class MosaicSource:
    """Batches of (patch pair, normalised IC-GN label) from speckle mosaics.

    Each mosaic is tiles x tiles tiles of ``tile`` px; one subset per tile,
    at the tile's centre node."""

    def __init__(self, ctx: Context, tiles=64, tile=96, seed=4, trans=12.0, grad=0.05,
                 label_ws=40, tol=1e-3):
        self.ctx = ctx
        self.tiles, self.tile = int(tiles), int(tile)
        self.M = self.N = self.tiles * self.tile
        self.seed = int(seed)
        self.trans, self.grad = float(trans), float(grad)
        self.label_ws = int(label_ws)
        self.tol = float(tol)
        self.sf, self.su = scales()
        n = self.tiles * self.tiles
        self.n = n
        tx, ty = np.meshgrid(np.arange(self.tiles), np.arange(self.tiles), indexing="ij")
        cx = (tx * self.tile + self.tile // 2).reshape(-1, order="F")
        cy = (ty * self.tile + self.tile // 2).reshape(-1, order="F")
        self.coords = np.concatenate([cx, cy]).astype(np.float64)  # n x 2 column-major
        self.d_ref = DeviceArray(ctx, self.M * self.N, np.float32)
        self.d_def = DeviceArray(ctx, self.M * self.N, np.float32)
        self.d_coords = DeviceArray.from_host(ctx, self.coords, np.float64)
        self.d_ws = DeviceArray.from_host(ctx, np.full(n, self.label_ws, np.int32))
        self.plan = Plan(self.label_ws, n, ctx=ctx)
        self.d_warps = DeviceArray(ctx, 6 * n, np.float32)
        self.d_U0 = DeviceArray(ctx, 2 * n, np.float32)
        self.d_F0 = DeviceArray(ctx, 4 * n, np.float32)
        self.d_U = DeviceArray(ctx, 2 * n, np.float32)
        self.d_F = DeviceArray(ctx, 4 * n, np.float32)
        self.d_conv = DeviceArray(ctx, n, np.int32)
        self.d_zncc = DeviceArray(ctx, n, np.float32)
        self.d_patches = DeviceArray(ctx, n * PATCH * PATCH, np.float32)  # bf16 pairs
        self.d_labels = DeviceArray(ctx, 6 * n, np.float32)
        self.d_kept = DeviceArray(ctx, 2 * n, np.float64)
        self.count = 0

    def warps(self, k):
        """Per-tile [u_x v_x u_y v_y u v] of mosaic k (tiles along dim 1 first)."""
        rng = np.random.default_rng([self.seed, k])
        w = np.empty((self.n, 6), dtype=np.float32)
        w[:, :4] = rng.uniform(-self.grad, self.grad, (self.n, 4))
        w[:, 4:] = rng.uniform(-self.trans, self.trans, (self.n, 2))
        return w

    def make(self, k):
        """Render mosaic k, label it with IC-GN, cut the patches; returns the
        number of kept samples and label statistics."""
        ctx = self.ctx
        w = self.warps(k)
        check(lib().aldic_memcpy_htod(ctx.ptr, self.d_warps.ptr, w.ctypes.data, w.nbytes))
        s0 = (self.seed * 1000003 + 2 * k) & 0xFFFFFFFF
        speckle_mosaic_dev(ctx, self.M, self.N, self.tile, None, s0, s0 + 11, self.d_ref.ptr)
        speckle_mosaic_dev(ctx, self.M, self.N, self.tile, self.d_warps.ptr, s0, s0 + 17,
                           self.d_def.ptr)
        im = Images.from_device(self.d_ref.ptr, self.d_def.ptr, self.M, self.N, ctx=ctx)
        # IC-GN from the true warp: the converged parameters are the labels
        U0 = np.ascontiguousarray(w[:, 4:]).reshape(-1)
        F0 = np.ascontiguousarray(w[:, :4]).reshape(-1)
        check(lib().aldic_memcpy_htod(ctx.ptr, self.d_U0.ptr, U0.ctypes.data, U0.nbytes))
        check(lib().aldic_memcpy_htod(ctx.ptr, self.d_F0.ptr, F0.ctypes.data, F0.nbytes))
        local_icgn_f0_dev(im, self.plan, self.d_coords.ptr, self.d_U0.ptr, self.d_F0.ptr,
                          self.d_ws.ptr, self.tol, self.d_U.ptr, self.d_F.ptr,
                          d_conv=self.d_conv.ptr, d_zncc=self.d_zncc.ptr)
        U = self.d_U.to_host().reshape(-1, 2)
        F = self.d_F.to_host().reshape(-1, 4)
        conv = self.d_conv.to_host()
        zn = self.d_zncc.to_host()
        keep = (conv >= 1) & (conv <= 100) & (zn >= 0.9) & np.all(np.isfinite(U), 1)
        idx = np.nonzero(keep)[0]
        lab = np.concatenate([F[idx] / self.sf, U[idx] / self.su], axis=1).astype(np.float32)
        kc = np.concatenate([self.coords[idx], self.coords[self.n + idx]])
        check(lib().aldic_memcpy_htod(ctx.ptr, self.d_kept.ptr, kc.ctypes.data, kc.nbytes))
        check(lib().aldic_memcpy_htod(ctx.ptr, self.d_labels.ptr, lab.ctypes.data, lab.nbytes))
        patches_dev(im, self.d_kept.ptr, len(idx), None, self.d_patches.ptr)
        ctx.synchronize()
        im.close()
        self.count = len(idx)
        err_u = np.sqrt(np.mean(np.sum((U[idx] - w[idx, 4:]) ** 2, 1))) if len(idx) else np.nan
        err_f = np.sqrt(np.mean((F[idx] - w[idx, :4]) ** 2)) if len(idx) else np.nan
        return dict(kept=len(idx), total=self.n, label_rms_u=float(err_u),
                    label_rms_f=float(err_f), truth_u=w[idx, 4:], truth_f=w[idx, :4],
                    labels=lab)

    def batch_ptrs(self, start, B):
        """Device pointers of samples start .. start+B-1 of the current mosaic."""
        pb = PATCH * PATCH * 2 * 2  # bytes per patch pair (bf16)
        return self.d_patches.ptr + start * pb, self.d_labels.ptr + start * 6 * 4

    def close(self):
        self.plan.close()


def train(net: Network, samples=2_000_000, batch=256, lr=2e-3, lr_min=5e-5, seed=4,
          tiles=64, log_every=50, log=None, time_limit=None, trans=12.0):
    """Adam over fresh mosaics until `samples` labelled pairs have been seen
    (cosine learning-rate decay).  trans: the tiles' translations are drawn
    within +-trans px (the range of the residual the network must predict
    around the rounded prior).  Returns a history dict."""
    ctx = net.ctx
    src = MosaicSource(ctx, tiles=tiles, seed=seed, trans=trans)
    d_loss = DeviceArray(ctx, 1, np.float32)
    hist = dict(step=[], loss=[], label_rms_u=[], kept=0, total=0)
    steps_total = max(1, samples // batch)
    step, seen, k = 0, 0, 0
    t0 = time.time()
    while seen < samples:
        st = src.make(k)
        k += 1
        hist["kept"] += st["kept"]
        hist["total"] += st["total"]
        hist["label_rms_u"].append(st["label_rms_u"])
        for s in range(0, st["kept"] - batch + 1, batch):
            frac = min(1.0, step / steps_total)
            rate = lr_min + 0.5 * (lr - lr_min) * (1 + np.cos(np.pi * frac))
            px, pl = src.batch_ptrs(s, batch)
            net.train_step_dev(px, pl, batch, rate, True, d_loss.ptr)
            step += 1
            seen += batch
            if step % log_every == 0:
                loss = float(d_loss.to_host()[0])
                hist["step"].append(step)
                hist["loss"].append(loss)
                if log:
                    log(f"step {step} seen {seen} loss {loss:.5f} lr {rate:.2e} "
                        f"mosaics {k} {time.time() - t0:.1f}s")
            if seen >= samples:
                break
        if time_limit and time.time() - t0 > time_limit:
            break
    ctx.synchronize()
    hist["seconds"] = time.time() - t0
    hist["steps"] = step
    hist["samples"] = seen
    hist["mosaics"] = k
    src.close()
    return hist


def to_bf16_bits(a):
    """float array -> bfloat16 bit patterns (round to nearest even), uint16."""
    u = np.ascontiguousarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def from_bf16_bits(b):
    return (np.asarray(b, dtype=np.uint32) << 16).view(np.float32)


def evaluate(net: Network, seed=99, tiles=32, trans=12.0):
    """Prediction error on a held-out mosaic (translations within +-trans
    px): RMS of (u, v) in px and of the gradients, against the IC-GN
    labels."""
    src = MosaicSource(net.ctx, tiles=tiles, seed=seed, trans=trans)
    st = src.make(0)
    n = st["kept"]
    out = DeviceArray(net.ctx, 6 * max(n, 1), np.float32)
    sf, su = scales()
    preds = []
    pb = PATCH * PATCH * 2 * 2
    for s in range(0, n, net.max_batch):
        B = min(net.max_batch, n - s)
        net.forward_dev(src.d_patches.ptr + s * pb, B, out.ptr + s * 24)
    y = out.to_host()[:6 * n].reshape(n, 6)
    lab = st["labels"]
    du = (y[:, 4:] - lab[:, 4:]) * su
    dF = (y[:, :4] - lab[:, :4]) * sf
    src.close()
    return dict(n=n, rms_u_px=float(np.sqrt(np.mean(np.sum(du ** 2, 1)))),
                rms_F=float(np.sqrt(np.mean(dF ** 2))),
                label_rms_u_px=st["label_rms_u"])


def icgn_gain(net: Network, seed=4321, tiles=24, ws=32, max_iter=30, tol=1e-3):
    """IC-GN on a held-out mosaic started three ways: from zero (the prior
    alone: U0 = 0, F0 = 0), from the network, and from the true warp.  For
    each: fraction of subsets that converge within max_iter evaluations
    with ZNCC >= 0.9 and |U - truth| < 0.1 px, and their mean evaluations."""
    ctx = net.ctx
    src = MosaicSource(ctx, tiles=tiles, seed=seed)
    src.make(0)
    w = src.warps(0)
    n = src.n
    im = Images.from_device(src.d_ref.ptr, src.d_def.ptr, src.M, src.N, ctx=ctx)
    d_ws = DeviceArray.from_host(ctx, np.full(n, ws, np.int32), np.int32)
    plan = Plan(ws, n, ctx=ctx)
    res = {}
    for mode in ("prior", "learned", "truth"):
        if mode == "truth":
            U0 = np.ascontiguousarray(w[:, 4:]).reshape(-1)
            F0 = np.ascontiguousarray(w[:, :4]).reshape(-1)
        else:
            U0 = np.zeros(2 * n, np.float32)
            F0 = np.zeros(4 * n, np.float32)
        dU0 = DeviceArray.from_host(ctx, U0, np.float32)
        dF0 = DeviceArray.from_host(ctx, F0, np.float32)
        if mode == "learned":
            net.predict_dev(im, src.d_coords.ptr, n, None, dU0.ptr, dF0.ptr)
        dU = DeviceArray(ctx, 2 * n, np.float32)
        dF = DeviceArray(ctx, 4 * n, np.float32)
        dcv = DeviceArray(ctx, n, np.int32)
        dz = DeviceArray(ctx, n, np.float32)
        local_icgn_f0_dev(im, plan, src.d_coords.ptr, dU0.ptr, dF0.ptr, d_ws.ptr, tol, dU.ptr,
                          dF.ptr, d_conv=dcv.ptr, d_zncc=dz.ptr)
        U = dU.to_host().reshape(-1, 2)
        conv = dcv.to_host()
        z = dz.to_host()
        err = np.sqrt(np.sum((U - w[:, 4:]) ** 2, 1))
        good = (conv >= 1) & (conv <= max_iter) & (z >= 0.9) & (err < 0.1)
        res[mode] = dict(good=round(float(np.mean(good)), 4),
                         iters=round(float(np.mean(conv[good])), 3) if good.any() else None)
        if mode == "learned":
            P = dU0.to_host().reshape(-1, 2)
            res[mode]["init_rms_px"] = round(float(np.sqrt(np.mean(
                np.sum((P - w[:, 4:]) ** 2, 1)))), 4)
    plan.close()
    im.close()
    src.close()
    return res


# ---------------------------------------------------------------------------
# inference: learned starting point + gate + fallback
# ---------------------------------------------------------------------------
class LearnedInit:
    """Device buffers and launch sequence of the learned-init local step on
    a fixed node set.

    run(k): sparse integer search -> prior at every node -> network ->
    IC-GN from [F0 U0] (at most max_iter evaluations) -> gate (convIt in
    [1, max_iter] and ZNCC >= min_zncc) -> the search-init start of the
    failed nodes (the points of the search grid their transfer reads) ->
    IC-GN rerun of those nodes (the reference's 100-evaluation cap).

    Steps of consecutive frames are pipelined over up to three contexts
    (streams): the prior search and network of frame k run on the images'
    context while IC-GN and the gate of frame k-1 run on ctx_icgn and the
    fallback of frame k-2 on ctx_fb (each solver context with its own plan
    of the same nodes).  The per-frame buffers are rotated over three sets;
    outputs(k) returns frame k's (U, F, convIt, ZNCC) device arrays.

    net_step (px, optional): the network runs on a regular grid of that
    spacing over the nodes' bounding box instead of at every node, and its
    [F0 U0] reach the nodes by bilinear interpolation
    (aldic_field_to_nodes_dev) -- fewer network passes where the nodes are
    denser than the field's variation."""

    NSETS = 3

    def __init__(self, net: Network, images: Images, d_f, d_g, plan: Plan, d_coords, n,
                 d_winsize, prior_grid, fallback_grid, region=(50, 50, 50, 50), max_iter=30,
                 min_zncc=0.9, tol=1e-3, ctx_fb: Context | None = None, plan_fb: Plan | None = None,
                 ctx_icgn: Context | None = None, net_step=None):
        ctx = images.ctx
        self.ctx = ctx
        self.ctx_icgn = ctx_icgn or ctx
        self.ctx_fb = ctx_fb or ctx
        self.net, self.images, self.plan = net, images, plan
        self.plan_fb = plan_fb or plan
        self.d_f, self.d_g = d_f, d_g  # FP32 planes the prior search runs on
        self.d_coords, self.n, self.d_winsize = d_coords, int(n), d_winsize
        self.region = tuple(int(r) for r in region)
        # fallback_grid = (gridx, gridy, winsize, step): the search-init grid;
        # a failed node restarts from exactly the U0 that grid gives it
        self.fb_grid = fallback_grid
        self.max_iter, self.min_zncc, self.tol = int(max_iter), float(min_zncc), float(tol)
        # prior_grid = (gridx, gridy, winsize, step): a sparse search grid
        self.gx, self.gy, self.pws, self.pstep = prior_grid
        # XList = gridx(1):step:gridx(2)-winsize-step (funIntegerSearch.m:106-107)
        nX = (self.gx[1] - self.pws - self.pstep - self.gx[0]) // self.pstep + 1
        nY = (self.gy[1] - self.pws - self.pstep - self.gy[0]) // self.pstep + 1
        self.nX, self.nY = int(nX), int(nY)
        npnt = self.nX * self.nY
        self.g = {k: DeviceArray(ctx, npnt, np.float64) for k in ("x", "y", "u", "v", "cc")}
        self.g_ipr = DeviceArray(ctx, npnt, np.int32)
        self.g_ipc = DeviceArray(ctx, npnt, np.int32)
        nn = self.n
        # the network's points: the nodes, or a regular grid over them
        self.net_step = None if net_step is None else float(net_step)
        self.n_net, self.d_net_coords = nn, d_coords
        if self.net_step is not None:
            xy = self._node_box()
            xs = np.arange(xy[0], xy[1] + self.net_step, self.net_step)
            ys = np.arange(xy[2], xy[3] + self.net_step, self.net_step)
            self.nXn, self.nYn = len(xs), len(ys)
            X, Y = np.meshgrid(xs, ys, indexing="xy")  # point ci + cj*nX
            gc = np.concatenate([X.reshape(-1), Y.reshape(-1)]).astype(np.float64)
            self.n_net = self.nXn * self.nYn
            self.x0n, self.y0n = float(xs[0]), float(ys[0])
            self.d_net_coords = DeviceArray.from_host(ctx, gc, np.float64)
        self.bufs = []
        for _ in range(self.NSETS):
            if self.net_step is not None:
                ng = self.n_net
                grid_bufs = dict(prior_g=DeviceArray(ctx, 2 * ng, np.float32),
                                 U0_g=DeviceArray(ctx, 2 * ng, np.float32),
                                 F0_g=DeviceArray(ctx, 4 * ng, np.float32))
            else:
                grid_bufs = {}
            self.bufs.append(dict(**grid_bufs,
                prior=DeviceArray(ctx, 2 * nn, np.float32), U0=DeviceArray(ctx, 2 * nn, np.float32),
                F0=DeviceArray(ctx, 4 * nn, np.float32), list=DeviceArray(ctx, max(nn, 1), np.int32),
                count=DeviceArray(ctx, 1, np.int32), U=DeviceArray(ctx, 2 * nn, np.float32),
                F=DeviceArray(ctx, 4 * nn, np.float32), conv=DeviceArray(ctx, nn, np.int32),
                z=DeviceArray(ctx, nn, np.float32)))

    def outputs(self, k):
        b = self.bufs[k % self.NSETS]
        return b["U"], b["F"], b["conv"], b["z"]

    def _node_box(self):
        """(xmin, xmax, ymin, ymax) of the nodes (read once from the device)."""
        c = np.empty(2 * self.n, dtype=np.float64)
        check(lib().aldic_memcpy_dtoh(self.ctx.ptr, c.ctypes.data, self.d_coords, c.nbytes))
        x, y = c[:self.n], c[self.n:]
        return float(x.min()), float(x.max()), float(y.min()), float(y.max())

    # ---- the three stages of one frame, each on its own context ----
    def _stage_net(self, sb):
        A, im, b = self.ctx, self.images, self.bufs[sb]
        # 1. coarse prior from a sparse integer search
        integer_search_dev(A, self.d_f, self.d_g, im.M, im.N, self.gx, self.gy, self.pws,
                           self.pstep, self.region, self.g["x"].ptr, self.g["y"].ptr,
                           self.g["u"].ptr, self.g["v"].ptr, self.g["cc"].ptr, self.g_ipr.ptr,
                           self.g_ipc.ptr)
        x0 = self.gx[0] + self.pws / 2.0
        y0 = self.gy[0] + self.pws / 2.0
        if self.net_step is None:
            grid_to_nodes_dev(A, self.g["u"].ptr, self.g["v"].ptr, self.nX, self.nY, x0, y0,
                              self.pstep, self.d_coords, self.n, b["prior"].ptr)
            # 2. network -> [F0 U0]
            self.net.predict_dev(im, self.d_coords, self.n, b["prior"].ptr, b["U0"].ptr,
                                 b["F0"].ptr)
            return
        # 2'. the network on its grid, [F0 U0] interpolated to the nodes
        grid_to_nodes_dev(A, self.g["u"].ptr, self.g["v"].ptr, self.nX, self.nY, x0, y0,
                          self.pstep, self.d_net_coords.ptr, self.n_net, b["prior_g"].ptr)
        self.net.predict_dev(im, self.d_net_coords.ptr, self.n_net, b["prior_g"].ptr,
                             b["U0_g"].ptr, b["F0_g"].ptr)
        field_to_nodes_dev(A, b["U0_g"].ptr, b["F0_g"].ptr, self.nXn, self.nYn, self.x0n, self.y0n,
                           self.net_step, self.d_coords, self.n, b["U0"].ptr, b["F0"].ptr)

    def _stage_icgn(self, sb):
        B, im, b = self.ctx_icgn, self.images, self.bufs[sb]
        # 3. IC-GN, 4. gate
        local_icgn_f0_dev(im, self.plan, self.d_coords, b["U0"].ptr, b["F0"].ptr, self.d_winsize,
                          self.tol, b["U"].ptr, b["F"].ptr, d_conv=b["conv"].ptr,
                          d_zncc=b["z"].ptr, max_iter=self.max_iter, ctx=B)
        gate_dev(B, self.n, b["conv"].ptr, b["z"].ptr, self.max_iter, self.min_zncc,
                 b["list"].ptr, b["count"].ptr)

    def _stage_fallback(self, sb):
        C, im, b = self.ctx_fb, self.images, self.bufs[sb]
        # 5. the search-init start of the failed nodes (only the grid points
        # they need), 6. IC-GN rerun of those nodes
        fgx, fgy, fws, fstep = self.fb_grid
        search_grid_nodes_dev(C, self.d_f, self.d_g, im.M, im.N, fgx, fgy, fws, fstep,
                              self.region, self.d_coords, self.n, b["list"].ptr,
                              b["count"].ptr, b["U0"].ptr, b["F0"].ptr)
        local_icgn_list_dev(im, self.plan_fb, self.d_coords, b["U0"].ptr, b["F0"].ptr,
                            self.d_winsize, b["list"].ptr, b["count"].ptr, self.tol, b["U"].ptr,
                            b["F"].ptr, d_conv=b["conv"].ptr, d_zncc=b["z"].ptr, ctx=C)

    def run(self, k=0, graphs=None):
        """Frame k; graphs (from capture()) replays the stages instead of
        issuing their launches."""
        A, B, C = self.ctx, self.ctx_icgn, self.ctx_fb
        sb = k % self.NSETS
        # marks: A(sb) network done, B(3+sb) gate done, C(6 if sb == 0 else sb) fallback done
        mA, mB, mC = sb, 3 + sb, 6 if sb == 0 else sb
        if C is not A:
            A.wait(C, mC)  # the fallback of frame k-3 has released set sb
        if graphs:
            graphs[0][sb].launch(A)
        else:
            self._stage_net(sb)
        if B is not A:
            A.record(mA)
            B.wait(A, mA)
        if graphs:
            graphs[1][sb].launch(B)
        else:
            self._stage_icgn(sb)
        if C is not B:
            B.record(mB)
            C.wait(B, mB)
        if graphs:
            graphs[2][sb].launch(C)
        else:
            self._stage_fallback(sb)
        if C is not A:
            C.record(mC)

    def capture(self):
        """One HIP graph per stage and buffer set (run every stage once
        first, so staging buffers and kernel attributes exist)."""
        out = [[], [], []]
        for sb in range(self.NSETS):
            out[0].append(self.ctx.capture(lambda: self._stage_net(sb)))
            out[1].append(self.ctx_icgn.capture(lambda: self._stage_icgn(sb)))
            out[2].append(self.ctx_fb.capture(lambda: self._stage_fallback(sb)))
        return out

    def fallback_count(self, k=0):
        return int(self.bufs[k % self.NSETS]["count"].to_host()[0])

    def close(self):
        """Free the device buffers (the caller synchronises first)."""
        for b in self.bufs:
            for a in b.values():
                a.free()
        self.bufs = []
        for a in (*self.g.values(), self.g_ipr, self.g_ipc):
            a.free()
        if self.net_step is not None:
            self.d_net_coords.free()
