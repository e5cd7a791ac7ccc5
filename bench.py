"""Benchmark of ALDIC's local matching path on one MI355X.

Headline workload (the large stereo configuration the metric is quoted on,
learned vs search start head to head): a 4096 x 3000 12-bit speckle pair
(0.01 spots/px^2, seed 5); the right view is the left one under a 25 px
x-disparity plus a Gaussian bulge of 15 px peak (sigma 600 px), so up to
40 px; ~175k nodes on an 8 px grid outside a centred circular hole of
radius 300 px, each node with its own subset size from 8 px at the hole
edge up to 64 px away from it (the ROI's high-x margin widened by the
largest disparity, so that deformed subsets stay inside the image).  The
2048 x 2048 tilted-plane pair (0.015 spots/px^2, seed 2, 40 px +/- 6 px
disparity, ~55k nodes) is reported beside it.

One step = the whole local matching pass over one pair with inputs resident
in HBM:

  search start:  integer search (ZNCC, radius 50, 33 px templates on a
                 32 px grid; numerators by 128 x 128 FFT blocks, the
                 library's default for a grid this size) -> initial guess of
                 every node from the search grid -> first-order IC-GN of
                 every node (tol 1e-3, <= 101 evaluations);
  learned start: sparse integer search (same templates and radius, 256 px
                 grid; direct sums, the default for a search this small) ->
                 coarse prior on a 32 px grid over the nodes (--net-step)
                 -> CNN on the 48 x 48 patch pair at each grid point and at
                 point + round(prior) -> [F0 U0] bilinear to every node (the
                 fields vary over tens of px, the nodes are 8 px apart)
                 -> IC-GN (at most 31 evaluations) -> gate (more than 30
                 evaluations or ZNCC < 0.9) -> the search start of the failed
                 nodes (only the grid points they read) -> IC-GN rerun.

Each start runs with three IC-GN stopping rules: the reference's relative
||b|| rule (parity), the step-size rule (the update moves the subset by
less than 1e-3 px) and its a-posteriori form (also converged once the
contraction of the last two updates predicts a distance to the fixed point
below 1e-3 px; aldic_plan_set_stopping), each with its distance from the
reference rule's answers.  The network is trained first
(not timed) on the device, on 2M patch pairs labelled by IC-GN on seeded
speckle mosaics, with tile translations within +-2 px: the network predicts
the residual around the rounded sparse-search prior, whose interpolation
error on these fields is below 1 px (a larger residual goes through the
gate and the search fallback).

Converged = status 1..100, ZNCC >= 0.9 and not a bad point by the
reference's rule (ConvIt > 99 or > max(mean + std, 6) over the good ones,
LocalICGNQuadtree.m:167-192).  The JSON line's value is the converged
subsets/s of the learned start with the a-posteriori rule on the headline
workload; the other starts, rules and the second workload sit beside it,
with one frame of the incremental sequence (temporal IC-GN and the two
augmented-Lagrangian iterations replayed as one HIP graph, `seq2048_alm`),
each with its RMS error against the known field; the learned start with
the network at every node is timed beside the grid (`network_at_every_node`).
The roofline is the
dominant kernel of that step (its HIP-event time, algorithmic bytes); on
rank 0 the FP64 CPU restatement (oracle, OpenMP) runs the search-start path
with the reference's rule on the same pair.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Hardware queues of this process (HIP reads this once, at its first call).
# The learned step runs on three contexts of two streams each, and the
# runtime maps streams to its hardware queues in turn: with HIP's default of
# four, the network's stream and the fallback's share a queue, so the next
# frame's network waits behind this frame's fallback, which waits for this
# frame's IC-GN (4.91 against 5.10 ms per step on the headline workload).
# A larger setting of the environment is kept.
try:
    _hw_queues = int(os.environ.get("GPU_MAX_HW_QUEUES", "") or 4)
except ValueError:
    _hw_queues = 4
if _hw_queues < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

import __graft_entry__ as ge  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
ROOFLINE_REPS = 5
TOL_DP = 1e-3  # step-size rule threshold (px)
# HBM bytes per launch of the dominant IC-GN kernel (65 px class, learned
# start, a-posteriori rule, headline workload), rocprofv3 --pmc FETCH_SIZE /
# WRITE_SIZE in KiB per dispatch, each scaled by the same counter's reading
# on a device-to-device copy of known size (FETCH x 1.9994, WRITE x 1);
# measured with this bench (profiles/icgn_c3_bulge4096_hbm_traffic.md)
ICGN_C3_TRAFFIC = (3471772.8 * 1.9994 + 22939.0) * 1024.0


def load(name, rel):
    return ge._load_file(name, os.path.join(ROOT, rel))


def icgn_flops(side, iters):
    """FP32 flops of one subset: 120 per pixel per evaluation (warp,
    weights, 4x4 separable combine, moments, residual, b) + 46 per pixel
    once (steepest-descent images, Hessian)."""
    s = np.asarray(side, dtype=np.float64)
    return (120.0 * np.asarray(iters, dtype=np.float64) + 46.0) * s * s


def icgn_bytes(side):
    """Algorithmic HBM bytes of one IC-GN subset of side s (N = s^2):
    f, DfDx, DfDy once (12 N), the deformed-image footprint with a 4 px
    margin (4 (s+8)^2), and 32 B of outputs."""
    s = np.asarray(side, dtype=np.float64)
    return 12.0 * s * s + 4.0 * (s + 8) ** 2 + 32.0


WORKLOADS = {
    "bulge4096x3000": dict(M=4096, N=3000, density=0.01, seed=5, field="bulge", max_disp=40),
    "tilt2048": dict(M=2048, N=2048, density=0.015, seed=2, field="stereo", max_disp=46),
}


def make_workload(name):
    synth = load("aldic_synth_b", "3d-stereo-aldic_amd/synth.py")
    mesh = load("aldic_mesh_b", "3d-stereo-aldic_amd/mesh.py")
    spec = WORKLOADS[name]
    M, N = spec["M"], spec["N"]
    uf, vf = (synth.field_bulge(M, N) if spec["field"] == "bulge" else synth.field_stereo(M, N))
    f12, g12, _ = synth.make_pair(M, N, uf, vf, density=spec["density"], seed=spec["seed"],
                                  bits=12)
    f = (f12 - f12.mean()) / f12.std(ddof=1)
    g = (g12 - g12.mean()) / g12.std(ddof=1)
    centre = ((M + 1) / 2.0, (N + 1) / 2.0)
    coords, ws = mesh.hole_nodes(M, N, centre, 300.0, extra=(0, spec["max_disp"] + 4, 0, 0))
    return dict(name=name, M=M, N=N, f=np.asfortranarray(f), g=np.asfortranarray(g),
                coords=coords, ws=ws, uf=uf, vf=vf, search_ws=32, search_step=32, radius=50,
                spec=spec)


def window_mean_truth(W, fun):
    """The known field averaged over each node's square subset (a 9 x 9
    Gauss-Legendre rule): what a first-order subset can represent of a
    curved field, beside the pointwise value."""
    x, wq = np.polynomial.legendre.leggauss(9)
    cx, cy = W["coords"][:, 0], W["coords"][:, 1]
    h = W["ws"] / 2.0
    acc = np.zeros(len(cx))
    for a, wa in zip(x, wq):
        for b, wb in zip(x, wq):
            acc += wa * wb * fun(cx + a * h, cy + b * h)
    return acc / 4.0


def host_info(threads):
    model = platform.processor() or ""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "affinity_cpus": avail,
            "threads": threads}


def cpu_threads():
    """Every core this process may run on, at most the 16 of the host's
    share for one GPU (OMP_NUM_THREADS when it says fewer)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    t = min(avail, 16)
    return min(t, env) if env > 0 else t


def cpu_baseline(W, n_sample, threads, U_gpu=None):
    """The FP64 oracle (C + OpenMP, test infrastructure) on the same pair:
    the whole search grid, then IC-GN (reference rule) of a sample of nodes
    (all when n_sample is 0).  Also returns the oracle's accuracy on that
    sample and its distance to the GPU's search-start result there."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    f, g = W["f"], W["g"]
    M, N = W["M"], W["N"]
    r = W["radius"]
    t0 = time.perf_counter()
    s = orc.integer_search(f, g, [1, M], [1, N], W["search_ws"], W["search_step"], [r, r, r, r],
                           nthreads=threads)
    t_search = time.perf_counter() - t0
    nY, nX = s["u"].shape
    n = len(W["ws"])
    if n_sample <= 0 or n_sample >= n:
        idx = np.arange(n)
    else:
        idx = np.sort(np.random.default_rng(0).choice(n, size=n_sample, replace=False))
    coords = np.asfortranarray(W["coords"][idx])
    x0 = W["search_ws"] / 2 + 1
    fx = np.clip((coords[:, 0] - x0) / W["search_step"], 0, nX - 1)
    fy = np.clip((coords[:, 1] - x0) / W["search_step"], 0, nY - 1)
    ix = np.minimum(np.floor(fx).astype(int), max(nX - 2, 0))
    iy = np.minimum(np.floor(fy).astype(int), max(nY - 2, 0))
    tx, ty = fx - ix, fy - iy
    U0 = np.zeros(2 * len(idx))
    for k, key in ((0, "u"), (1, "v")):
        a = s[key]
        U0[k::2] = ((1 - tx) * (1 - ty) * a[iy, ix] + tx * (1 - ty) * a[iy, ix + 1] +
                    (1 - tx) * ty * a[iy + 1, ix] + tx * ty * a[iy + 1, ix + 1])
    im = orc.Images(f, g)
    t0 = time.perf_counter()
    U, F, H, conv, hole, z = orc.local_icgn(im, coords, U0, W["ws"][idx], 1e-3, nthreads=threads)
    t_icgn = time.perf_counter() - t0
    pkg = ge.load_package()
    bad, _ = pkg.icgn_bad_points(conv)
    ok = (conv >= 1) & (conv <= 100) & (z >= 0.9)
    ok[bad] = False
    t_total = t_icgn + t_search * len(idx) / n  # the search pro-rated to the sample
    e = np.hypot(U[0::2] - W["uf"](coords[:, 0], coords[:, 1]),
                 U[1::2] - W["vf"](coords[:, 0], coords[:, 1]))
    diag = {"oracle_rms_px": float(np.sqrt(np.mean(e[ok] ** 2))),
            "oracle_converged_frac": float(ok.mean()), "nodes": int(len(idx))}
    if U_gpu is not None:
        du = np.abs(U_gpu.reshape(-1, 2)[idx] - U.reshape(-1, 2)).max(axis=1)
        both = ok & np.isfinite(du)
        diag["gpu_vs_oracle_max_du_px"] = float(np.nanmax(du[both]))
        diag["gpu_vs_oracle_p99_du_px"] = float(np.quantile(du[both], 0.99))
    info = host_info(threads)
    return diag, {"value": float(ok.sum() / t_total), "unit": "converged subsets/s",
                  "cores": threads, "kind": "port", **info,
                  "sample": (f"FP64 oracle, search-start path, reference stopping rule: integer "
                             f"search over the full {nX}x{nY} grid ({t_search:.1f} s"
                             + (", pro-rated to the sample" if len(idx) < n else "")
                             + f") + IC-GN of {len(idx)} of the {n} nodes ({t_icgn:.1f} s), "
                             f"{threads} OpenMP threads")}


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def subset_class(side):
    return np.where(side <= 11, 0, np.where(side <= 17, 1, np.where(side <= 33, 2,
                    np.where(side <= 65, 3, 4))))


def accuracy(api, W, U, conv, z, truth_mean):
    """Converged mask (status 1..100, ZNCC >= 0.9, not a bad point by the
    reference's rule), RMS against the pointwise field and against its
    subset means, and per-size statistics."""
    cx, cy = W["coords"][:, 0], W["coords"][:, 1]
    bad, _ = api.icgn_bad_points(conv)
    status_ok = (conv >= 1) & (conv <= 100) & (z >= 0.9)
    ok = status_ok.copy()
    ok[bad] = False
    e = np.hypot(U[0::2] - W["uf"](cx, cy), U[1::2] - W["vf"](cx, cy))
    em = np.hypot(U[0::2] - truth_mean[0], U[1::2] - truth_mean[1])

    def rms(a, sel):
        return float(np.sqrt(np.mean(a[sel] ** 2))) if sel.any() else float("nan")

    by_size = {}
    for w in np.unique(W["ws"]):
        sel = W["ws"] == w
        oks = ok & sel
        by_size[int(w)] = {"nodes": int(sel.sum()), "converged": int(oks.sum()),
                           "rms_px": round(rms(e, oks), 5),
                           "rms_vs_subset_mean_px": round(rms(em, oks), 5),
                           "mean_evals": round(float(conv[oks].mean()), 3) if oks.any() else None}
    return dict(ok=ok, status_ok=int(status_ok.sum()), bad_points=int(len(bad)),
                rms=rms(e, ok), rms_mean=rms(em, ok), by_size=by_size)


def timed(ctxs, api, step, steps, warmup):
    """Wall time per step of `steps` calls step(k) after `warmup` untimed
    ones, every context drained on both sides; stage times summed over the
    contexts."""
    for k in range(warmup):
        step(k)
    for c in ctxs:
        c.synchronize()
    for c in ctxs:
        api.set_timing(c, True)
        api.stage_times(c, reset=True)
    t0 = time.perf_counter()
    for k in range(warmup, warmup + steps):
        step(k)
    for c in ctxs:
        c.synchronize()
    dt = (time.perf_counter() - t0) / steps
    stages = {}
    for c in ctxs:
        for name, (ms, cnt) in api.stage_times(c, reset=True).items():
            m0, c0 = stages.get(name, (0.0, 0))
            stages[name] = (m0 + ms, c0 + cnt)
        api.set_timing(c, False)
    return dt, stages, warmup + steps - 1


def global_step_mode(pkg, args):
    """The rows around the local step on the stereo workload's geometry:
    the quadtree mesh of the 2048 x 2048 hole mask (host), Subpb2 on it
    (device, FP64), the NaN fill (host) and the raw 16-bit decode (device)
    against the FP64 upload.  Prints one JSON line of timings."""
    api = pkg.api
    from aldic_amd import mesh as mesh_mod, quadtree
    synth = load("aldic_synth_b", "3d-stereo-aldic_amd/synth.py")
    M = N = 2048
    uf, vf = synth.field_stereo(M, N)
    mask = mesh_mod.hole_mask(M, N, (1024.5, 1024.5), 300.0)
    step, ws = 16, 32
    coords0, nX, nY = mesh_mod.uniform_grid([1, M], [1, N], ws, step)
    x = coords0[:, 0].reshape(nY, nX)
    y = coords0[:, 1].reshape(nY, nX)
    out = {"mode": "global-step", "image": [M, N], "winstepsize": step, "winsize": ws,
           "elementMinSize": 4}
    t0 = time.perf_counter()
    DICmesh = quadtree.MeshSetUp(x, y, {"ImgSize": (M, N), "winstepsize": step})
    DICmesh["elementMinSize"] = 4
    U0 = np.empty(2 * len(coords0))
    U0[0::2] = uf(coords0[:, 0], coords0[:, 1])
    U0[1::2] = vf(coords0[:, 0], coords0[:, 1])
    qm, _, U0q = quadtree.GenerateQuadtreeMesh(U0, None, mask, DICmesh,
                                               {"winstepsize": step, "winsize": ws})
    out["mesh_s"] = round(time.perf_counter() - t0, 3)
    c, el = qm["coordinatesFEM"], qm["elementsFEM"]
    n = len(c)
    out.update(nodes=n, elements=len(el), hanging=int((el[:, 4:] > 0).sum()),
               dirichlet=len(qm["dirichlet"]))
    progress(f"quadtree mesh: {n} nodes, {len(el)} elements ({out['mesh_s']} s)")
    ctx = pkg.default_context()
    t0 = time.perf_counter()
    fem = pkg.FEMesh(c, el, qm["dirichlet"], None, 2, ctx=ctx)
    out["fem_setup_s"] = round(time.perf_counter() - t0, 3)
    info = fem.info()
    out["nnz"] = info["nnz"]
    h = 1e-3
    U = np.empty(2 * n)
    U[0::2] = uf(c[:, 0], c[:, 1])
    U[1::2] = vf(c[:, 0], c[:, 1])
    F = np.zeros(4 * n)
    F[0::4] = (uf(c[:, 0] + h, c[:, 1]) - uf(c[:, 0] - h, c[:, 1])) / (2 * h)
    F[1::4] = (vf(c[:, 0] + h, c[:, 1]) - vf(c[:, 0] - h, c[:, 1])) / (2 * h)
    F[2::4] = (uf(c[:, 0], c[:, 1] + h) - uf(c[:, 0], c[:, 1] - h)) / (2 * h)
    F[3::4] = (vf(c[:, 0], c[:, 1] + h) - vf(c[:, 0], c[:, 1] - h)) / (2 * h)
    rng = np.random.default_rng(1)
    Un = U + rng.normal(scale=0.02, size=2 * n)  # a local step's noise
    DA = api.DeviceArray
    dU = DA.from_host(ctx, Un, np.float64)
    dF = DA.from_host(ctx, F, np.float64)
    dO = DA(ctx, 2 * n, np.float64)
    mu = 1e-3
    beta = 1e-3 * step * step * mu
    iters = fem.solve_dev(ctx, beta, mu, dU.ptr, dF.ptr, dO.ptr, winstepsize=step)
    ctx.synchronize()
    reps = max(args.steps, 5)
    api.set_timing(ctx, True)
    api.stage_times(ctx, reset=True)
    t0 = time.perf_counter()
    for _ in range(reps):
        fem.solve_dev(ctx, beta, mu, dU.ptr, dF.ptr, dO.ptr, winstepsize=step)
    ctx.synchronize()
    dt = (time.perf_counter() - t0) / reps
    st = api.stage_times(ctx, reset=True)
    api.set_timing(ctx, False)
    Uh = dO.to_host()
    # nodes in no element (inside the hole) get 0, as in the reference
    used = np.zeros(n, bool)
    used[el[el > 0] - 1] = True
    uu = np.repeat(used, 2)
    out["subpb2_ms"] = round(dt * 1e3, 4)
    out["subpb2_event_ms"] = round(st["subpb2"][0] / st["subpb2"][1], 4)
    out["cg_iterations"] = iters
    # bytes the CG iterations must touch: CSR values + columns, five 2-vectors
    per_it = 12 * info["nnz"] + n * (5 * 16 + 8 + 1)
    out["cg_bytes_per_iteration"] = per_it
    out["cg_gbs"] = round(per_it * iters / dt / 1e9, 2) if dt > 0 else None
    out["solved_nodes"] = int(used.sum())
    out["subpb2_rms_vs_field_px"] = round(float(np.sqrt(np.mean((Uh - U)[uu] ** 2))), 5)
    out["subpb2_rms_input_px"] = round(float(np.sqrt(np.mean((Un - U)[uu] ** 2))), 5)
    progress(f"subpb2 {out['subpb2_ms']} ms, {iters} CG iterations")
    fem.close()
    # NaN fill of 2% bad nodes (host)
    Ub = Un.copy()
    bad = rng.choice(n, n // 50, replace=False)
    Ub[2 * bad] = np.nan
    Ub[2 * bad + 1] = np.nan
    t0 = time.perf_counter()
    Uf = pkg.nan_fill(c, Ub, None, mask)
    out["nan_fill_ms"] = round((time.perf_counter() - t0) * 1e3, 2)
    out["nan_fill_nodes"] = int(len(bad))
    # raw 16-bit decode vs the FP64 upload
    f12, g12, _ = synth.make_pair(M, N, uf, vf, density=0.015, seed=2, bits=12)
    rawL = np.ascontiguousarray(f12.T).astype(np.uint16)
    rawR = np.ascontiguousarray(g12.T).astype(np.uint16)
    roi = {"gridx": [1, M], "gridy": [1, N]}
    (fn, gn), _ = pkg.funNormalizeImg([f12, g12], roi)
    for name, make in (("raw16", lambda: api.Images.from_raw(rawL, rawR, mask, roi, ctx=ctx)),
                       ("fp64", lambda: api.Images(fn, gn, mask, ctx=ctx))):
        make().close()
        t0 = time.perf_counter()
        for _ in range(3):
            make().close()
        out[f"images_{name}_ms"] = round((time.perf_counter() - t0) / 3 * 1e3, 2)
    progress(f"images: raw16 {out['images_raw16_ms']} ms, fp64 {out['images_fp64_ms']} ms")
    print(json.dumps(out), flush=True)




def search_mode(pkg, args):
    """The integer search alone, direct lag tiles against FFT lag blocks, on
    the headline image size (4096 x 3000, a smoothed random field and its
    (25, -3) px shift; normalised): per configuration (template side, grid
    step, radius) the device time per search over the whole grid (5
    searches after one warmup, stage timer) and the number of points whose
    integer peaks differ between the methods.  Prints one JSON line."""
    api = pkg.api
    from scipy.ndimage import gaussian_filter
    M, N = 4096, 3000
    rng = np.random.default_rng(0)
    base = gaussian_filter(rng.normal(size=(M + 60, N + 60)), 1.2)
    f = np.asfortranarray(base[30:-30, 30:-30])
    g = np.asfortranarray(np.roll(base, (-25, 3), (0, 1))[30:-30, 30:-30])
    f = (f - f.mean()) / f.std()
    g = (g - g.mean()) / g.std()
    out = {"workload": "4096x3000 smoothed random field, (25, -3) px shift"}
    for ws, step, r in ((32, 32, 50), (64, 32, 50), (32, 32, 20), (16, 16, 10)):
        nX = len(range(1, M - ws - step + 1, step))
        nY = len(range(1, N - ws - step + 1, step))
        res, peaks = {}, {}
        for method in ("direct", "fft"):
            ctx = api.Context(0).set_search_method(method)
            DA = api.DeviceArray
            d_f = DA.from_host(ctx, f, np.float32)
            d_g = DA.from_host(ctx, g, np.float32)
            o = [DA(ctx, nX * nY, np.float64) for _ in range(5)]
            pk = [DA(ctx, nX * nY, np.int32) for _ in range(2)]

            def run():
                api.integer_search_dev(ctx, d_f.ptr, d_g.ptr, M, N, [1, M], [1, N], ws, step,
                                       [r, r, r, r], *[a.ptr for a in o], *[a.ptr for a in pk])
            run()
            ctx.synchronize()
            api.set_timing(ctx, True)
            api.stage_times(ctx, reset=True)
            for _ in range(5):
                run()
            ms, cnt = api.stage_times(ctx, reset=True)["search"]
            res[method] = round(ms / cnt, 3)
            peaks[method] = (pk[0].to_host(), pk[1].to_host())
            ctx.close()
        diff = int(np.sum((peaks["direct"][0] != peaks["fft"][0]) |
                          (peaks["direct"][1] != peaks["fft"][1])))
        out[f"side{ws + 1}_step{step}_r{r}"] = {"points": nX * nY, "lags": (2 * r + 3) ** 2,
                                                "direct_ms": res["direct"], "fft_ms": res["fft"],
                                                "speedup": round(res["direct"] / res["fft"], 2),
                                                "peaks_differ": diff}
        progress(f"search side {ws + 1}, r {r}: direct {res['direct']} ms, fft {res['fft']} ms")
    print(json.dumps(out), flush=True)


RULES = {"reference": (0, 0.0), "step": (1, TOL_DP), "posterior": (2, TOL_DP)}
HEAD_RULE = "posterior"  # the rule of the headline value and the side runs


def run_workload(pkg, W, params, args, headline):
    """Every start x stopping rule on one workload; returns the summaries,
    the roofline of the headline configuration's dominant kernel and the
    search start's U (reference rule) for the CPU comparison."""
    api = pkg.api
    from aldic_amd import learned
    M, N = W["M"], W["N"]
    n = len(W["ws"])
    ctx = api.Context(0)
    DA = api.DeviceArray
    d_f = DA.from_host(ctx, W["f"], np.float32)
    d_g = DA.from_host(ctx, W["g"], np.float32)
    im = api.Images.from_device(d_f.ptr, d_g.ptr, M, N, ctx=ctx)
    d_coords = DA.from_host(ctx, W["coords"], np.float64)  # n x 2 column-major
    d_ws = DA.from_host(ctx, W["ws"], np.int32)
    sws, sst, r = W["search_ws"], W["search_step"], W["radius"]
    nX = len(range(1, M - sws - sst + 1, sst))
    nY = len(range(1, N - sws - sst + 1, sst))
    ng = nX * nY
    s_out = [DA(ctx, ng, np.float64) for _ in range(5)]  # x y u v cc
    s_pk = [DA(ctx, ng, np.int32) for _ in range(2)]
    x0 = sws / 2 + 1
    # three contexts (streams) pipeline consecutive frames; each solver
    # context has its own plan of the same nodes
    ctx2, ctx3 = api.Context(0), api.Context(0)
    plan = api.Plan(W["ws"], n, ctx=ctx)
    plan2 = api.Plan(W["ws"], n, ctx=ctx2)
    plan3 = api.Plan(W["ws"], n, ctx=ctx3)
    sets = [dict(U0=DA(ctx, 2 * n, np.float32), U=DA(ctx, 2 * n, np.float32),
                 F=DA(ctx, 4 * n, np.float32), conv=DA(ctx, n, np.int32),
                 z=DA(ctx, n, np.float32)) for _ in range(2)]
    truth_mean = (window_mean_truth(W, W["uf"]), window_mean_truth(W, W["vf"]))

    def search_stage(sb):  # on ctx
        api.integer_search_dev(ctx, d_f.ptr, d_g.ptr, M, N, [1, M], [1, N], sws, sst,
                               [r, r, r, r], *[o.ptr for o in s_out], *[p.ptr for p in s_pk])
        api.grid_to_nodes_dev(ctx, s_out[2].ptr, s_out[3].ptr, nX, nY, x0, x0, sst,
                              d_coords.ptr, n, sets[sb]["U0"].ptr)

    def icgn_stage(sb):  # on ctx2
        b = sets[sb]
        api.local_icgn_dev(im, plan2, d_coords.ptr, b["U0"].ptr, d_ws.ptr, 1e-3, b["U"].ptr,
                           b["F"].ptr, None, b["conv"].ptr, None, b["z"].ptr, ctx=ctx2)

    def step_search(k):
        sb = k % 2
        ctx.wait(ctx2, sb)  # the IC-GN of step k-2 has released set sb
        search_stage(sb)
        ctx.record(2 + sb)
        ctx2.wait(ctx, 2 + sb)
        icgn_stage(sb)
        ctx2.record(sb)

    # plan-time autotune of the launch shapes on this node set (not timed)
    search_stage(0)
    b0 = sets[0]
    chosen = api.local_icgn_autotune_dev(im, plan, d_coords.ptr, b0["U0"].ptr, d_ws.ptr, 1e-3,
                                         b0["U"].ptr, b0["F"].ptr, b0["conv"].ptr, b0["z"].ptr,
                                         reps=3)
    for c in range(4):
        for pl in (plan2, plan3):
            pl.set_variant(0, c, chosen[c])
            pl.set_variant(1, c, chosen[c])
    progress(f"{W['name']}: autotuned launch shapes {list(chosen)}")
    inet = learned.Network(ctx, max_batch=(n + 15) // 16 * 16, params=params)
    li = learned.LearnedInit(inet, im, d_f.ptr, d_g.ptr, plan3, d_coords.ptr, n, d_ws.ptr,
                             prior_grid=([1, M], [1, N], sws, 256),
                             fallback_grid=([1, M], [1, N], sws, sst), region=(r, r, r, r),
                             max_iter=30, min_zncc=0.9, tol=1e-3, ctx_fb=ctx2, plan_fb=plan2,
                             ctx_icgn=ctx3, net_step=args.net_step or None)
    out = {}
    U_search_ref = None
    for rule, (code, tol_dp) in RULES.items():
        for pl in (plan2, plan3):
            pl.set_stopping(code, tol_dp)
        dt, st, last = timed([ctx, ctx2], api, step_search, args.steps, args.warmup)
        b = sets[last % 2]
        U = b["U"].to_host().astype(np.float64)
        acc = accuracy(api, W, U, b["conv"].to_host(), b["z"].to_host(), truth_mean)
        out[("search", rule)] = dict(dt=dt, stages=st, acc=acc, U=U)
        if rule == "reference":
            U_search_ref = U
        progress(f"{W['name']} search start, {rule} rule: {dt * 1e3:.3f} ms, "
                 f"{acc['ok'].sum()} converged, rms {acc['rms']:.4f} px")
        dt, st, last = timed([ctx, ctx2, ctx3], api, li.run, args.steps, args.warmup)
        lU, lF, lconv, lz = li.outputs(last)
        Ul = lU.to_host().astype(np.float64)
        acc = accuracy(api, W, Ul, lconv.to_host(), lz.to_host(), truth_mean)
        out[("learned", rule)] = dict(dt=dt, stages=st, acc=acc, fallback=li.fallback_count(last),
                                      U=Ul)
        progress(f"{W['name']} learned start, {rule} rule: {dt * 1e3:.3f} ms, "
                 f"{acc['ok'].sum()} converged, rms {acc['rms']:.4f} px, fallback "
                 f"{li.fallback_count(last)}")
    if args.net_step:
        # ---- the network at every node instead of on its grid, headline
        # rule: what the grid saves, and that the start is as good ----
        li_node = learned.LearnedInit(inet, im, d_f.ptr, d_g.ptr, plan3, d_coords.ptr, n, d_ws.ptr,
                                      prior_grid=([1, M], [1, N], sws, 256),
                                      fallback_grid=([1, M], [1, N], sws, sst),
                                      region=(r, r, r, r), max_iter=30, min_zncc=0.9, tol=1e-3,
                                      ctx_fb=ctx2, plan_fb=plan2, ctx_icgn=ctx3)
        for pl in (plan2, plan3):
            pl.set_stopping(*RULES[HEAD_RULE])
        dt, st, last = timed([ctx, ctx2, ctx3], api, li_node.run, args.steps, args.warmup)
        lU, lF, lconv, lz = li_node.outputs(last)
        acc = accuracy(api, W, lU.to_host().astype(np.float64), lconv.to_host(), lz.to_host(),
                       truth_mean)
        out[("learned", "posterior_per_node")] = dict(dt=dt, stages=st, acc=acc,
                                                 fallback=li_node.fallback_count(last))
        progress(f"{W['name']} learned start at every node, a-posteriori rule: {dt * 1e3:.3f} ms, "
                 f"{acc['ok'].sum()} converged, rms {acc['rms']:.4f} px, fallback "
                 f"{li_node.fallback_count(last)}")
        for c in (ctx, ctx2, ctx3):
            c.synchronize()
        li_node.close()
    # ---- the interpolating cubic B-spline option (not the reference's
    # interpolant; no parity with it): both starts under the headline rule.
    # Each frame needs its coefficient plane: the prefilter is timed on its
    # own and added to the step ----
    d_gc = DA(ctx, M * N, np.float32)
    dtp, _, _ = timed([ctx], api, lambda k: api.bspline_prefilter_dev(ctx, d_g.ptr, d_gc.ptr, M, N),
                      args.steps, args.warmup)
    d_gc.free()
    im.set_interp("bspline")
    for pl in (plan2, plan3):
        pl.set_stopping(*RULES[HEAD_RULE])
    dt, st, last = timed([ctx, ctx2], api, step_search, args.steps, args.warmup)
    b = sets[last % 2]
    acc = accuracy(api, W, b["U"].to_host().astype(np.float64), b["conv"].to_host(),
                   b["z"].to_host(), truth_mean)
    out[("search", "posterior_bspline")] = dict(dt=dt + dtp, stages=st, acc=acc, prefilter_dt=dtp)
    progress(f"{W['name']} search start, a-posteriori rule, B-spline: {(dt + dtp) * 1e3:.3f} ms "
             f"(prefilter {dtp * 1e3:.3f} ms), rms {acc['rms']:.4f} px")
    dt, st, last = timed([ctx, ctx2, ctx3], api, li.run, args.steps, args.warmup)
    lU, lF, lconv, lz = li.outputs(last)
    acc = accuracy(api, W, lU.to_host().astype(np.float64), lconv.to_host(), lz.to_host(),
                   truth_mean)
    out[("learned", "posterior_bspline")] = dict(dt=dt + dtp, stages=st, acc=acc, prefilter_dt=dtp,
                                           fallback=li.fallback_count(last))
    progress(f"{W['name']} learned start, a-posteriori rule, B-spline: {(dt + dtp) * 1e3:.3f} ms, "
             f"rms {acc['rms']:.4f} px, fallback {li.fallback_count(last)}")
    im.set_interp("local")
    for c in (ctx, ctx2, ctx3):
        c.synchronize()
    # ---- roofline of the headline step's dominant kernel: the learned
    # start's first IC-GN pass (a-posteriori rule), ROOFLINE_REPS launches
    # alone on one stream, HIP events on that stream ----
    roof = None
    if headline:
        plan.set_stopping(*RULES[HEAD_RULE])
        li._stage_net(0)
        ctx.synchronize()
        bb = li.bufs[0]
        ctx.synchronize()
        api.set_timing(ctx, True)
        api.stage_times(ctx, reset=True)
        for _ in range(ROOFLINE_REPS):
            api.local_icgn_f0_dev(im, plan, d_coords.ptr, bb["U0"].ptr, bb["F0"].ptr, d_ws.ptr,
                                  1e-3, bb["U"].ptr, bb["F"].ptr, d_conv=bb["conv"].ptr,
                                  d_zncc=bb["z"].ptr, max_iter=30)
        rst = api.stage_times(ctx, reset=True)
        api.set_timing(ctx, False)
        side = W["ws"] + 1
        cls = subset_class(side)
        # the dominant class: the most pixels (a profiler that serialises
        # the launches would skew stage times)
        dom_c = max((c for c in range(4) if f"icgn_c{c}" in rst),
                    key=lambda c: float((side[cls == c].astype(np.float64) ** 2).sum()))
        # the same whole solve (every size class, the workgroup passes) as
        # one launch per class on two streams (autotuned shapes) and as the
        # persistent launch fed by the size-sorted queue: wall time per
        # solve over ROOFLINE_REPS solves, the stream drained on both sides
        plan_p = api.Plan(W["ws"], n, ctx=ctx).set_persistent(True)
        plan_p.set_stopping(*RULES[HEAD_RULE])
        solve = {}
        for name, pl in (("per_class", plan), ("persistent", plan_p)):
            def one():
                api.local_icgn_f0_dev(im, pl, d_coords.ptr, bb["U0"].ptr, bb["F0"].ptr,
                                      d_ws.ptr, 1e-3, bb["U"].ptr, bb["F"].ptr,
                                      d_conv=bb["conv"].ptr, d_zncc=bb["z"].ptr, max_iter=30)
            one()
            ctx.synchronize()
            t0 = time.perf_counter()
            for _ in range(ROOFLINE_REPS):
                one()
            ctx.synchronize()
            solve[name] = round((time.perf_counter() - t0) / ROOFLINE_REPS * 1e3, 4)
        plan_p.close()
        progress(f"whole IC-GN solve: per class {solve['per_class']} ms, "
                 f"persistent {solve['persistent']} ms")
        # the dominant class alone (a plan over its nodes only, so that no
        # other class's blocks share the GPU with it), last in the run:
        # ROOFLINE_REPS launches from the same starts, HIP events on its stream
        idx = np.nonzero(cls == dom_c)[0]
        m = len(idx)
        U0h, F0h = bb["U0"].to_host(), bb["F0"].to_host()
        sub = dict(c=DA.from_host(ctx, np.asfortranarray(W["coords"][idx]), np.float64),
                   u0=DA.from_host(ctx, U0h.reshape(-1, 2)[idx].reshape(-1), np.float32),
                   f0=DA.from_host(ctx, F0h.reshape(-1, 4)[idx].reshape(-1), np.float32),
                   ws=DA.from_host(ctx, W["ws"][idx], np.int32),
                   U=DA(ctx, 2 * m, np.float32), F=DA(ctx, 4 * m, np.float32),
                   conv=DA(ctx, m, np.int32), z=DA(ctx, m, np.float32))
        plan_d = api.Plan(W["ws"][idx], m, ctx=ctx)
        plan_d.set_stopping(*RULES[HEAD_RULE])
        plan_d.set_variant(0, dom_c, chosen[dom_c])
        ctx.synchronize()
        api.set_timing(ctx, True)
        for rep in range(ROOFLINE_REPS + 1):
            if rep == 1:
                ctx.synchronize()
                api.stage_times(ctx, reset=True)  # the first launch warms up
            api.local_icgn_f0_dev(im, plan_d, sub["c"].ptr, sub["u0"].ptr, sub["f0"].ptr,
                                  sub["ws"].ptr, 1e-3, sub["U"].ptr, sub["F"].ptr,
                                  d_conv=sub["conv"].ptr, d_zncc=sub["z"].ptr, max_iter=30)
        ctx.synchronize()
        dst = api.stage_times(ctx, reset=True)
        api.set_timing(ctx, False)
        dom = f"icgn_c{dom_c}"
        ms = dst[dom][0] / dst[dom][1]
        it = np.clip(sub["conv"].to_host(), 1, 31)
        nbytes = float(icgn_bytes(side[idx]).sum())
        flops = float(icgn_flops(side[idx], it).sum())
        units, evals = m, float(it.mean())
        plan_d.close()
        for a in sub.values():
            a.free()
        achieved = nbytes / (ms * 1e-3) / 1e9
        tflops = flops / (ms * 1e-3) / 1e12
        roof = {"bound": "hbm", "limiter": "VALU issue (the fp32 entry)", "kernel": dom,
                "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": ICGN_C3_TRAFFIC if dom == "icgn_c3" else None,
                "kernel_ms": round(ms, 4), "launches": ROOFLINE_REPS,
                "algorithmic_bytes": nbytes, "subsets": units, "mean_evaluations": round(evals, 3),
                "fp32": {"achieved": round(tflops, 2), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(tflops / FP32_PEAK_TFLOPS, 4)},
                "pass": "learned start, first IC-GN pass (<= 31 evaluations), a-posteriori "
                        "step-size rule; the class's nodes alone",
                "in_whole_solve_ms": round(rst[dom][0] / rst[dom][1], 4),
                "whole_solve_ms": solve}
    inet.close()
    for pl in (plan, plan2, plan3):
        pl.close()
    im.close()
    return out, roof, U_search_ref, dict(nodes=n, search_points=ng, grid=(nX, nY),
                                         shapes=[int(c) for c in chosen])


def alm_block(pkg, args):
    """The incremental sequence's per-frame work on the device: a 2048 x 2048
    temporal pair (0.1% x-stretch, 0.7 px y-translation; 12-bit speckle,
    0.015 spots/px^2, seed 3) over the quadtree mesh of a centred hole
    (r = 300 px; 64 px elements refined to 8 px at the hole, 64 px
    subsets).  Timed: the temporal IC-GN of every node (device, the plan's
    size classes), and the two ALM iterations of TemporalMatch_quadtree_
    ST1.m:337-387 (Subpb1, bad-point rule, Subpb2, hole-edge copy, UpdateY,
    duals, early exit) replayed as one captured graph.  The starting
    USubpb1/USubpb2 come from the host driver (LocalICGNQuadtree with its
    NaN fill, then Subpb2), as the reference's loop receives them."""
    api = pkg.api
    from aldic_amd import quadtree
    synth = load("aldic_synth_b", "3d-stereo-aldic_amd/synth.py")
    mesh = load("aldic_mesh_b", "3d-stereo-aldic_amd/mesh.py")
    M = N = 2048
    xc = (M + 1) / 2.0
    uf = lambda X, Y: 0.001 * (X - xc)  # noqa: E731
    vf = lambda X, Y: 0.7 + 0.0 * X  # noqa: E731
    f12, g12, _ = synth.make_pair(M, N, uf, vf, density=0.015, seed=3, bits=12)
    f = np.asfortranarray((f12 - f12.mean()) / f12.std(ddof=1))
    g = np.asfortranarray((g12 - g12.mean()) / g12.std(ddof=1))
    mask = mesh.hole_mask(M, N, (1024.5, 1024.5), 300.0)
    ws = step = 64
    c0, nX, nY = mesh.uniform_grid([1, M], [1, N], ws, step)
    DICmesh = quadtree.MeshSetUp(c0[:, 0].reshape(nY, nX), c0[:, 1].reshape(nY, nX),
                                 {"ImgSize": (M, N), "winstepsize": step})
    DICmesh["elementMinSize"] = 8
    qm, _, _ = quadtree.GenerateQuadtreeMesh(np.zeros(2 * len(c0)), None, mask, DICmesh,
                                             {"winstepsize": step, "winsize": ws})
    coords = np.asfortranarray(qm["coordinatesFEM"])
    n = len(coords)
    ctx = api.Context(0)
    DA = api.DeviceArray
    DICpara = {"winsize": ws, "winstepsize": step, "ImgRefMask": mask}
    Df = pkg.funImgGradient(f, g, mask, ctx=ctx)
    im = api.Images(f, g, mask, Df, ctx=ctx)
    U0 = np.empty(2 * n)
    U0[0::2] = np.round(uf(coords[:, 0], coords[:, 1]))
    U0[1::2] = np.round(vf(coords[:, 0], coords[:, 1]))
    # the temporal IC-GN of the frame on the device
    wsn = np.full(n, ws, np.int32)
    plan = api.Plan(wsn, n, ctx=ctx)
    d_c = DA.from_host(ctx, coords, np.float64)
    d_ws = DA.from_host(ctx, wsn, np.int32)
    d_U0 = DA.from_host(ctx, U0, np.float32)
    d_Uf, d_Ff = DA(ctx, 2 * n, np.float32), DA(ctx, 4 * n, np.float32)
    d_cv, d_z = DA(ctx, n, np.int32), DA(ctx, n, np.float32)
    icgn = lambda: api.local_icgn_dev(im, plan, d_c.ptr, d_U0.ptr, d_ws.ptr, 1e-2,  # noqa: E731
                                      d_Uf.ptr, d_Ff.ptr, None, d_cv.ptr, None, d_z.ptr)
    icgn()
    ctx.synchronize()
    reps = max(args.steps, 5)
    t0 = time.perf_counter()
    for _ in range(reps):
        icgn()
    ctx.synchronize()
    icgn_ms = (time.perf_counter() - t0) / reps * 1e3
    # the ALM loop's start, as the reference's loop receives it
    U1, F1, _, _, _, _, hole = pkg.LocalICGNQuadtree(U0, coords, Df, f, g, DICpara, "GaussNewton",
                                                      1e-2, 1, ctx=ctx)
    mu = 1e-3
    beta = 1e-3 * step * step * mu
    fem = pkg.FEMesh(coords, qm["elementsFEM"], qm["dirichlet"], None, 2, ctx=ctx)
    U2 = fem.solve(beta, mu, U1, F1, 0.0, step)
    U2.reshape(-1, 2)[hole] = U1.reshape(-1, 2)[hole]
    vdual = U2 - U1
    wl = np.full((n, 2), ws, np.int32)
    plan1 = api.Plan(np.asfortranarray(wl), n, rect=True, ctx=ctx)
    d_wl = DA.from_host(ctx, np.asfortranarray(wl), np.int32)
    d_F1 = DA.from_host(ctx, F1, np.float64)
    hole8 = np.zeros(n, np.uint8)
    hole8[hole] = 1
    d_hole = DA.from_host(ctx, hole8, np.uint8)
    d_U1, d_U2, d_vd = (DA(ctx, 2 * n, np.float64) for _ in range(3))
    d_upd, d_state = DA(ctx, 2, np.float64), DA(ctx, 2, np.int32)

    def reload_inputs():
        d_U1.upload(U1)
        d_U2.upload(U2)
        d_vd.upload(vdual)

    def iterations():
        for k in range(2):
            api.alm_iteration_dev(im, plan1, fem, d_c.ptr, d_wl.ptr, mu, beta, 1e-2, 1e-2, True,
                                  k == 0, d_F1.ptr, d_hole.ptr, d_U1.ptr, d_U2.ptr, d_vd.ptr,
                                  d_upd.ptr, d_state.ptr, winstepsize=step, cg_max_iter=512)

    reload_inputs()
    iterations()
    ctx.synchronize()
    state = d_state.to_host()
    graph = ctx.capture(iterations)
    tot = 0.0
    for _ in range(reps):
        reload_inputs()
        ctx.synchronize()
        t0 = time.perf_counter()
        graph.launch(ctx)
        ctx.synchronize()
        tot += time.perf_counter() - t0
    alm_ms = tot / reps * 1e3
    upd = d_upd.to_host()
    graph.close()
    for o in (plan, plan1, fem, im):
        o.close()
    return {"workload": "2048x2048 temporal pair, quadtree mesh of a r=300 hole (64 px "
                        "elements to 8 px), 64 px subsets",
            "nodes": n, "elements": int(len(qm["elementsFEM"])),
            "temporal_icgn_ms": round(icgn_ms, 4),
            "alm_two_iterations_graph_ms": round(alm_ms, 4),
            "frame_ms": round(icgn_ms + alm_ms, 4),
            "frames_per_s": round(1e3 / (icgn_ms + alm_ms), 1),
            "alm_done_after_first": bool(state[0]), "bad_subpb1_nodes": int(state[1]),
            "update_y": [float(upd[0]), float(upd[1])]}


def summary(d, n, steps):
    acc = d["acc"]
    nc = int(acc["ok"].sum())
    s = {"value": round(nc / d["dt"], 1), "ms_per_step": round(d["dt"] * 1e3, 4),
         "converged": nc, "converged_frac": round(nc / n, 5),
         "status_ok_before_bad_point_rule": acc["status_ok"], "bad_points": acc["bad_points"],
         "rms_px": round(acc["rms"], 5), "rms_ok": bool(acc["rms"] <= 0.01),
         "rms_vs_subset_mean_px": round(acc["rms_mean"], 5), "by_subset_size": acc["by_size"],
         "stage_ms_per_step": {k: round(v[0] / steps, 4) for k, v in d["stages"].items()}}
    if "fallback" in d:
        s["fallback"] = d["fallback"]
        s["fallback_rate"] = round(d["fallback"] / n, 5)
    if "prefilter_dt" in d:
        s["prefilter_ms_in_step"] = round(d["prefilter_dt"] * 1e3, 4)
    return s


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-sample", type=int, default=0,
                    help="IC-GN nodes of the CPU baseline (0 = all)")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0: every core of the share")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-alm", action="store_true", help="skip the incremental-sequence frame")
    ap.add_argument("--train-samples", type=int, default=2_000_000)
    ap.add_argument("--net-step", type=float, default=32.0,
                    help="run the network on a regular grid of this spacing (px) and "
                         "interpolate its [F0 U0] to the nodes; 0: at every node.  The "
                         "per-node network is timed beside it under the a-posteriori rule")
    ap.add_argument("--train-trans", type=float, default=2.0,
                    help="training translations within +- this many px: the residual the "
                         "network predicts around the rounded prior (the prior's bilinear "
                         "error on these smooth fields is below 1 px)")
    ap.add_argument("--workloads", default="bulge4096x3000,tilt2048",
                    help="comma-separated; the first is the headline")
    ap.add_argument("--search", action="store_true",
                    help="diagnostic: the integer search alone, direct tiles against FFT blocks, "
                         "at the headline image size, and exit (no metric line)")
    ap.add_argument("--global-step", action="store_true",
                    help="diagnostic: time the rows around the local step on the 2048 x 2048 "
                         "workload (quadtree mesh, Subpb2, NaN fill, raw decode) and exit "
                         "(no metric line)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU: the local matching path is "
                          "measured on one MI355X (see DESIGN.md for how it would shard)",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    pkg = ge.load_package()
    if args.search:
        search_mode(pkg, args)
        return
    if args.global_step:
        global_step_mode(pkg, args)
        return
    from aldic_amd import learned
    # ---- train the network on the device (not timed) ----
    t0 = time.perf_counter()
    tctx = pkg.api.Context(0)
    net = learned.Network(tctx, max_batch=256, seed=0)
    hist = learned.train(net, samples=args.train_samples, batch=256, lr=2e-3, tiles=64, seed=4,
                         log_every=2000, log=progress, trans=args.train_trans)
    ev = learned.evaluate(net, seed=12345, tiles=32, trans=args.train_trans)
    params = net.get_params()
    net.close()
    tctx.close()
    train = {"samples": hist["samples"], "steps": hist["steps"],
             "kept_frac": round(hist["kept"] / max(hist["total"], 1), 5),
             "seconds": round(time.perf_counter() - t0, 1),
             "final_loss": round(hist["loss"][-1], 5) if hist["loss"] else None,
             "translation_range_px": args.train_trans,
             "grid_step_px": args.net_step or None,
             "heldout_rms_u_px": round(ev["rms_u_px"], 4), "heldout_rms_F": round(ev["rms_F"], 5)}
    progress(f"network trained ({train['seconds']} s)")
    names = [w for w in args.workloads.split(",") if w]
    results = {}
    for i, name in enumerate(names):
        t0 = time.perf_counter()
        W = make_workload(name)
        gen_s = time.perf_counter() - t0
        progress(f"workload {name} ready ({gen_s:.1f} s, {len(W['ws'])} nodes)")
        out, roof, Uref, meta = run_workload(pkg, W, params, args, headline=(i == 0))
        results[name] = dict(W=W, out=out, roof=roof, Uref=Uref, meta=meta, gen_s=gen_s)
        if i > 0:
            W["f"] = W["g"] = None  # keep the headline pair only
    head_name = names[0]
    H = results[head_name]
    n = H["meta"]["nodes"]
    headd = H["out"][("learned", HEAD_RULE)]
    value = headd["acc"]["ok"].sum() / headd["dt"]

    def block(res):
        nn = res["meta"]["nodes"]
        blk = {}
        for (start, rule), d in res["out"].items():
            blk[f"{start}_start_{rule}_rule"] = summary(d, nn, args.steps)
            ref = res["out"].get((start, "reference"))
            if rule in RULES and rule != "reference" and ref is not None:
                # how far the rule's answer lies from the reference rule's
                # on the nodes both count as converged (px, per node |dU|)
                both = d["acc"]["ok"] & ref["acc"]["ok"]
                du = np.abs(d["U"] - ref["U"]).reshape(-1, 2)[both].max(axis=1)
                blk[f"{start}_start_{rule}_rule"]["du_vs_reference_rule_px"] = {
                    "nodes": int(both.sum()), "mean": float(du.mean()),
                    "p99": float(np.percentile(du, 99)), "max": float(du.max())}
        for rule in (*RULES, "posterior_bspline"):
            a = res["out"][("learned", rule)]
            b = res["out"][("search", rule)]
            blk[f"learned_vs_search_{rule}_rule"] = round(
                (a["acc"]["ok"].sum() / a["dt"]) / (b["acc"]["ok"].sum() / b["dt"]), 3)
        blk["nodes"] = nn
        blk["search_points"] = res["meta"]["search_points"]
        blk["icgn_launch_shapes"] = res["meta"]["shapes"]
        blk["gen_s"] = round(res["gen_s"], 1)
        return blk

    spec = H["W"]["spec"]
    net_where = (f"on a {args.net_step:g} px grid interpolated to the nodes" if args.net_step
                 else "at every node")
    line = {
        "metric": "converged IC-GN subsets/s at <=0.01 px RMS, learned vs FFT init, 1 MI355X",
        "value": round(float(value), 1),
        "unit": "converged subsets/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(headd["dt"] * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": (f"large stereo pair {spec['M']}x{spec['N']} 12-bit speckle "
                                f"(0.01 spots/px^2, seed 5), 25 px disparity + 15 px Gaussian "
                                f"bulge (sigma 600 px), hole r=300, {n} nodes with 8-64 px "
                                f"subsets; learned start (CNN on a 256 px sparse-search prior, "
                                f"{net_where}, gate + search fallback) vs FFT ZNCC search start (r=50, "
                                f"{H['meta']['grid'][0]}x{H['meta']['grid'][1]} grid); IC-GN "
                                f"tol 1e-3"),
                   "init": "learned", "stop": f"a-posteriori step size < {TOL_DP} px",
                   "nodes": n},
        "converged": int(headd["acc"]["ok"].sum()),
        "converged_frac": round(float(headd["acc"]["ok"].sum()) / n, 5),
        "rms_px": round(headd["acc"]["rms"], 5),
        "rms_ok": bool(headd["acc"]["rms"] <= 0.01),
        "roofline": H["roof"],
        "network": train,
    }
    line[head_name] = block(H)
    for name in names[1:]:
        line[name] = block(results[name])
    line["learned_vs_fft"] = line[head_name][f"learned_vs_search_{HEAD_RULE}_rule"]
    bsp = H["out"][("learned", "posterior_bspline")]
    line["interp_bspline"] = {
        "note": "option, not the reference's interpolant: the interpolating cubic B-spline on "
                "prefiltered coefficients; learned start, a-posteriori rule, prefilter included",
        "value": round(float(bsp["acc"]["ok"].sum() / bsp["dt"]), 1),
        "ms_per_step": round(bsp["dt"] * 1e3, 4),
        "prefilter_ms": round(bsp["prefilter_dt"] * 1e3, 4),
        "rms_px": round(bsp["acc"]["rms"], 5), "rms_ok": bool(bsp["acc"]["rms"] <= 0.01),
        "learned_vs_search": line[head_name]["learned_vs_search_posterior_bspline_rule"]}
    if ("learned", "posterior_per_node") in H["out"]:
        pn = H["out"][("learned", "posterior_per_node")]
        line["network_at_every_node"] = {
            "note": "the same step with the network at every node instead of on its grid",
            "value": round(float(pn["acc"]["ok"].sum() / pn["dt"]), 1),
            "ms_per_step": round(pn["dt"] * 1e3, 4), "rms_px": round(pn["acc"]["rms"], 5),
            "fallback": pn["fallback"], "grid_speedup": round(pn["dt"] / headd["dt"], 3)}
    if not args.no_alm:
        progress("incremental-sequence frame: temporal IC-GN + ALM graph")
        line["seq2048_alm"] = alm_block(pkg, args)
    if not args.no_cpu:
        threads = args.cpu_threads or cpu_threads()
        progress(f"CPU baseline (FP64 oracle, {threads} threads)")
        diag, line["cpu_baseline"] = cpu_baseline(H["W"], args.cpu_sample, threads, H["Uref"])
        line["oracle_sample"] = diag
    else:
        line["cpu_baseline"] = None
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
