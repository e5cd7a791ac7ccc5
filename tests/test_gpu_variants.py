"""GPU parity of every kernel instance the timed path can run, against the
FP64 oracle, through the C ABI's device-resident entry points.

* Every launch shape (waves per subset x pixels per lane) of every size
  class, for IC-GN and for Subpb1, forced with aldic_plan_set_variant: the
  compile-time-size instances (every subset at the class's full side) and
  the run-time-size ones (mixed sides within the class), each through the
  one-entry kernels (aldic_local_icgn_dev) and the list-walk kernels
  (aldic_local_icgn_list_dev, count read on the device).
* The autotuner: its picks are legal launch shapes and the solve it leaves
  the plan with matches the oracle.
* Evaluation-capped runs (aldic_local_icgn_f0_dev, max_iter) and starts
  with a gradient guess F0, against the oracle run with the same cap/start.
* The mask branch: a constructed mask whose grown window and 4-connected
  component fix the Hessian (funICGNQuadtree.m:66-103), and Subpb1 on
  hole-edge nodes with the mask on both sides (funICGN_Subpb1quadtree.m:
  158-203).

Tolerances as tests/test_gpu_parity.py: |dU| <= 1e-3 px on subsets both
converge, statuses equal up to the one-evaluation rule on <= 5%, |dF| <=
1e-5 at equal counts, H within 1e-5 of the node's largest entry.  The
one-entry and list-walk kernels run the same arithmetic: their outputs are
compared bit for bit.
"""
import numpy as np
import pytest

from _cases import hole_mask, small_pair
from test_gpu_parity import _cmp_icgn

pytestmark = pytest.mark.gpu

CLASS_SIDES = (11, 17, 33, 65)
# winsize values (side - 1) inside each class: the full side and a mixture
MIXED = {0: [8, 10], 1: [12, 14, 16], 2: [18, 24, 30, 32], 3: [34, 48, 60, 64]}


def _say(msg):
    print(msg, flush=True)


def _nodes(ws, n, M, N, seed):
    rng = np.random.default_rng(seed)
    w = np.asarray(ws, np.int32)
    h = int(w.max()) // 2 + 8
    cx = rng.integers(h, M - h + 1, n).astype(float)
    cy = rng.integers(h, N - h + 1, n).astype(float)
    return np.asfortranarray(np.stack([cx, cy], 1))


class _Dev:
    """Device copies of a node set and its outputs (IC-GN)."""

    def __init__(self, api, ctx, coords, U0, ws, F0=None):
        DA = api.DeviceArray
        n = coords.shape[0]
        self.n = n
        self.coords = DA.from_host(ctx, coords, np.float64)
        self.U0 = DA.from_host(ctx, U0, np.float32)
        self.F0 = None if F0 is None else DA.from_host(ctx, F0, np.float32)
        self.ws = DA.from_host(ctx, ws, np.int32)
        self.U = DA(ctx, 2 * n, np.float32)
        self.F = DA(ctx, 4 * n, np.float32)
        self.H = DA(ctx, 21 * n, np.float32)
        self.conv = DA(ctx, n, np.int32)
        self.hole = DA(ctx, n, np.uint8)
        self.z = DA(ctx, n, np.float32)
        self.list = DA.from_host(ctx, np.arange(n, dtype=np.int32), np.int32)
        self.count = DA.from_host(ctx, np.array([n], np.int32), np.int32)

    def outs(self):
        return (self.U.ptr, self.F.ptr, self.H.ptr, self.conv.ptr, self.hole.ptr, self.z.ptr)

    def host(self):
        n = self.n
        return (self.U.to_host().astype(np.float64), self.F.to_host().astype(np.float64),
                self.H.to_host().reshape(21, n).T.astype(np.float64), self.conv.to_host(),
                self.hole.to_host(), self.z.to_host().astype(np.float64))


def _icgn_case(ws_values, n, seed):
    f, g, uf, vf = small_pair(M=256, N=240, seed=10, shift=(0.6, -0.2), strain=0.002)
    rng = np.random.default_rng(seed)
    ws = rng.choice(np.asarray(ws_values, np.int32), size=n).astype(np.int32)
    coords = _nodes(ws, n, 256, 240, seed + 1)
    U0 = np.zeros(2 * n)
    U0[0::2] = np.round(uf(coords[:, 0], coords[:, 1]))
    U0[1::2] = np.round(vf(coords[:, 0], coords[:, 1]))
    return f, g, coords, U0, ws


@pytest.fixture(scope="module")
def images(pkg):
    f, g, _, _ = small_pair(M=256, N=240, seed=10, shift=(0.6, -0.2), strain=0.002)
    im = pkg.Images(f, g)
    yield im
    im.close()


@pytest.mark.parametrize("cls", [0, 1, 2, 3])
@pytest.mark.parametrize("full", [True, False])
def test_icgn_every_launch_shape(pkg, orc, images, cls, full):
    api = pkg.api
    ctx = images.ctx
    values = [CLASS_SIDES[cls] - 1] if full else MIXED[cls]
    f, g, coords, U0, ws = _icgn_case(values, 72, 100 + cls)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8)
    d = _Dev(api, ctx, coords, U0, ws)
    plan = api.Plan(ws, len(ws), ctx=ctx)
    for var in range(3):
        plan.set_variant(0, cls, var)
        api.local_icgn_dev(images, plan, d.coords.ptr, d.U0.ptr, d.ws.ptr, 1e-3, *d.outs())
        got = d.host()
        _say(f"class {cls} ({'full' if full else 'mixed'}) variant {var}: one-entry kernel")
        _cmp_icgn(got, want, conv_tol_frac=0.05)
        # the list-walk instance over every node: the same arithmetic
        for a in (d.U, d.F, d.H, d.conv, d.z):
            a.upload(np.zeros(a.count))
        api.local_icgn_list_dev(images, plan, d.coords.ptr, d.U0.ptr, None, d.ws.ptr,
                                d.list.ptr, d.count.ptr, 1e-3, *d.outs())
        walk = d.host()
        for k, name in enumerate(("U", "F", "H", "conv", "hole", "zncc")):
            assert np.array_equal(walk[k], got[k], equal_nan=True), (cls, var, name)
    plan.close()


@pytest.mark.parametrize("cls", [0, 1, 2, 3])
def test_subpb1_every_launch_shape(pkg, orc, images, cls):
    api = pkg.api
    ctx = images.ctx
    DA = api.DeviceArray
    f, g, coords, U0, ws = _icgn_case(MIXED[cls] + [CLASS_SIDES[cls] - 1], 64, 200 + cls)
    n = len(ws)
    base = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8)
    UOld = np.nan_to_num(base[0])
    FOld = np.nan_to_num(base[1])
    rng = np.random.default_rng(cls)
    vdual = rng.normal(scale=0.02, size=2 * n)
    udual = np.zeros(4 * n)
    # rectangular windows: winsize_List(:,2) one step smaller where possible
    wl = np.stack([ws, np.maximum(ws - 2, 8)], 1).astype(np.int32)
    Uo, convo = orc.subpb1(orc.Images(f, g), coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2,
                           nthreads=8)
    plan = api.Plan(np.asfortranarray(wl), n, rect=True, ctx=ctx)
    d_c = DA.from_host(ctx, coords, np.float64)
    d_uo = DA.from_host(ctx, UOld, np.float32)
    d_fo = DA.from_host(ctx, FOld, np.float32)
    d_vd = DA.from_host(ctx, vdual, np.float32)
    d_wl = DA.from_host(ctx, np.asfortranarray(wl), np.int32)
    d_U = DA(ctx, 2 * n, np.float32)
    d_conv = DA(ctx, n, np.int32)
    for var in range(3):
        plan.set_variant(1, cls, var)
        api.subpb1_dev(images, plan, d_c.ptr, d_uo.ptr, d_fo.ptr, d_vd.ptr, d_wl.ptr, 1e-3,
                       1e-2, d_U.ptr, d_conv.ptr)
        U = d_U.to_host().astype(np.float64)
        conv = d_conv.to_host()
        ok = (convo >= 1) & (convo <= 100) & (conv >= 1) & (conv <= 100)
        du = np.abs(U - Uo).reshape(-1, 2)[ok].max(initial=0)
        _say(f"Subpb1 class {cls} variant {var}: {ok.sum()} of {n} converged in both, "
             f"max |dU| {du:.2e}, counts differ {np.sum(conv != convo)}")
        assert np.array_equal(conv == 0, convo == 0)
        assert np.mean(conv != convo) <= 0.05
        assert ok.sum() >= 0.8 * n and du <= 1e-3
    plan.close()


def test_autotuner_picks_and_outputs(pkg, orc, images):
    api = pkg.api
    ctx = images.ctx
    vals = sum(([CLASS_SIDES[c] - 1] + MIXED[c] for c in range(4)), [])
    f, g, coords, U0, ws = _icgn_case(vals, 160, 300)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8)
    d = _Dev(api, ctx, coords, U0, ws)
    plan = api.Plan(ws, len(ws), ctx=ctx)
    chosen = api.local_icgn_autotune_dev(images, plan, d.coords.ptr, d.U0.ptr, d.ws.ptr, 1e-3,
                                         d.U.ptr, d.F.ptr, d.conv.ptr, d.z.ptr, reps=2)
    _say(f"autotuner picks {list(chosen)}")
    assert chosen.shape == (4,) and np.all((chosen >= 0) & (chosen < 3))
    api.local_icgn_dev(images, plan, d.coords.ptr, d.U0.ptr, d.ws.ptr, 1e-3, *d.outs())
    _cmp_icgn(d.host(), want, conv_tol_frac=0.05)
    plan.close()


@pytest.mark.parametrize("cap", [1, 2, 3])
def test_evaluation_capped_runs(pkg, orc, images, cap):
    """max_iter = cap stops the loop after cap + 1 evaluations; a subset not
    converged by then reports 101 with the warp after cap + 1 updates, in
    both (the warps the oracle's stopping-rule pin reads)."""
    api = pkg.api
    ctx = images.ctx
    vals = [16, 30, 32, 64]
    f, g, coords, U0, ws = _icgn_case(vals, 96, 400)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8, max_iter=cap)
    d = _Dev(api, ctx, coords, U0, ws)
    plan = api.Plan(ws, len(ws), ctx=ctx)
    api.local_icgn_f0_dev(images, plan, d.coords.ptr, d.U0.ptr, None, d.ws.ptr, 1e-3, *d.outs()[:2],
                          d_H21=d.H.ptr, d_conv=d.conv.ptr, d_hole=d.hole.ptr, d_zncc=d.z.ptr,
                          max_iter=cap)
    got = d.host()
    conv, convo = got[3], want[3]
    _say(f"cap {cap}: status 101 in {np.sum(conv == 101)} (oracle {np.sum(convo == 101)})")
    assert np.sum(convo == 101) > 10  # the cap stops most subsets
    both = (conv == 101) & (convo == 101)
    assert np.mean(conv != convo) <= 0.05
    dU = np.abs(got[0] - want[0]).reshape(-1, 2)[both].max()
    dF = np.abs(got[1] - want[1]).reshape(-1, 4)[both].max()
    _say(f"  capped warps: max |dU| {dU:.2e} px, max |dF| {dF:.2e}")
    assert dU <= 1e-3 and dF <= 1e-4
    plan.close()


def test_gradient_start(pkg, orc, images):
    """P0 = [F0 U0] (the learned start's entry point) against the oracle
    started the same way."""
    api = pkg.api
    ctx = images.ctx
    f, g, coords, U0, ws = _icgn_case([10, 16, 30, 32, 48, 64], 96, 500)
    n = len(ws)
    rng = np.random.default_rng(5)
    F0 = rng.uniform(-0.01, 0.01, 4 * n)
    U0 = U0 + rng.uniform(-0.4, 0.4, 2 * n)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8, F0=F0)
    d = _Dev(api, ctx, coords, U0, ws, F0=F0)
    plan = api.Plan(ws, n, ctx=ctx)
    api.local_icgn_f0_dev(images, plan, d.coords.ptr, d.U0.ptr, d.F0.ptr, d.ws.ptr, 1e-3,
                          *d.outs()[:2], d_H21=d.H.ptr, d_conv=d.conv.ptr, d_hole=d.hole.ptr,
                          d_zncc=d.z.ptr)
    _cmp_icgn(d.host(), want, conv_tol_frac=0.05)
    plan.close()


def test_mask_branch_window_growth(pkg, orc):
    """The constructed case of the oracle's window-growth pin: a 3 px zero
    stripe through a 21 px window grows it to 23 px and cuts the component;
    ImgDef = ImgRef converges at the first evaluation, so H is returned.
    The workgroup kernel against the oracle."""
    f, _, _, _ = small_pair(seed=21)
    M, N = f.shape
    mask = np.ones((M, N))
    mask[:, 74:77] = 0  # MATLAB columns 75..77
    coords = np.array([[80.0, 70.0], [60.0, 68.0], [100.0, 72.0]])
    im = pkg.Images(f, f, mask)
    got = pkg.local_icgn_raw(im, coords, np.zeros(6), 20, 1e-3)
    want = orc.local_icgn(orc.Images(f, f, mask), coords, np.zeros(6),
                          np.full(3, 20, np.int32), 1e-3)
    assert np.all(want[3] == 1) and np.all(got[3] == 1), (got[3], want[3])
    _cmp_icgn(got, want, conv_tol_frac=0.0)
    im.close()


def test_subpb1_mask_branch(pkg, orc):
    """Subpb1 on nodes around a hole with the mask on both sides: the
    masked recompute of f, its gradients and H (funICGN_Subpb1quadtree.m:
    158-203) in the workgroup kernel against the oracle."""
    f, g, _, _ = small_pair(M=200, N=180, seed=7, shift=(0.8, -0.6))
    M, N = f.shape
    mask = hole_mask(M, N, (100, 90), 28)
    fm = np.asfortranarray(f * mask)
    gm = np.asfortranarray(g * hole_mask(M, N, (100.8, 89.4), 28))
    xs = np.arange(40, 161, 8, dtype=float)
    ys = np.arange(40, 141, 8, dtype=float)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    coords = np.asfortranarray(np.stack([X.reshape(-1), Y.reshape(-1)], 1))
    r = np.hypot(coords[:, 0] - 100, coords[:, 1] - 90)
    coords = np.asfortranarray(coords[(r > 20) & (r < 48)])
    n = coords.shape[0]
    oim = orc.Images(fm, gm, mask)
    base = orc.local_icgn(oim, coords, np.tile([0.8, -0.6], n), 16, 1e-3)
    UOld = np.where(np.isnan(base[0]), np.tile([0.8, -0.6], n), base[0])
    FOld = np.nan_to_num(base[1])
    rng = np.random.default_rng(8)
    vdual = rng.normal(scale=0.02, size=2 * n)
    udual = np.zeros(4 * n)
    wl = np.full((n, 2), 16, np.int32)
    Uo, convo = orc.subpb1(oim, coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2)
    # how many of these take the mask branch: some pixel of the window masked
    h = 8
    masked = np.array([(mask[int(x) - 1 - h:int(x) + h, int(y) - 1 - h:int(y) + h] == 0).any()
                       for x, y in coords])
    assert masked.sum() >= 10
    im = pkg.Images(fm, gm, mask)
    U, conv = pkg.subpb1_raw(im, coords, UOld, FOld, udual, vdual, wl, 1e-3, 1e-2)
    ok = (convo >= 1) & (convo <= 100) & (conv >= 1) & (conv <= 100)
    du = np.abs(U - Uo).reshape(-1, 2)[ok]
    _say(f"Subpb1 masked: {n} nodes ({masked.sum()} with masked pixels), {ok.sum()} converged "
         f"in both ({(ok & masked).sum()} masked), max |dU| {du.max():.2e}, statuses "
         f"{dict(zip(*np.unique(convo, return_counts=True)))}")
    for code in (0, 102, 103):
        assert np.array_equal(conv == code, convo == code), code
    assert np.mean(conv != convo) <= 0.05
    assert (ok & masked).sum() >= 5
    assert du.max() <= 1e-3
    im.close()


@pytest.mark.parametrize("rule,tol_dp", [(1, 1e-3), (1, 1e-2), (2, 1e-3)])
def test_step_size_rule(pkg, orc, images, rule, tol_dp):
    """The plan's step-size stopping rules (aldic_plan_set_stopping, rule 1
    and its a-posteriori form, rule 2) on every size class and the
    workgroup kernel (> 65 px), from the search start, against the oracle
    run with the same rule; switching the plan back gives the reference
    rule's results."""
    api = pkg.api
    ctx = images.ctx
    f, g, coords, U0, ws = _icgn_case([8, 10, 16, 30, 32, 48, 64, 80], 120, 600)
    n = len(ws)
    d = _Dev(api, ctx, coords, U0, ws)
    plan = api.Plan(ws, n, ctx=ctx)
    plan.set_stopping(rule, tol_dp)
    want = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8, stop_rule=rule,
                          tol_dp=tol_dp)
    api.local_icgn_dev(images, plan, d.coords.ptr, d.U0.ptr, d.ws.ptr, 1e-3, *d.outs())
    got = d.host()
    conv, convo = got[3], want[3]
    ok = (conv >= 1) & (conv <= 100) & (convo >= 1) & (convo <= 100)
    du = np.abs(got[0] - want[0]).reshape(-1, 2)[ok].max()
    same = ok & (conv == convo)
    dF = np.abs(got[1] - want[1]).reshape(-1, 4)[same].max()
    _say(f"stopping rule {rule}, {tol_dp}: {ok.sum()} of {n} converged in both, counts differ "
         f"{np.sum(conv != convo)}, mean evaluations {conv[ok].mean():.2f} (oracle "
         f"{convo[ok].mean():.2f}), max |dU| {du:.2e}, max |dF| same count {dF:.2e}")
    assert ok.sum() >= 0.9 * n and np.mean(conv != convo) <= 0.05
    assert du <= 1e-3 and dF <= 1e-4
    for code in (-1, 102, 103):
        assert np.array_equal(conv == code, convo == code)
    # back to the reference rule: the reference's results
    plan.set_stopping(0)
    api.local_icgn_dev(images, plan, d.coords.ptr, d.U0.ptr, d.ws.ptr, 1e-3, *d.outs())
    ref = orc.local_icgn(orc.Images(f, g), coords, U0, ws, 1e-3, nthreads=8)
    _cmp_icgn(d.host(), ref, conv_tol_frac=0.05)
    plan.close()
