"""The augmented-Lagrangian loop on the device (aldic_alm_iteration_dev)
against the same loop driven from the host through the reference's
MATLAB-level functions (TemporalMatch_quadtree_ST1.m:337-387: Subpb1Quadtree
with its bad-point rule and NaN fill, Subpb2Quadtree, the hole-edge copy,
UpdateY, the duals, the early exit), on one temporal frame of a 1024 x 1024
pair (0.1% x-stretch, 0.7 px y-translation) over a quadtree mesh refined at
a centred hole.

The device loop gives a bad Subpb1 node (Subpb1Quadtree.m:167-182) the
global estimate it started from, where the reference fills it by natural-
neighbour interpolation (host code).  Against a host loop that applies the
same rule the two run the same kernels on the same data: USubpb1, USubpb2,
the duals and the UpdateY norms must agree to 1e-9 (the FP64 conjugate
gradients stop at a relative residual of 1e-13).  Against the host loop
with the reference's fill, the nodes that are good in Subpb1 must agree to
within the stated bound: the fill's values reach them only through the
global step, whose free rows do not see the Dirichlet rows' values
(Subpb2Quadtree.m:319), and on this mesh every bad node is a Dirichlet
(hole-edge or border) node, so they agree exactly.  The two iterations of a frame are recorded as one HIP graph;
every replay, with host work and fresh inputs between them, must give the
eager result bit for bit."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _norm(a):
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray((a - a.mean()) / a.std(ddof=1))


@pytest.fixture(scope="module")
def frame(pkg, synth, mesh):
    from aldic_amd import quadtree
    M = N = 1024
    xc = (M + 1) / 2.0
    uf = lambda X, Y: 0.001 * (X - xc)  # noqa: E731
    vf = lambda X, Y: 0.7 + 0.0 * X  # noqa: E731
    f8, g8, _ = synth.make_pair(M, N, uf, vf, density=0.015, seed=31, bits=12)
    f, g = _norm(f8), _norm(g8)
    mask = mesh.hole_mask(M, N, (512.5, 512.5), 120.0)
    ws = step = 32
    c0, nX, nY = mesh.uniform_grid([1, M], [1, N], ws, step)
    DICmesh = quadtree.MeshSetUp(c0[:, 0].reshape(nY, nX), c0[:, 1].reshape(nY, nX),
                                 {"ImgSize": (M, N), "winstepsize": step})
    DICmesh["elementMinSize"] = 8
    qm, _, _ = quadtree.GenerateQuadtreeMesh(np.zeros(2 * len(c0)), None, mask, DICmesh,
                                             {"winstepsize": step, "winsize": ws})
    coords = np.asfortranarray(qm["coordinatesFEM"])
    n = len(coords)
    DICpara = {"winsize": ws, "winstepsize": step, "ImgRefMask": mask}
    Df = pkg.funImgGradient(f, g, mask)
    U0 = np.empty(2 * n)
    U0[0::2] = np.round(uf(coords[:, 0], coords[:, 1]))
    U0[1::2] = np.round(vf(coords[:, 0], coords[:, 1]))
    U1, F1, _, _, _, _, hole = pkg.LocalICGNQuadtree(U0, coords, Df, f, g, DICpara,
                                                      "GaussNewton", 1e-2, 1)
    return dict(f=f, g=g, mask=mask, coords=coords, qm=qm, n=n, DICpara=DICpara, Df=Df, U1=U1,
                F1=F1, hole=hole, step=step, ws=ws)


def _host_loop(pkg, fr, fem, mu, beta, tol2, check, reference_fill):
    """The reference's loop from the host (the bookkeeping of
    TemporalMatch_quadtree_ST1.m:333-387); reference_fill False gives the
    bad Subpb1 nodes the global estimate instead of the natural fill."""
    U1, F1, hole = fr["U1"], fr["F1"], fr["hole"]
    U2 = fem.solve(beta, mu, U1, F1, 0.0, fr["step"])
    U2.reshape(-1, 2)[hole] = U1.reshape(-1, 2)[hole]
    vdual = U2 - U1
    udual = np.zeros(4 * fr["n"])
    U1p, U2p = U1.copy(), U2.copy()
    start = (U1.copy(), U2.copy(), vdual.copy())
    upd = (1e4, np.nan)
    bad_all = np.zeros(fr["n"], bool)
    for al_step in (2, 3):
        U1, _, _, conv, badnum = pkg.Subpb1Quadtree(
            U2, F1, udual, vdual, fr["coords"], fr["Df"], fr["f"], fr["g"], mu, beta, None,
            al_step, fr["DICpara"], "GaussNewton", 1e-2, fill_nans=reference_fill)
        bad, _ = pkg.subpb1_bad_points(conv)
        nan = np.isnan(U1[0::2]) | np.isnan(U1[1::2])
        bad_all[bad] = True
        bad_all |= nan
        if not reference_fill:
            U1 = U1.copy()
            U1.reshape(-1, 2)[nan] = U2.reshape(-1, 2)[nan]
        U2 = fem.solve(beta, mu, U1, F1, 0.0, fr["step"])
        U2.reshape(-1, 2)[hole] = U1.reshape(-1, 2)[hole]
        if check:
            upd = (np.linalg.norm(U2p - U2) / np.sqrt(U2p.size),
                   np.linalg.norm(U1p - U1) / np.sqrt(U1p.size))
        U1p, U2p = U1, U2
        vdual = U2 - U1
        if check and (upd[0] < tol2 or upd[1] < tol2):
            break
    return start, (U1, U2, vdual, upd, al_step, bad_all)


@pytest.mark.parametrize("tol2", [1e-2, 1.0])
def test_device_alm_matches_the_host_loop(pkg, frame, tol2):
    api = pkg.api
    fr = frame
    n = fr["n"]
    ctx = api.Context(0)
    DA = api.DeviceArray
    mu = 1e-3
    beta = 1e-3 * fr["step"] ** 2 * mu
    fem = pkg.FEMesh(fr["coords"], fr["qm"]["elementsFEM"], fr["qm"]["dirichlet"], None, 2,
                     ctx=ctx)
    start, (U1h, U2h, vdh, updh, last_step, bad) = _host_loop(pkg, fr, fem, mu, beta, tol2, True,
                                                              False)
    _, (U1r, U2r, _, _, _, _) = _host_loop(pkg, fr, fem, mu, beta, tol2, True, True)
    print(f"{n} nodes, host loop ended at ALSolveStep {last_step}, UpdateY {updh[0]:.3e}, "
          f"UpdateY2 {updh[1]:.3e}, bad Subpb1 nodes {bad.sum()}", flush=True)
    # the planes the host driver builds (its Df included)
    im = api.Images(fr["f"], fr["g"], fr["mask"], fr["Df"], ctx=ctx)
    wl = np.full((n, 2), fr["ws"], np.int32)
    plan = api.Plan(np.asfortranarray(wl), n, rect=True, ctx=ctx)
    d_c = DA.from_host(ctx, fr["coords"], np.float64)
    d_wl = DA.from_host(ctx, np.asfortranarray(wl), np.int32)
    d_F1 = DA.from_host(ctx, fr["F1"], np.float64)
    hole8 = np.zeros(n, np.uint8)
    hole8[fr["hole"]] = 1
    d_hole = DA.from_host(ctx, hole8, np.uint8)
    d_U1, d_U2, d_vd = DA(ctx, 2 * n, np.float64), DA(ctx, 2 * n, np.float64), DA(ctx, 2 * n,
                                                                                  np.float64)
    d_upd, d_state = DA(ctx, 2, np.float64), DA(ctx, 2, np.int32)

    def load():
        d_U1.upload(start[0])
        d_U2.upload(start[1])
        d_vd.upload(start[2])

    def frame_iterations():
        for k in range(2):
            api.alm_iteration_dev(im, plan, fem, d_c.ptr, d_wl.ptr, mu, beta, 1e-2, tol2, True,
                                  k == 0, d_F1.ptr, d_hole.ptr, d_U1.ptr, d_U2.ptr, d_vd.ptr,
                                  d_upd.ptr, d_state.ptr, winstepsize=fr["step"], cg_max_iter=512)

    load()
    frame_iterations()
    ctx.synchronize()
    eager = [a.to_host() for a in (d_U1, d_U2, d_vd, d_upd, d_state)]
    state = eager[4]
    print(f"device loop: done {state[0]}, bad nodes of its last iteration {state[1]}, "
          f"UpdateY {eager[3][0]:.3e}, UpdateY2 {eager[3][1]:.3e}", flush=True)
    assert state[0] == (1 if last_step == 2 else state[0])
    for got, want, name in ((eager[0], U1h, "USubpb1"), (eager[1], U2h, "USubpb2"),
                            (eager[2], vdh, "vdual")):
        d = np.abs(got - want).max()
        print(f"  {name}: max |device - host| {d:.2e}", flush=True)
        assert d <= 1e-9, name
    assert np.allclose(eager[3], updh, rtol=1e-9, atol=1e-12)
    # against the loop with the reference's natural-neighbour fill, on the
    # nodes Subpb1 solved
    good = ~bad
    dref = np.abs(eager[1] - U2r).reshape(-1, 2)[good].max()
    print(f"  USubpb2 vs the reference-fill loop on {good.sum()} good nodes: max {dref:.2e} px",
          flush=True)
    assert dref <= 0.05
    # the frame's two iterations as one graph, replayed with fresh inputs
    g = ctx.capture(frame_iterations)
    for rep in range(3):
        load()
        _ = np.linalg.norm(start[1])  # host work between replays
        g.launch(ctx)
        ctx.synchronize()
        for got, want in zip((a.to_host() for a in (d_U1, d_U2, d_vd, d_upd, d_state)), eager):
            assert np.array_equal(got, want, equal_nan=True), rep
    g.close()
    plan.close()
    fem.close()
    im.close()
