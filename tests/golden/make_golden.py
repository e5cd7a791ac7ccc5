"""Writes tests/golden/oracle_small.npz: inputs and FP64 oracle outputs of
the local matching path on seeded small cases.

These are regression fixtures of the oracle (oracle/), not vectors of the
reference (the reference ships none and MATLAB cannot run here): the CPU
tests check that the oracle still reproduces them exactly, the GPU tests
check the library against them with the tolerances of the parity tests.

Run from the repository root (after `make -C oracle`):

    python tests/golden/make_golden.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.dirname(HERE))

import fem_oracle  # noqa: E402
import natural_oracle  # noqa: E402
import oracle  # noqa: E402
from _cases import quadtree_mesh  # noqa: E402

import __graft_entry__ as ge  # noqa: E402


def build():
    synth = ge._load_file("aldic_synth_golden", os.path.join(ge.PKG_DIR, "synth.py"))
    M, N = 96, 80
    uf = lambda X, Y: 1.3 + 0.004 * (X - M / 2) + 0.3 * np.sin(2 * np.pi * Y / 64.0)
    vf = lambda X, Y: -0.7 + 0.003 * (Y - N / 2)
    f8, g8, _ = synth.make_pair(M, N, uf, vf, density=0.02, seed=11)
    # stored as float32, used as the FP64 values of those floats
    f = ((f8 - f8.mean()) / f8.std(ddof=1)).astype(np.float32)
    g = ((g8 - g8.mean()) / g8.std(ddof=1)).astype(np.float32)
    f64, g64 = np.asfortranarray(f, np.float64), np.asfortranarray(g, np.float64)
    rng = np.random.default_rng(11)
    out = {"f": f, "g": g}
    # ba_interp2_spline: 64 points, interior, edges and outside-clamped
    X = np.concatenate([rng.uniform(1, N, 48), [1.0, N, 0.3, N + 2.5, 5.0, 40.25,
                                                 12.0, 70.5, 33.0, 2.0, 79.9, 60.0, 1.5,
                                                 7.75, 20.0, 50.0]])
    Y = np.concatenate([rng.uniform(1, M, 48), [1.0, M, 4.0, 50.0, 0.2, M + 3.0, 95.5, 1.25,
                                                 48.0, 2.0, 9.0, 96.0, 30.5, 88.0, 44.0,
                                                 17.5]])
    out["interp_X"], out["interp_Y"] = X, Y
    out["interp_out"] = oracle.interp_spline(f64, X, Y)
    # funIntegerSearch, ws 16, step 12, radius 6
    s = oracle.integer_search(f64, g64, [1, M], [1, N], 16, 12, [6, 6, 6, 6])
    for k in ("x", "y", "u", "v", "ccmax", "ipeak_row", "ipeak_col"):
        out["search_" + k] = s[k]
    # IC-GN from the rounded true displacement, 20 nodes, winsize 20
    im = oracle.Images(f64, g64)
    xs, ys = np.meshgrid(np.arange(16.0, 81.0, 16.0), np.arange(16.0, 65.0, 16.0),
                         indexing="ij")
    coords = np.asfortranarray(np.stack([xs.ravel(order="F"), ys.ravel(order="F")], 1))
    n = len(coords)
    U0 = np.empty(2 * n)
    U0[0::2] = np.round(uf(coords[:, 0], coords[:, 1]))
    U0[1::2] = np.round(vf(coords[:, 0], coords[:, 1]))
    ws = np.full(n, 20, np.int32)
    U, F, H, conv, hole, z = oracle.local_icgn(im, coords, U0, ws, 1e-3)
    out.update(icgn_coords=coords, icgn_U0=U0, icgn_ws=ws, icgn_U=U, icgn_F=F, icgn_conv=conv,
               icgn_zncc=z)
    # Subpb1 from the IC-GN result with a dual term
    vdual = rng.normal(scale=0.02, size=2 * n)
    udual = np.zeros(4 * n)
    wl = np.full((n, 2), 20, np.int32)
    U1, conv1 = oracle.subpb1(im, coords, U, F, udual, vdual, wl, 1e-3, 1e-2)
    out.update(subpb1_vdual=vdual, subpb1_U=U1, subpb1_conv=conv1)
    # Subpb2 on a one-level quadtree mesh with hanging nodes
    qc, qe, qb = quadtree_mesh(5, 4, 4.0, refine=[(1, 1), (3, 2)])
    nq = len(qc)
    Uq = np.empty(2 * nq)
    Uq[0::2] = 0.4 + 0.01 * qc[:, 0] + 0.2 * np.sin(qc[:, 1] / 7.0)
    Uq[1::2] = -0.3 + 0.02 * qc[:, 1]
    Fq = rng.normal(scale=0.01, size=4 * nq)
    Uh = fem_oracle.subpb2(qc, qe, qb, None, 2, 2e-3, 1e-3, Uq, Fq, 0.0, 8.0)
    out.update(fem_coords=qc, fem_elements=qe, fem_dirichlet=qb, fem_U=Uq, fem_F=Fq,
               fem_Uhat=Uh)
    # the NaN fill: 40 nodes on a grid, a mask with two regions
    gx, gy = np.meshgrid(np.arange(6.0, 90.0, 12.0), np.arange(6.0, 78.0, 12.0), indexing="ij")
    nc = np.asfortranarray(np.stack([gx.ravel(order="F"), gy.ravel(order="F")], 1))
    Un = np.empty(2 * len(nc))
    Un[0::2] = 0.2 + 0.01 * nc[:, 0] + np.sin(nc[:, 1] / 15.0)
    Un[1::2] = -0.1 + 0.004 * nc[:, 1]
    bad = rng.choice(len(nc), 7, replace=False)
    Un[2 * bad] = np.nan
    Un[2 * bad + 1] = np.nan
    mask = np.ones((M, N))
    mask[50:53, :] = 0
    out.update(fill_coords=nc, fill_U=Un, fill_mask=mask.astype(np.uint8),
               fill_out=natural_oracle.nan_fill(nc, Un, None, mask))
    return out


if __name__ == "__main__":
    out = build()
    np.savez_compressed(os.path.join(HERE, "oracle_small.npz"), **out)
    print("wrote", os.path.join(HERE, "oracle_small.npz"),
          {k: np.shape(v) for k, v in out.items()})
