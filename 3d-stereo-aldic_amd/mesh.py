"""Node lists for the solvers.

``uniform_grid`` reproduces the node layout of funIntegerSearch +
MeshSetUp (func/funIntegerSearch.m:106-112, func/MeshSetUp.m:23-32): node
centres x = ii + ws/2 over XList = gridx(1):step:gridx(2)-ws-step, x-index
fastest.  ``hole_nodes`` builds the node set of the benchmark's stereo
workload: nodes on a regular grid outside a circular hole, each with its
own subset size chosen by its distance to the hole edge (8 px subsets at
the edge doubling up to 64 px away from it), sorted by size so that one
launch per size class sees neighbouring nodes together.
"""
from __future__ import annotations

import numpy as np


def grid_list(a, b, ws, step):
    return np.arange(a, b - ws - step + 1, step, dtype=np.int64)


def uniform_grid(gridx, gridy, winsize, winstepsize):
    XList = grid_list(gridx[0], gridx[1], winsize, winstepsize)
    YList = grid_list(gridy[0], gridy[1], winsize, winstepsize)
    xx, yy = np.meshgrid(XList + winsize / 2.0, YList + winsize / 2.0, indexing="xy")
    coords = np.stack([xx.reshape(-1), yy.reshape(-1)], axis=1)  # x fastest
    return np.asfortranarray(coords), len(XList), len(YList)


def hole_nodes(M, N, centre, radius, step=8, ws_min=8, ws_max=64, margin=None,
               extra=(0, 0, 0, 0)):
    """Returns (coords n x 2, winsize n) for nodes outside the hole whose
    subsets (grown up to 1.3x for masked pixels) stay inside the image.
    ``extra`` widens the margin at the low-x, high-x, low-y and high-y
    edges, e.g. by the largest disparity towards that edge, so that the
    deformed subsets stay inside the image as well (the ROI a user sets
    for a stereo pair)."""
    if margin is None:
        margin = int(0.65 * ws_max) + 8
    xs = np.arange(margin + extra[0], M - margin - extra[1] + 1, step, dtype=np.float64)
    ys = np.arange(margin + extra[2], N - margin - extra[3] + 1, step, dtype=np.float64)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    d = np.hypot(X - centre[0], Y - centre[1]) - radius
    keep = d > 0
    X, Y, d = X[keep], Y[keep], d[keep]
    ws = np.full(X.shape, ws_max, dtype=np.int32)
    lev = np.clip(np.floor(np.log2(np.maximum(d, 1.0))), np.log2(ws_min), np.log2(ws_max))
    ws = (2 ** lev).astype(np.int32)
    order = np.lexsort((X, Y, ws))
    coords = np.stack([X[order], Y[order]], axis=1)
    return np.asfortranarray(coords), ws[order].copy()


def hole_mask(M, N, centre, radius):
    i = np.arange(1, M + 1, dtype=np.float64)[:, None]
    j = np.arange(1, N + 1, dtype=np.float64)[None, :]
    return np.asfortranarray(((i - centre[0]) ** 2 + (j - centre[1]) ** 2 > radius ** 2)
                             .astype(np.float64))
