"""Seeded inputs shared by the CPU and GPU tests (built once per process)."""
from __future__ import annotations

import functools
import os

import numpy as np

import __graft_entry__ as ge

_synth = ge._load_file("aldic_synth_cases", os.path.join(ge.PKG_DIR, "synth.py"))
_mesh = ge._load_file("aldic_mesh_cases", os.path.join(ge.PKG_DIR, "mesh.py"))


def normalize(a):
    a = np.asarray(a, dtype=np.float64)
    return np.asfortranarray((a - a.mean()) / a.std(ddof=1))


@functools.lru_cache(maxsize=None)
def c1_pair():
    """512 x 512 uint8 speckle, 0.02 spots/px^2, seed 1, C1 field; normalised."""
    uf, vf = _synth.field_c1()
    f8, g8, _ = _synth.make_pair(512, 512, uf, vf, density=0.02, seed=1)
    return normalize(f8), normalize(g8), uf, vf


@functools.lru_cache(maxsize=None)
def small_pair(M=160, N=144, seed=5, shift=(0.0, 0.0), strain=0.0):
    ux, uy = shift
    uf = lambda X, Y: ux + strain * (X - M / 2)
    vf = lambda X, Y: uy - 0.5 * strain * (Y - N / 2)
    f8, g8, _ = _synth.make_pair(M, N, uf, vf, density=0.02, seed=seed)
    return normalize(f8), normalize(g8), uf, vf


def c1_grid():
    return _mesh.uniform_grid([1, 512], [1, 512], 30, 16)


def hole_mask(M, N, centre, radius):
    return _mesh.hole_mask(M, N, centre, radius)


def quadtree_mesh(nx, ny, h, refine=(), x0=10.0, y0=10.0):
    """A one-level quadtree Q4 mesh: an nx x ny grid of 2h cells with the
    cells in ``refine`` ((i, j) pairs) split into four h cells.  Returns
    coordinatesFEM (n x 2), elementsFEM (ne x 8, 1-based, node order
    (xmin, ymin), (xmax, ymin), (xmax, ymax), (xmin, ymax) as the reference
    reorders them, GenerateQuadtreeMesh.m:36-47; columns 5..8 hold the
    hanging node on edge 2-3, 3-4, 4-1, 1-2 or 0) and the outer boundary
    nodes (1-based)."""
    refine = set(refine)
    ids = {}
    pts = []

    def node(x, y):
        k = (round(x * 4), round(y * 4))
        if k not in ids:
            ids[k] = len(pts) + 1
            pts.append((x, y))
        return ids[k]

    H = 2 * h
    # all corner nodes first (the coarse grid), then refinement nodes
    for j in range(ny + 1):
        for i in range(nx + 1):
            node(x0 + i * H, y0 + j * H)
    elems = []
    for j in range(ny):
        for i in range(nx):
            xa, ya = x0 + i * H, y0 + j * H
            if (i, j) in refine:
                for (a, b) in ((0, 0), (1, 0), (1, 1), (0, 1)):
                    xs, ys = xa + a * h, ya + b * h
                    elems.append([node(xs, ys), node(xs + h, ys), node(xs + h, ys + h),
                                  node(xs, ys + h), 0, 0, 0, 0])
            else:
                e = [node(xa, ya), node(xa + H, ya), node(xa + H, ya + H), node(xa, ya + H),
                     0, 0, 0, 0]
                # hanging nodes: the edge midpoint of a refined neighbour
                if (i + 1, j) in refine:
                    e[4] = node(xa + H, ya + h)
                if (i, j + 1) in refine:
                    e[5] = node(xa + h, ya + H)
                if (i - 1, j) in refine:
                    e[6] = node(xa, ya + h)
                if (i, j - 1) in refine:
                    e[7] = node(xa + h, ya)
                elems.append(e)
    coords = np.array(pts, dtype=np.float64)
    xmax, ymax = x0 + nx * H, y0 + ny * H
    bnd = np.flatnonzero((np.isclose(coords[:, 0], x0) | np.isclose(coords[:, 0], xmax) |
                          np.isclose(coords[:, 1], y0) | np.isclose(coords[:, 1], ymax))) + 1
    return np.asfortranarray(coords), np.array(elems, dtype=np.int64), bnd
