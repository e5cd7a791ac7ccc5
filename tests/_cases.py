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
