"""Cell-tree restatement of GenerateQuadtreeMesh's refinement
(func_quadtree/GenerateQuadtreeMesh.m:19-104 with QrefineR.m), for
checking the library's edge-based version geometrically.

TEST INFRASTRUCTURE ONLY: imported by tests/ (never by the product package).

Independent of the product's arithmetic: cells are (level, i, j) boxes of a
tree over the uniform mesh; a round marks the leaves whose bounding box sees
both mask values (sliced directly from the mask, no summed-area table) and
whose side exceeds elementMinSize, then closes the marking with the rules
QrefineR applies through its edge flags, stated on cells:
  * a leaf that would end with a midpoint on three or four of its sides is
    refined (a side has a midpoint when the cell across it, at the leaf's
    level, is refined now or already was);
  * a leaf next to a finer cell that is refined now is refined (no side may
    carry two levels of hanging nodes).
Rounds repeat until nothing is marked.  The leaves that keep less than half
of the mask over their box are dropped (funMarkInside.m).  The result is a
set of element boxes and, per side, whether it carries a hanging node:
QrefineR's mesh up to node numbering.

Pinned by tests/test_quadtree.py: on a mask with no edge the mesh is the
uniform one; one masked pixel refines its cells to elementMinSize and the
closure is checked by hand on a small case (every leaf has at most one
hanging node per side, and no leaf sits next to a leaf two levels finer).
"""
from __future__ import annotations

import numpy as np


class Tree:
    def __init__(self, x0, y0, H, nx, ny):
        self.x0, self.y0, self.H = float(x0), float(y0), float(H)
        self.nx, self.ny = nx, ny
        self.leaves = {(0, i, j) for i in range(nx) for j in range(ny)}

    def box(self, c):
        l, i, j = c
        h = self.H / 2 ** l
        return self.x0 + i * h, self.y0 + j * h, self.x0 + (i + 1) * h, self.y0 + (j + 1) * h

    def covering_leaf(self, l, i, j):
        """The leaf that covers cell (l, i, j), if that cell or a parent of
        it is a leaf; None if the region is subdivided finer (or outside)."""
        while l >= 0:
            if (l, i, j) in self.leaves:
                return (l, i, j)
            l, i, j = l - 1, i // 2, j // 2
        return None

    def inside(self, l, i, j):
        n = 2 ** l
        return 0 <= i < self.nx * n and 0 <= j < self.ny * n

    def finer_along(self, c, side):
        """Leaves finer than c touching side (0 left, 1 right, 2 bottom, 3 top)."""
        l, i, j = c
        di, dj = [(-1, 0), (1, 0), (0, -1), (0, 1)][side]
        ni, nj = i + di, j + dj
        if not self.inside(l, ni, nj) or self.covering_leaf(l, ni, nj) is not None:
            return []
        out = []
        stack = [(l, ni, nj)]
        while stack:
            ll, a, b = stack.pop()
            for ca in (0, 1):
                for cb in (0, 1):
                    ch = (ll + 1, 2 * a + ca, 2 * b + cb)
                    # keep children that touch the shared side
                    if side == 0 and ca != 1 or side == 1 and ca != 0:
                        continue
                    if side == 2 and cb != 1 or side == 3 and cb != 0:
                        continue
                    if ch in self.leaves:
                        out.append(ch)
                    else:
                        stack.append(ch)
        return out

    def refine(self, marked):
        M = set(marked)
        changed = True
        while changed:
            changed = False
            for c in list(self.leaves - M):
                l, i, j = c
                cnt, force = 0, False
                for side, (di, dj) in enumerate([(-1, 0), (1, 0), (0, -1), (0, 1)]):
                    ni, nj = i + di, j + dj
                    if not self.inside(l, ni, nj):
                        continue
                    cov = self.covering_leaf(l, ni, nj)
                    if cov is not None:
                        if cov == (l, ni, nj) and cov in M:
                            cnt += 1
                    else:  # finer across: a hanging node already
                        cnt += 1
                        if any(f in M for f in self.finer_along(c, side)):
                            force = True
                if cnt >= 3 or force:
                    M.add(c)
                    changed = True
        for c in M:
            l, i, j = c
            self.leaves.remove(c)
            for a in (0, 1):
                for b in (0, 1):
                    self.leaves.add((l + 1, 2 * i + a, 2 * j + b))

    def hanging(self, c):
        """Per side (right, top, left, bottom = element columns 5..8): the
        side's midpoint when finer leaves lie across it, else None."""
        x0, y0, x1, y1 = self.box(c)
        res = []
        for side in (1, 3, 0, 2):
            if self.finer_along(c, side):
                res.append({1: (x1, (y0 + y1) / 2), 3: ((x0 + x1) / 2, y1),
                            0: (x0, (y0 + y1) / 2), 2: ((x0 + x1) / 2, y0)}[side])
            else:
                res.append(None)
        return res


def mark_edge(tree, mask, min_size):
    out = []
    for c in tree.leaves:
        x0, y0, x1, y1 = (int(round(v)) for v in tree.box(c))
        sub = mask[x0 - 1:x1, y0 - 1:y1]
        if sub.max() - sub.min() > 0 and min(x1 - x0, y1 - y0) > min_size:
            out.append(c)
    return out


def generate(mask, x0, y0, H, nx, ny, min_size):
    """Leaves after refinement and removal: {box: [hanging per column 5..8]}."""
    mask = np.asarray(mask, dtype=np.float64)
    t = Tree(x0, y0, H, nx, ny)
    while True:
        m = mark_edge(t, mask, min_size)
        if not m:
            break
        t.refine(m)
    out = {}
    for c in t.leaves:
        bx = t.box(c)
        xa, ya, xb, yb = (int(round(v)) for v in bx)
        if mask[xa - 1:xb, ya - 1:yb].sum() < 0.5 * (xb - xa) * (yb - ya):
            continue
        out[bx] = t.hanging(c)
    return out, t
