"""The drivers' NaN fill (LocalICGNQuadtree.m:200-325, Subpb1Quadtree.m:
189-254) and its interpolant, scatteredInterpolant(..., 'natural',
'nearest').  Host code only: these run without a GPU.

The oracle (oracle/natural_oracle.py) is pinned first by properties that
fix Sibson's interpolant independently of any implementation:
  * linear precision: a linear field is reproduced exactly inside the hull;
  * symmetric configurations with known weights (square centre 1/4, regular
    hexagon centre 1/6);
  * the stolen areas against a raster count of nearest sites.
The library's C++ is then checked against the oracle on random, gridded
(cocircular, the quadtree case) and duplicated sites, inside, on and
outside the hull, to 1e-9 of the value range.  Parity with MATLAB's own
builtins (scatteredInterpolant, bwconncomp) is unpinned: no reference
fixture holds their outputs.
"""
import numpy as np
import pytest

import natural_oracle as no


def _lin(x, y):
    return 3.0 + 0.5 * x - 0.25 * y


# ---------------------------------------------------------------------------
# oracle pins
# ---------------------------------------------------------------------------
def test_oracle_square_and_hexagon_weights():
    I = no.Natural([0, 1, 0, 1], [0, 0, 1, 1], [1.0, 2.0, 3.0, 4.0])
    w = I.weights(0.5, 0.5)
    assert sorted(w) == [0, 1, 2, 3]
    assert np.allclose(list(w.values()), 0.25, atol=1e-12)
    t = np.arange(6) * np.pi / 3
    I = no.Natural(np.cos(t), np.sin(t), np.arange(6.0))
    w = I.weights(0.0, 0.0)
    assert len(w) == 6 and np.allclose(list(w.values()), 1 / 6, atol=1e-12)
    assert abs(I.eval1(0.0, 0.0)[0] - 2.5) < 1e-12


def test_oracle_linear_precision():
    rng = np.random.default_rng(1)
    x, y = rng.uniform(0, 100, 60), rng.uniform(0, 100, 60)
    I = no.Natural(x, y, _lin(x, y))
    for qx, qy in rng.uniform(30, 70, (25, 2)):
        assert abs(I.eval1(qx, qy)[0] - _lin(qx, qy)) < 1e-9


def test_oracle_weights_match_raster_count():
    rng = np.random.default_rng(2)
    x, y = rng.uniform(0, 10, 14), rng.uniform(0, 10, 14)
    I = no.Natural(x, y, np.zeros(14))
    qx, qy = 5.1, 4.7
    w = I.weights(qx, qy)
    h = 0.004
    gx, gy = np.meshgrid(np.arange(-6, 16, h), np.arange(-6, 16, h), indexing="ij")
    d_new = (gx - qx) ** 2 + (gy - qy) ** 2
    d_old = (gx[..., None] - x) ** 2 + (gy[..., None] - y) ** 2
    inq = d_new < d_old.min(axis=-1)
    owner = np.argmin(d_old[inq], axis=-1)
    cnt = np.bincount(owner, minlength=14) / inq.sum()
    for i in range(14):
        assert abs(cnt[i] - w.get(i, 0.0)) < 3e-3, (i, cnt[i], w.get(i, 0.0))


def test_oracle_outside_and_edge():
    I = no.Natural([0, 2, 0, 2], [0, 0, 2, 2], [1.0, 5.0, 3.0, 7.0])
    assert I.eval1(-1.0, -0.2)[0] == 1.0  # nearest
    assert abs(I.eval1(1.0, 0.0)[0] - 3.0) < 1e-12  # on the hull edge: linear
    assert abs(I.eval1(0.5, 0.0)[0] - 2.0) < 1e-12


# ---------------------------------------------------------------------------
# library vs oracle
# ---------------------------------------------------------------------------
def _queries(rng, x, y):
    lo, hi = min(x.min(), y.min()), max(x.max(), y.max())
    q = [rng.uniform(lo - 10, hi + 10, (150, 2))]  # inside and outside
    q.append(np.stack([x[:20], y[:20]], axis=1))  # at sites
    I = no.Natural(x, y, np.zeros_like(x))
    h = I.hull
    for k in range(len(h)):  # on hull edges
        a, b = h[k], h[(k + 1) % len(h)]
        t = rng.uniform(0.1, 0.9)
        q.append(np.array([[x[a] + t * (x[b] - x[a]), y[a] + t * (y[b] - y[a])]]))
    return np.concatenate(q)


def _site_sets():
    rng = np.random.default_rng(7)
    out = []
    x, y = rng.uniform(0, 100, 200), rng.uniform(0, 100, 200)
    out.append(("uniform", x, y))
    # a quadtree-like grid: 16 px cells, one quadrant refined to 8 px
    g = np.arange(16.0, 257.0, 16.0)
    X, Y = np.meshgrid(g, g, indexing="ij")
    f = np.arange(136.0, 257.0, 8.0)
    Xf, Yf = np.meshgrid(f, f, indexing="ij")
    P = np.unique(np.concatenate([np.stack([X.ravel(), Y.ravel()], 1),
                                  np.stack([Xf.ravel(), Yf.ravel()], 1)]), axis=0)
    keep = rng.random(len(P)) > 0.1  # holes: NaN nodes
    out.append(("quadtree-grid", P[keep, 0], P[keep, 1]))
    # clustered
    c = rng.normal(0, 1, (120, 2)) * np.repeat([[3.0], [20.0]], 60, axis=0) + 50
    out.append(("clustered", c[:, 0], c[:, 1]))
    return out


@pytest.mark.parametrize("name", ["uniform", "quadtree-grid", "clustered"])
def test_natural_interp_matches_oracle(pkg, name):
    rng = np.random.default_rng(11)
    _, x, y = [s for s in _site_sets() if s[0] == name][0]
    v = np.stack([np.sin(x / 17.0) + 0.01 * y, np.cos(y / 23.0) - 0.02 * x], axis=1)
    q = _queries(rng, x, y)
    if name == "quadtree-grid":  # cell centres: four cocircular sites
        q = np.concatenate([q, np.stack([np.arange(24.0, 240, 16), np.arange(24.0, 240, 16)], 1)])
    got = pkg.scattered_natural(x, y, v, q[:, 0], q[:, 1])
    want = no.Natural(x, y, v)(q[:, 0], q[:, 1])
    scale = np.abs(v).max()
    assert np.max(np.abs(got - want)) <= 1e-9 * scale, np.max(np.abs(got - want))
    # linear precision through the library
    got = pkg.scattered_natural(x, y, _lin(x, y), q[:, 0], q[:, 1])
    inside = [no.Natural(x, y, np.zeros_like(x))._locate(a, b)[0] >= 0 for a, b in q]
    assert np.max(np.abs(got[inside] - _lin(q[inside, 0], q[inside, 1]))) < 1e-9


def test_natural_interp_duplicates_and_errors(pkg):
    x = np.array([0.0, 1.0, 0.0, 1.0, 1.0])
    y = np.array([0.0, 0.0, 1.0, 1.0, 1.0])
    v = np.array([1.0, 2.0, 3.0, 4.0, 6.0])  # (1, 1) twice: mean 5
    got = pkg.scattered_natural(x, y, v, [1.0, 0.5], [1.0, 0.5])
    assert got[0] == 5.0 and abs(got[1] - (1 + 2 + 3 + 5) / 4) < 1e-12
    with pytest.raises(pkg.AldicError):
        pkg.scattered_natural([0.0, 1.0, 2.0], [0.0, 1.0, 2.0], [1.0, 2.0, 3.0], [0.5], [0.5])
    with pytest.raises(pkg.AldicError):
        pkg.scattered_natural([0.0, 1.0], [0.0, 1.0], [1.0, 2.0], [0.5], [0.5])


def _fill_case(seed):
    rng = np.random.default_rng(seed)
    M, N = 150, 130
    g = np.arange(8.0, 150.0, 10.0)
    h = np.arange(8.0, 130.0, 10.0)
    X, Y = np.meshgrid(g, h, indexing="ij")
    coords = np.asfortranarray(np.stack([X.ravel(order="F"), Y.ravel(order="F")], 1))
    n = len(coords)
    U = np.empty(2 * n)
    U[0::2] = 0.3 + 0.01 * coords[:, 0] + np.sin(coords[:, 1] / 20)
    U[1::2] = -0.2 + 0.005 * coords[:, 1]
    mask = np.ones((M, N))
    mask[70:74, :] = 0  # two regions
    mask[20:40, 20:40] = 0  # a hole
    bad = rng.choice(n, 25, replace=False)
    bad = np.union1d(bad, np.arange(0, n, 15)[:5])  # a few along the x = 8 edge
    U[2 * bad] = np.nan
    U[2 * bad + 1] = np.nan
    return coords, U, mask, bad


@pytest.mark.parametrize("with_F", [False, True])
def test_nan_fill_matches_oracle(pkg, with_F):
    coords, U, mask, bad = _fill_case(3)
    n = len(coords)
    F = np.zeros(4 * n)
    if with_F:
        F[4 * 7] = np.nan  # a thrown subset: the global pass runs
    got = pkg.nan_fill(coords, U, F if with_F else None, mask)
    want = no.nan_fill(coords, U, F if with_F else None, mask)
    assert np.all(np.isfinite(got))
    assert np.max(np.abs(got - want)) < 1e-9
    good = np.setdiff1d(np.arange(n), bad)
    assert np.array_equal(got[2 * good], U[2 * good])  # valid nodes keep their values
    # no NaN and no NaN in F: nothing changes
    clean = np.nan_to_num(U)
    assert np.array_equal(pkg.nan_fill(coords, clean, None, mask), clean)


def test_nan_fill_region_rules(pkg):
    """A region with fewer than three valid nodes is skipped by the region
    pass; the global pass then fills from every valid node."""
    coords, U, mask, _ = _fill_case(5)
    n = len(coords)
    # a region (x >= 74) whose only valid nodes are two
    right = np.flatnonzero(coords[:, 0] >= 74)
    U2 = U.copy()
    U2[2 * right[2:]] = np.nan
    U2[2 * right[2:] + 1] = np.nan
    got = pkg.nan_fill(coords, U2, None, mask)
    want = no.nan_fill(coords, U2, None, mask)
    assert np.max(np.abs(got - want)) < 1e-9
    with pytest.raises(pkg.AldicError):  # every node NaN
        pkg.nan_fill(coords, np.full(2 * n, np.nan), None, mask)


def test_nan_fill_diagonal_touch_is_one_region(pkg):
    """Two mask blocks that touch only at a corner are one 8-connected
    region: its NaN nodes are filled from both blocks' valid nodes and not
    from the separate third block (whose values are far off)."""
    M, N = 120, 120
    mask = np.zeros((M, N))
    mask[0:50, 0:50] = 1
    mask[50:100, 50:100] = 1  # touches the first block at one corner
    mask[105:120, 0:40] = 1   # separate
    pts = [(x, y) for x in np.arange(5.0, 50, 9) for y in np.arange(5.0, 50, 9)]
    pts += [(x, y) for x in np.arange(54.0, 100, 9) for y in np.arange(54.0, 100, 9)]
    pts += [(x, y) for x in (108.0, 116.0) for y in np.arange(5.0, 40, 8)]
    coords = np.asfortranarray(np.array(pts, dtype=float))
    n = len(coords)
    U = np.empty(2 * n)
    U[0::2] = 0.02 * coords[:, 0] + 0.01 * coords[:, 1]
    U[1::2] = -0.01 * coords[:, 0]
    U[0::2][coords[:, 0] > 104] += 50.0
    second = np.flatnonzero((coords[:, 0] > 50) & (coords[:, 0] < 100))
    bad = second[::3]
    U[2 * bad] = np.nan
    U[2 * bad + 1] = np.nan
    got = pkg.nan_fill(coords, U, None, mask)
    want = no.nan_fill(coords, U, None, mask)
    assert np.max(np.abs(got - want)) < 1e-9
    assert np.max(np.abs(got[2 * bad])) < 10.0  # not pulled toward the third block


def test_natural_interp_large_void_many_queries(pkg):
    """A grid with a large empty disc: queries at its rim and inside it have
    Voronoi cells reaching across the void; enough queries (300) that the
    library splits them over threads.  Eight are checked against the oracle,
    all 300 by linear precision."""
    g = np.arange(0.0, 321.0, 16.0)
    X, Y = np.meshgrid(g, g, indexing="ij")
    x, y = X.ravel(), Y.ravel()
    keep = np.hypot(x - 160, y - 160) > 90
    x, y = x[keep], y[keep]
    v = np.stack([np.sin(x / 31.0) + 0.01 * y, np.cos(y / 17.0)], axis=1)
    rng = np.random.default_rng(5)
    t = rng.uniform(0, 2 * np.pi, 300)
    rr = np.concatenate([rng.uniform(80, 100, 150), rng.uniform(0, 80, 150)])
    q = np.stack([160 + rr * np.cos(t), 160 + rr * np.sin(t)], 1)
    got = pkg.scattered_natural(x, y, v, q[:, 0], q[:, 1])
    pick = np.r_[0:4, 150:154]
    want = no.Natural(x, y, v)(q[pick, 0], q[pick, 1])
    assert np.max(np.abs(got[pick] - want)) <= 1e-9 * np.abs(v).max()
    got = pkg.scattered_natural(x, y, _lin(x, y), q[:, 0], q[:, 1])
    assert np.max(np.abs(got - _lin(q[:, 0], q[:, 1]))) < 1e-9
