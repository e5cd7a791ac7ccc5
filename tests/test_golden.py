"""The committed golden vectors (tests/golden/oracle_small.npz, written by
tests/golden/make_golden.py from the FP64 oracle on seeded small cases):
the oracle must still reproduce them exactly, and the library's host code
(the NaN fill) within 1e-9.  The GPU side is tests/test_gpu_golden.py."""
import os

import fem_oracle
import natural_oracle
import numpy as np
import pytest

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def _img(gold):
    return (np.asfortranarray(gold["f"], np.float64), np.asfortranarray(gold["g"], np.float64))


def test_oracle_reproduces_interp_and_search(orc, gold):
    f, g = _img(gold)
    assert np.array_equal(orc.interp_spline(f, gold["interp_X"], gold["interp_Y"]),
                          gold["interp_out"])
    M, N = f.shape
    s = orc.integer_search(f, g, [1, M], [1, N], 16, 12, [6, 6, 6, 6])
    for k in ("x", "y", "u", "v", "ccmax", "ipeak_row", "ipeak_col"):
        assert np.array_equal(s[k], gold["search_" + k]), k


def test_oracle_reproduces_icgn_and_subpb1(orc, gold):
    f, g = _img(gold)
    im = orc.Images(f, g)
    U, F, _, conv, _, z = orc.local_icgn(im, gold["icgn_coords"], gold["icgn_U0"],
                                         gold["icgn_ws"], 1e-3)
    assert np.array_equal(conv, gold["icgn_conv"])
    assert np.array_equal(U, gold["icgn_U"]) and np.array_equal(F, gold["icgn_F"])
    assert np.array_equal(z, gold["icgn_zncc"])
    n = len(gold["icgn_ws"])
    U1, conv1 = orc.subpb1(im, gold["icgn_coords"], U, F, np.zeros(4 * n), gold["subpb1_vdual"],
                           np.full((n, 2), 20, np.int32), 1e-3, 1e-2)
    assert np.array_equal(conv1, gold["subpb1_conv"]) and np.array_equal(U1, gold["subpb1_U"])


def test_oracle_reproduces_subpb2_and_fill(gold):
    Uh = fem_oracle.subpb2(gold["fem_coords"], gold["fem_elements"], gold["fem_dirichlet"], None,
                           2, 2e-3, 1e-3, gold["fem_U"], gold["fem_F"], 0.0, 8.0)
    assert np.allclose(Uh, gold["fem_Uhat"], rtol=0, atol=1e-12)
    out = natural_oracle.nan_fill(gold["fill_coords"], gold["fill_U"], None, gold["fill_mask"])
    assert np.array_equal(out, gold["fill_out"])


def test_library_nan_fill_matches_golden(pkg, gold):
    got = pkg.nan_fill(gold["fill_coords"], gold["fill_U"], None, gold["fill_mask"])
    assert np.max(np.abs(got - gold["fill_out"])) < 1e-9
