"""The library against the committed golden vectors of the FP64 oracle
(tests/golden/oracle_small.npz), with the tolerances of
tests/test_gpu_parity.py: FP64 interpolation 1e-12; integer search peaks
bit-exact, sub-pixel u, v within 1e-3 + 1e-9 (the fit is rounded to 1e-3);
IC-GN and Subpb1 |dU| <= 1e-3 px on the subsets both converge, |dF| <= 1e-5
where both stop after the same number of evaluations; Subpb2 1e-9
relative."""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "oracle_small.npz")


@pytest.fixture(scope="module")
def gold():
    with np.load(GOLD, allow_pickle=False) as d:
        return {k: d[k] for k in d.files}


def test_golden_local_path(pkg, gold):
    f = np.asfortranarray(gold["f"], np.float64)
    g = np.asfortranarray(gold["g"], np.float64)
    M, N = f.shape
    got = pkg.ba_interp2_spline(f, gold["interp_X"], gold["interp_Y"])
    assert np.max(np.abs(got - gold["interp_out"])) <= 1e-12 * np.abs(gold["interp_out"]).max()
    x, y, u, v, cc, pr, pc = pkg.funIntegerSearch(f, g, [6, 6], [1, M], [1, N], 16, 12,
                                                  return_peaks=True)
    assert np.array_equal(x, gold["search_x"]) and np.array_equal(y, gold["search_y"])
    assert np.array_equal(pr, gold["search_ipeak_row"])
    assert np.array_equal(pc, gold["search_ipeak_col"])
    assert np.max(np.abs(u - gold["search_u"])) <= 1e-3 + 1e-9
    assert np.max(np.abs(v - gold["search_v"])) <= 1e-3 + 1e-9
    im = pkg.Images(f, g)
    U, F, H, conv, hole, z = pkg.local_icgn_raw(im, gold["icgn_coords"], gold["icgn_U0"],
                                                gold["icgn_ws"], 1e-3)
    cw = gold["icgn_conv"]
    ok = (conv >= 1) & (conv <= 100) & (cw >= 1) & (cw <= 100)
    assert ok.sum() >= 18 and np.mean(conv != cw) <= 0.1
    assert np.abs(U - gold["icgn_U"]).reshape(-1, 2)[ok].max() <= 1e-3
    same = ok & (conv == cw)
    assert np.abs(F - gold["icgn_F"]).reshape(-1, 4)[same].max() <= 1e-5
    n = len(gold["icgn_ws"])
    U1, conv1 = pkg.subpb1_raw(im, gold["icgn_coords"], gold["icgn_U"], gold["icgn_F"],
                               np.zeros(4 * n), gold["subpb1_vdual"],
                               np.full((n, 2), 20, np.int32), 1e-3, 1e-2)
    ok1 = (conv1 >= 1) & (conv1 <= 100) & (gold["subpb1_conv"] >= 1)
    assert ok1.sum() >= 18
    assert np.abs(U1 - gold["subpb1_U"]).reshape(-1, 2)[ok1].max() <= 1e-3
    im.close()


def test_golden_subpb2(pkg, gold):
    mesh = pkg.FEMesh(gold["fem_coords"], gold["fem_elements"], gold["fem_dirichlet"], None, 2)
    Uh = mesh.solve(2e-3, 1e-3, gold["fem_U"], gold["fem_F"], 0.0, 8.0)
    assert np.max(np.abs(Uh - gold["fem_Uhat"])) <= 1e-9 * np.abs(gold["fem_U"]).max()
    mesh.close()
