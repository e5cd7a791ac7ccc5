"""aldic_params: the DICpara fields the path reads as a plain C struct
(CPU, host code of the library): setDICparaDefaults.m's values,
validateDICpara.m's rules, and the search region of IntegerSearchQuadtree.m:
52-90 / funIntegerSearch.m:12, which must equal the host driver's choice."""
import numpy as np
import pytest


def test_defaults(pkg):
    from aldic_amd._capi import AldicParams, lib
    import ctypes as C
    p = AldicParams()
    assert lib().aldic_params_default(C.byref(p)) == 0
    assert (p.winsize, p.winstepsize, p.winsizeMin) == (0, 0, 8)
    assert (p.DICIncOrNot, p.UseGlobal, p.GaussPtOrder, p.ClusterNo, p.InitFFTSearchMethod) == \
        (0, 1, 2, 1, 1)
    assert (p.nStereoSearchDistance, p.nTemporalSearchDistance, p.nNewFFTSearchDistance) == (0, 0, 0)
    assert (p.tol_stereo, p.tol_temporal, p.mu, p.beta, p.alpha, p.tol2) == \
        (1e-3, 1e-2, 1e-3, 0.0, 0.0, 1e-2)
    assert p.fixSearchDistanceOrNot == 0


@pytest.mark.parametrize("bad", [
    {"winsize": 31}, {"winsize": -2}, {"winstepsize": 24}, {"winstepsize": 16, "winsizeMin": 12},
    {"winstepsize": 8, "winsizeMin": 16}, {"DICIncOrNot": 2}, {"InitFFTSearchMethod": 4},
    {"ClusterNo": 0}, {"UseGlobal": 3}])
def test_validation_rejects(pkg, bad):
    with pytest.raises(pkg.AldicError):
        pkg.api.validate_dicpara(bad)


def test_validation_accepts(pkg):
    for ok in ({"winsize": 32, "winstepsize": 16}, {"winsize": 30, "winstepsize": 16, "winsizeMin": 16},
               {}, {"NewFFTSearchDistance": [5, 6, 7, 8]}, {"winsize": 64, "winstepsize": 64}):
        pkg.api.validate_dicpara(ok)
    with pytest.raises(ValueError):
        pkg.api.validate_dicpara({"StereoSearchDistance": [1, 2, 3]})


def test_search_region_equals_the_host_driver(pkg, monkeypatch):
    """The C rule and IntegerSearchQuadtree's (host Python) pick the same
    radius for every case of the reference's branches."""
    api = pkg.api
    seen = {}

    def fake_search(f, g, r, gridx, gridy, ws, step, tempNoOfInitPt=0, ctx=None):
        seen["r"] = np.asarray(r).copy()
        x = np.ones((2, 2))
        return x, x, 0 * x, 0 * x, {"max": x, "A": []}

    monkeypatch.setattr(api, "funIntegerSearch", fake_search)
    base = {"gridxyROIRange": {"gridx": [1, 100], "gridy": [1, 90]}, "winsize": 20,
            "winstepsize": 16, "NewFFTSearchDistance": 7}
    cases = [(-1, {"StereoSearchDistance": 30}), (-1, {"StereoSearchDistance": [40, 12]}),
             (-1, {"StereoSearchDistance": []}), (2, {"TemporalSearchDistance": 9}), (3, {}),
             (-1, {"StereoSearchDistance": 0}), (4, {"TemporalSearchDistance": [3, 4]}),
             (-1, {"StereoSearchDistance": [9, 8, 7, 6]})]
    for seq, extra in cases:
        D = {**base, **extra}
        api.IntegerSearchQuadtree(None, None, None, D, seq)
        r = np.asarray(seen["r"]).reshape(-1)
        want = list(np.concatenate([r, r])[:4]) if r.size != 4 else list(r)
        assert api.search_region(D, seq) == [int(v) for v in want], (seq, extra)
    for seq in (0, 1):
        D = {k: v for k, v in base.items() if k != "NewFFTSearchDistance"}
        with pytest.raises(pkg.AldicError):
            api.search_region(D, seq)
