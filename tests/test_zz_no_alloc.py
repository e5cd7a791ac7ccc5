"""The no-allocation integer search (aldic_integer_search_noalloc_dev)
against the allocating one (aldic_integer_search_dev) through the C ABI:
the same kernels on the same inputs, so every output must agree bit for
bit, on two calls in a row with the second over a smaller grid (the shared
per-GPU scratch keeps the first call's contents); a grid whose scratch
exceeds the array is an argument error, not a fault."""
import numpy as np
import pytest

from _cases import c1_pair

pytestmark = pytest.mark.gpu


def _grid(M, N, gridx, gridy, ws, step):
    nX = len(range(gridx[0], gridx[1] - ws - step + 1, step))
    nY = len(range(gridy[0], gridy[1] - ws - step + 1, step))
    return nX, nY


def test_noalloc_search_matches_the_allocating_one(pkg):
    api = pkg.api
    DA = api.DeviceArray
    ctx = api.Context(0)
    f, g, _, _ = c1_pair()
    M, N = f.shape
    d_f = DA.from_host(ctx, f, np.float32)
    d_g = DA.from_host(ctx, g, np.float32)
    cases = [([1, M], [1, N], 30, 16, [10, 10, 10, 10]),   # 900 points
             ([41, 400], [17, 300], 20, 12, [6, 9, 6, 9])]  # a smaller second call
    for gx, gy, ws, step, reg in cases:
        nX, nY = _grid(M, N, gx, gy, ws, step)
        n = nX * nY
        print(f"grid {nX} x {nY}, winsize {ws}, region {reg}", flush=True)
        outs = []
        for fn in (api.integer_search_dev, api.integer_search_noalloc_dev):
            o = [DA(ctx, n, np.float64) for _ in range(5)] + [DA(ctx, n, np.int32)
                                                              for _ in range(2)]
            for a in o:  # distinct garbage, so nothing passes by leftover agreement
                a.upload(np.full(n, 7 if fn is api.integer_search_dev else -3))
            fn(ctx, d_f.ptr, d_g.ptr, M, N, gx, gy, ws, step, reg, *[a.ptr for a in o])
            ctx.synchronize()
            outs.append([a.to_host() for a in o])
        for k, (a, b) in enumerate(zip(*outs)):
            assert np.array_equal(a, b, equal_nan=True), k
        assert np.isfinite(outs[1][2]).sum() > 0.9 * n
    # a grid whose scratch exceeds the fixed array: an error with the size
    gx, gy, ws, step = [1, M], [1, N], 64, 4
    nX, nY = _grid(M, N, gx, gy, ws, step)
    o = [DA(ctx, nX * nY, np.float64) for _ in range(5)] + [DA(ctx, nX * nY, np.int32)
                                                            for _ in range(2)]
    with pytest.raises(pkg.AldicError, match="scratch"):
        api.integer_search_noalloc_dev(ctx, d_f.ptr, d_g.ptr, M, N, gx, gy, ws, step,
                                       [50, 50, 50, 50], *[a.ptr for a in o])
    ctx.close()
