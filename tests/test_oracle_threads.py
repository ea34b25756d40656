"""The oracle's results do not depend on its thread count (CPU): the forward
depth test is an order-free minimum and the backward passes sum per-band
buffers in band order."""
import numpy as np

import oracle
from tests.scenes import noisy_clip


def _run(pos, tri, attr, target):
    H = W = 96
    rast = oracle.rasterize(pos, tri, (H, W))
    opp = oracle.topology(tri)
    col = oracle.interpolate(attr, rast, tri)
    y = oracle.antialias(col, rast, pos, tri, opp)
    gcol, gpos_aa = oracle.antialias_bwd(col, rast, pos, tri, 2 * (y - target), opp)
    gattr, grast = oracle.interpolate_bwd(attr, rast, tri, gcol)
    gpos = oracle.rasterize_bwd(pos, tri, rast, grast)
    return rast, y, gcol, gpos_aa, gattr, grast, gpos


def test_results_independent_of_thread_count(oracle_lib):
    pos, tri, _ = noisy_clip(3, 0.01, 96, views=3)
    pos, tri = pos.numpy(), tri.numpy()
    rng = np.random.default_rng(50)
    attr = rng.random((pos.shape[1], 4)).astype(np.float32)
    target = rng.random((3, 96, 96, 4)).astype(np.float32)
    try:
        oracle.set_threads(1)
        one = _run(pos, tri, attr, target)
        oracle.set_threads(5)
        many = _run(pos, tri, attr, target)
    finally:
        oracle.set_threads(0)
    for a, b in zip(one, many):
        np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))
    assert np.abs(one[-1]).max() > 0 and np.abs(one[3]).max() > 0
