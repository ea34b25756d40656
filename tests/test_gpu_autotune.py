"""The binning autotuner: installs a valid configuration, and results never depend on it."""
import numpy as np
import pytest
import torch

from diffrp_amd import autotune, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def test_autotune_installs_config_and_results_are_invariant(cuda):
    pos, tri, _ = noisy_clip(5, 0.01, 160, views=3)
    pos, tri = pos.cuda(), tri.cuda()
    ctx = nvdr.RasterizeCudaContext()
    base, _ = nvdr.rasterize(ctx, pos, tri, (160, 160), grad_db=False)
    res = autotune.tune_rasterize(ctx, pos, tri, (160, 160), iters=2)
    tile = tuple(res['best']['tile'])
    assert tile in autotune.TILES and res['best']['cap'] >= res['best']['peak_bin']
    assert ctx.config(3, 160, 160, tri.shape[0]) == (tile, res['best']['cap'])
    tuned, _ = nvdr.rasterize(ctx, pos, tri, (160, 160), grad_db=False)
    assert torch.equal(base, tuned)
    for t in autotune.TILES:
        other, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(tile=t, bin_capacity=8), pos, tri, (160, 160),
                                  grad_db=False)
        assert torch.equal(base, other), t


def test_bin_load_reports_spill(cuda):
    pos, tri, _ = noisy_clip(5, 0.01, 128)
    peak, spilled = autotune.bin_load(pos.cuda(), tri.cuda(), (128, 128), (16, 16), probe_cap=4)
    assert peak > 4 and spilled > 0
    peak2, spilled2 = autotune.bin_load(pos.cuda(), tri.cuda(), (128, 128), (16, 16), probe_cap=4096)
    assert peak2 == peak and spilled2 == 0


def test_capacity_learnt_from_first_call(cuda):
    """A crumpled mesh loads its tiles far beyond the default bin capacity: the
    first eager call's exact bin counts (read back at the next call, without a
    host wait) size the capacity of the calls that follow; every call matches
    the oracle bit for bit."""
    import oracle
    res = (128, 128)
    pos, tri, _ = noisy_clip(7, 0.05, res[0])
    ctx = nvdr.RasterizeCudaContext()
    F = tri.shape[0]
    tile0 = nvdr.default_tile(*res, F)
    cap0 = nvdr.default_bin_capacity(tile0, *res, F)
    peak, spilled = autotune.bin_load(pos.cuda(), tri.cuda(), res, tile0, probe_cap=cap0)
    assert peak > cap0 and spilled > 0  # the heuristic's bins overflow here
    ref = oracle.rasterize(pos.numpy(), tri.numpy(), res)
    bits = np.ascontiguousarray(ref, np.float32).view(np.uint32)
    for call in range(3):
        r, _ = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), res, grad_db=False)
        assert np.array_equal(r.cpu().numpy().view(np.uint32), bits), f'call {call}'  # .cpu() waits
        if call >= 1:
            tile, cap = ctx.config(1, *res, F)
            assert tile == tile0 and cap >= peak, (tile, cap, peak)
