"""Host-side sizing rules of the rasteriser context (CPU, no GPU calls)."""
import pytest

from diffrp_amd import nvdr
from diffrp_amd.nvdr import learnt_bin_capacity


def test_learnt_capacity_keeps_a_capacity_that_did_not_overflow():
    assert learnt_bin_capacity(peak=100, cap=512, n_tiles=4096, items=81920) == 512
    assert learnt_bin_capacity(peak=512, cap=512, n_tiles=4096, items=81920) == 512


def test_learnt_capacity_is_the_power_of_two_above_the_peak_plus_margin():
    # small image: the budget does not bind
    assert learnt_bin_capacity(peak=1000, cap=256, n_tiles=256, items=1280) == 2048   # 1151 -> 2048
    assert learnt_bin_capacity(peak=800, cap=256, n_tiles=256, items=1280) == 1024    # 921 -> 1024
    assert learnt_bin_capacity(peak=28000, cap=256, n_tiles=256, items=1310720) == 32768


@pytest.mark.parametrize('n_tiles,items', [(16384, 1310720), (65536, 64 * 81920), (1 << 20, 5242880), (4096, 81920)])
def test_learnt_capacity_is_bounded_by_the_memory_budget(n_tiles, items):
    # one dense tile (a distant high-poly object) must not size every tile's bin
    cap = learnt_bin_capacity(peak=60000, cap=64, n_tiles=n_tiles, items=items)
    budget = (256 << 20) + 64 * items
    assert cap * n_tiles * 4 <= budget or cap == 64
    assert cap & (cap - 1) == 0
    assert cap >= 64


def test_learnt_capacity_budget_examples():
    # 2048^2 in 16x16 tiles, 1.3M triangles: 4096 slots per tile (64 MiB of bins), not 65536 (4 GiB)
    assert learnt_bin_capacity(peak=60000, cap=512, n_tiles=16384, items=1310720) == 4096
    # 64 views at 512^2 in 32x8 tiles: 2048
    assert learnt_bin_capacity(peak=60000, cap=1024, n_tiles=65536, items=64 * 81920) == 2048


def test_tuned_shapes_are_bounded_least_recently_used():
    ctx = nvdr.RasterizeCudaContext(device='cuda:0')
    ctx.MAX_TUNED = 3
    for k in range(3):
        ctx.set_tuning((1, 64, 64, k), (8, 8), 32)
    assert ctx.config(1, 64, 64, 0) == ((8, 8), 32)   # touch shape 0
    ctx.set_tuning((1, 64, 64, 3), (16, 16), 64)
    assert (1, 64, 64, 1) not in ctx._tuned          # the least recently used one went
    assert ctx.config(1, 64, 64, 0) == ((8, 8), 32)
    assert ctx.config(1, 64, 64, 3) == ((16, 16), 64)


def test_installed_tuning_is_not_replaced_by_a_pending_first_call(monkeypatch):
    monkeypatch.setattr(nvdr.torch.cuda, 'is_current_stream_capturing', lambda: False)
    ctx = nvdr.RasterizeCudaContext(device='cuda:0')

    class _Done:
        def query(self):
            return True
    key = (3, 160, 160, 20480)
    ctx._pending[key] = ((8, 8), 512, [100000], _Done(), 400, 61440)
    ctx.set_tuning(key, (32, 8), 1024)
    assert key not in ctx._pending
    ctx._pending[key] = ((8, 8), 512, [100000], _Done(), 400, 61440)  # a first call racing the tuner
    assert ctx.config(*key) == ((32, 8), 1024)
    assert not ctx._pending
    # without an installed tuning, the completed read-back is learnt
    key2 = (1, 128, 128, 1310720)
    ctx._pending[key2] = ((16, 16), 64, [1000], _Done(), 64, 1310720)
    assert ctx.config(*key2) == ((16, 16), 2048)
