"""The allocation-free rasterize entry point (drp_rasterize_fwd_noalloc),
through the C-ABI, against drp_rasterize_fwd with a caller workspace: two
calls in a row, the second with a smaller input, equal bit for bit (rast,
rast_db and the per-triangle boxes), and a call beyond the library's scratch
fails with a message."""
import numpy as np
import pytest
import torch

from diffrp_amd import _lib
from diffrp_amd._lib import ptr, stream_of
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _call(noalloc, pos, tri, H, W, tile, cap, with_db, prev=None):
    L = _lib.lib()
    n, V, F = pos.shape[0], pos.shape[1], tri.shape[0]
    rast = torch.full((n, H, W, 4), float('nan'), device=pos.device)
    db = torch.full((n, H, W, 4), float('nan'), device=pos.device) if with_db else None
    boxes = torch.full((n, F, 3), -7, dtype=torch.int32, device=pos.device)
    s = stream_of(pos)
    if noalloc:
        rc = L.drp_rasterize_fwd_noalloc(s, ptr(pos), n, V, 1, ptr(tri), F, None, H, W, ptr(prev), ptr(rast),
                                         ptr(db), tile[0], tile[1], cap, n * F, F, ptr(boxes))
    else:
        ws_bytes = L.drp_rasterize_workspace_size(n, H, W, n * F, tile[0], tile[1], cap)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=pos.device)
        rc = L.drp_rasterize_fwd(s, ptr(pos), n, V, 1, ptr(tri), F, None, H, W, ptr(prev), ptr(rast), ptr(db),
                                 ptr(ws), ws_bytes, tile[0], tile[1], cap, n * F, F, ptr(boxes))
    assert rc == 0, L.drp_last_error()
    torch.cuda.synchronize()
    return rast, db, boxes


def _bits(t):
    return None if t is None else t.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize('tile,cap', [((16, 16), 64), ((32, 8), 8)])
def test_noalloc_rasterize_equals_workspace_entry_point(cuda, tile, cap):
    cases = [(noisy_clip(5, 0.01, 384, views=3), 384, 320, True),   # first call: 3 views, 20k triangles
             (noisy_clip(3, 0.01, 96, views=1), 96, 80, False)]     # second call: smaller
    for (pos, tri, _), H, W, with_db in cases:
        print(f'noalloc rasterize: {pos.shape[0]} views, {tri.shape[0]} triangles, {H}x{W}, tile {tile}, cap {cap}',
              flush=True)
        p, t = pos.to(cuda).contiguous(), tri.to(cuda).int().contiguous()
        want = _call(False, p, t, H, W, tile, cap, with_db)
        got = _call(True, p, t, H, W, tile, cap, with_db)
        for a, b in zip(got, want):
            if b is not None:
                np.testing.assert_array_equal(_bits(a), _bits(b))
        assert (want[0][..., 3] > 0).any()
        # a peeled second layer through the same scratch
        want2 = _call(False, p, t, H, W, tile, cap, False, prev=want[0])
        got2 = _call(True, p, t, H, W, tile, cap, False, prev=got[0])
        np.testing.assert_array_equal(_bits(got2[0]), _bits(want2[0]))


def test_noalloc_rejects_work_beyond_its_scratch(cuda):
    L = _lib.lib()
    n, H, W, F = 64, 512, 512, 81920
    need = L.drp_rasterize_workspace_size(n, H, W, n * F, 32, 8, 4096)
    assert need > (64 << 20)
    pos = torch.zeros((n, 3, 4), device=cuda)
    tri = torch.zeros((F, 3), dtype=torch.int32, device=cuda)
    rast = torch.empty((n, H, W, 4), device=cuda)
    rc = L.drp_rasterize_fwd_noalloc(stream_of(pos), ptr(pos), n, 3, 1, ptr(tri), F, None, H, W, None, ptr(rast),
                                     None, 32, 8, 4096, n * F, F, None)
    assert rc == -3
    assert b'scratch' in L.drp_last_error()
