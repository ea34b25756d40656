"""
Autotuner for the rasteriser's binning: tile shape and bin capacity.

For each tile shape it rasterises the caller's workload once with a generous
capacity, reads the per-tile bin counts back from the workspace, sizes the
capacity to the largest bin (so no bin spills), times the rasterisation with
that capacity, and installs the fastest (tile, capacity) in the context for
this (views, H, W, F) shape.  Results never depend on the choice: a spilled
bin only costs time.
"""
import time
from typing import Dict, Iterable, Tuple

import torch

from . import _lib, nvdr
from ._lib import check, ptr, stream_of

TILES = ((8, 8), (16, 8), (16, 16), (32, 8))


def _next_pow2(x: int) -> int:
    p = 1
    while p < x:
        p <<= 1
    return p


def _run(pos, tri, H, W, tile, cap, ws, boxes, rast):
    n = pos.shape[0] if pos.dim() == 3 else 1
    V, F = pos.shape[-2], tri.shape[0]
    L = _lib.lib()
    rc = L.drp_rasterize_fwd(stream_of(pos), ptr(pos), n, V, int(pos.dim() == 3), ptr(tri), F, None, H, W, None,
                             ptr(rast), None, ptr(ws), ws.numel(), tile[0], tile[1], cap, n * F, F, ptr(boxes))
    check(rc, 'rasterize (autotune)')


def bin_load(pos, tri, resolution, tile, probe_cap=4096) -> Tuple[int, int]:
    """(largest bin count, spilled triangles) of one rasterisation with ``tile``."""
    H, W = resolution
    n = pos.shape[0] if pos.dim() == 3 else 1
    F = tri.shape[0]
    L = _lib.lib()
    ws_bytes = L.drp_rasterize_workspace_size(n, H, W, n * F, tile[0], tile[1], probe_cap)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=pos.device)
    rast = torch.empty((n, H, W, 4), dtype=torch.float32, device=pos.device)
    boxes = torch.empty((n, F, 3), dtype=torch.int32, device=pos.device)
    _run(pos, tri, H, W, tile, probe_cap, ws, boxes, rast)
    n_tiles = n * ((W + tile[0] - 1) // tile[0]) * ((H + tile[1] - 1) // tile[1])
    # workspace layout: int32 bin counts [n_tiles], then spill-list counts [n]
    counts = ws[:4 * (n_tiles + n)].view(torch.int32)
    return int(counts[:n_tiles].max().item()), int(counts[n_tiles:].sum().item())


def tune_rasterize(ctx: 'nvdr.RasterizeCudaContext', pos: torch.Tensor, tri: torch.Tensor, resolution,
                   tiles: Iterable[Tuple[int, int]] = TILES, iters: int = 5, max_cap: int = 4096) -> Dict:
    """Pick the fastest (tile, capacity) for this workload and install it in ``ctx``."""
    pos = pos.detach().contiguous()
    tri = tri.contiguous()
    H, W = int(resolution[0]), int(resolution[1])
    n = pos.shape[0] if pos.dim() == 3 else 1
    F = tri.shape[0]
    L = _lib.lib()
    results = {}
    for tile in tiles:
        peak, _ = bin_load(pos, tri, (H, W), tile, probe_cap=max_cap)
        cap = min(max_cap, max(32, _next_pow2(int(peak * 1.15) + 1)))
        ws_bytes = L.drp_rasterize_workspace_size(n, H, W, n * F, tile[0], tile[1], cap)
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=pos.device)
        rast = torch.empty((n, H, W, 4), dtype=torch.float32, device=pos.device)
        boxes = torch.empty((n, F, 3), dtype=torch.int32, device=pos.device)
        _run(pos, tri, H, W, tile, cap, ws, boxes, rast)
        torch.cuda.synchronize(pos.device)
        st = torch.cuda.current_stream(pos.device)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(st)
        for _ in range(iters):
            _run(pos, tri, H, W, tile, cap, ws, boxes, rast)
        e1.record(st)
        torch.cuda.synchronize(pos.device)
        results[tile] = {'cap': cap, 'peak_bin': peak, 'ms': e0.elapsed_time(e1) / iters,
                         'workspace_mb': ws_bytes / 2 ** 20}
    best = min(results, key=lambda k: results[k]['ms'])
    ctx.set_tuning(ctx.tuning_key(n, H, W, F), best, results[best]['cap'])
    return {'best': {'tile': best, **results[best]}, 'all': {f'{k[0]}x{k[1]}': v for k, v in results.items()}}
