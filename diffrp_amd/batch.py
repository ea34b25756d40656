"""
Uneven batches: many meshes, each with its own triangle count and resolution,
rasterised in one launch.

diffrp renders one mesh set per ``DepthPeeler`` call
(diffrp/rendering/surface_deferred.py:259-265); a training loop over many
objects repeats that call per object.  ``rasterize_batch`` runs the binning and
the fine raster of all of them in one launch (one tile queue over all images,
each tile reading its image's descriptor), and returns one ``rast`` per item.
Each item's ``rast`` is exactly what ``nvdr.rasterize`` gives for that item
alone (triangle ids are item-local), so ``nvdr.interpolate`` /
``nvdr.antialias`` / the backward passes apply to it unchanged.

    rasts, dbs = rasterize_batch(ctx, [pos_0, pos_1, ...], [tri_0, tri_1, ...],
                                 [(H_0, W_0), (H_1, W_1), ...])

pos_i are [V_i, 4] clip-space vertices, tri_i [F_i, 3] int32 with item-local
indices.  Gradients flow to every pos_i (per-item backward launches).
"""
import ctypes
from typing import List, Sequence, Tuple

import numpy as np
import torch

from . import _lib, nvdr
from ._lib import check, ptr, stream_of, require_device

_c_ll = ctypes.c_longlong


class BatchPlan:
    """Per-item layout of an uneven batch for one tile shape (host and device copies).

    caps: bin capacity per item (None or 0 entries: sized from the item's
    triangle density by the planner)."""

    def __init__(self, items: np.ndarray, tile: Tuple[int, int], device, caps=None):
        items = np.asarray(items, dtype=np.int32).reshape(-1, 4)  # V, F, H, W
        caps = np.zeros(len(items), np.int32) if caps is None else np.asarray(caps, np.int32).reshape(-1)
        self.items = np.ascontiguousarray(items)
        self.caps = caps
        self.items5 = np.ascontiguousarray(np.concatenate([items, caps[:, None]], 1), dtype=np.int32)
        self.tile = tuple(tile)
        L = _lib.lib()
        n = len(self.items)
        ip = self.items5.ctypes.data_as(ctypes.c_void_p)
        nbytes = L.drp_batch_plan_size(ip, n, self.tile[0], self.tile[1])
        if nbytes == 0:
            raise ValueError('empty batch or unsupported tile shape')
        self.host = np.zeros(nbytes, dtype=np.uint8)
        check(L.drp_batch_plan(ip, n, self.tile[0], self.tile[1], self.host.ctypes.data_as(ctypes.c_void_p), nbytes),
              'batch plan')
        self.dev = torch.from_numpy(self.host).to(device)
        vals = [_c_ll() for _ in range(4)]
        check(L.drp_batch_plan_info(self.host_ptr(), *[ctypes.byref(v) for v in vals]), 'batch plan info')
        self.n_pix, self.n_tiles, self.n_tri, self.n_vert = (v.value for v in vals)
        V, F, H, W = (self.items[:, k].astype(np.int64) for k in range(4))
        tiles = ((W + self.tile[0] - 1) // self.tile[0]) * ((H + self.tile[1] - 1) // self.tile[1])
        self.tbase = np.concatenate([[0], np.cumsum(tiles)[:-1]])
        self.tiles = tiles
        self.vbase = np.concatenate([[0], np.cumsum(V)[:-1]])
        self.fbase = np.concatenate([[0], np.cumsum(F)[:-1]])
        self.pbase = np.concatenate([[0], np.cumsum(H * W)[:-1]])

    def host_ptr(self):
        return self.host.ctypes.data_as(ctypes.c_void_p)

    def __len__(self):
        return len(self.items)

    def item(self, i):
        """(vertex base, V, triangle base, F, pixel base, H, W) of item i."""
        V, F, H, W = (int(x) for x in self.items[i])
        return int(self.vbase[i]), V, int(self.fbase[i]), F, int(self.pbase[i]), H, W


def default_batch_tile(items: np.ndarray) -> Tuple[int, int]:
    """Tile shape for a batch without tuning (capacities: per item, from density)."""
    return (16, 16) if float(np.mean(items[:, 2].astype(np.float64) * items[:, 3])) >= 256 * 256 else (8, 8)


def batch_key(items: np.ndarray):
    return ('batch',) + tuple(int(x) for x in np.asarray(items).reshape(-1))


def _launch(plan: BatchPlan, pos, tri, prev, want_db):
    dev = pos.device
    L = _lib.lib()
    rast = torch.empty((plan.n_pix, 4), dtype=torch.float32, device=dev)
    db = torch.empty((plan.n_pix, 4), dtype=torch.float32, device=dev) if want_db else None
    ws_bytes = L.drp_rasterize_batch_workspace_size(plan.host_ptr())
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    rc = L.drp_rasterize_batch_fwd(stream_of(pos), ptr(pos), ptr(tri), plan.host_ptr(), ptr(plan.dev), ptr(prev),
                                   ptr(rast), ptr(db), ptr(ws), ws_bytes, None)
    check(rc, 'batch rasterize')
    return rast, db, ws


class _RasterizeBatchFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pos, tri, plan, prev, want_db):
        rast, db, _ = _launch(plan, pos, tri, prev, want_db)
        fctx.save_for_backward(pos, tri, rast)
        fctx.plan = plan
        fctx.want_db = want_db
        if db is None:
            db = torch.empty((plan.n_pix, 0), dtype=torch.float32, device=pos.device)
        return rast, db

    @staticmethod
    def backward(fctx, grast, gdb):
        pos, tri, rast = fctx.saved_tensors
        plan = fctx.plan
        gpos = torch.zeros_like(pos)
        if grast is None:
            return gpos, None, None, None, None
        grast = nvdr._as_f32(grast, 'grad_rast')
        gdb = nvdr._as_f32(gdb, 'grad_db') if (fctx.want_db and gdb is not None and gdb.numel()) else None
        L = _lib.lib()
        s = stream_of(pos)
        for i in range(len(plan)):
            vb, V, fb, F, pb, H, W = plan.item(i)
            if V == 0 or F == 0:
                continue
            # per item: the per-pixel chain with float atomics into the item's vertices
            rc = L.drp_rasterize_bwd(s, ptr(pos[vb:vb + V]), 1, V, 0, ptr(tri[fb:fb + F]), F, None, H, W,
                                     ptr(rast[pb:pb + H * W]), ptr(grast[pb:pb + H * W]),
                                     ptr(gdb[pb:pb + H * W]) if gdb is not None else None, ptr(gpos[vb:vb + V]),
                                     None, None, None, None, 0)
            check(rc, 'batch rasterize backward')
        return gpos, None, None, None, None


def _prepare(pos_list, tri_list, resolutions):
    if not (len(pos_list) == len(tri_list) == len(resolutions)) or not pos_list:
        raise ValueError('pos_list, tri_list and resolutions must be non-empty and of equal length')
    require_device(*pos_list, *tri_list)
    items = []
    for p, t, (H, W) in zip(pos_list, tri_list, resolutions):
        if p.dim() != 2 or p.shape[1] != 4:
            raise ValueError(f'each pos must be [V, 4] (got {tuple(p.shape)})')
        items.append((p.shape[0], nvdr._as_tri(t).shape[0], int(H), int(W)))
    items = np.asarray(items, np.int32)
    pos = nvdr._as_f32(torch.cat([p for p in pos_list]), 'pos')
    tri = torch.cat([nvdr._as_tri(t) for t in tri_list]).contiguous()
    return items, pos, tri


def _split(plan: BatchPlan, packed: torch.Tensor, ch: int):
    out = []
    for i in range(len(plan)):
        _, _, _, _, pb, H, W = plan.item(i)
        out.append(packed[pb:pb + H * W].view(1, H, W, ch))
    return out


def rasterize_batch(ctx: 'nvdr.RasterizeCudaContext', pos_list: Sequence[torch.Tensor],
                    tri_list: Sequence[torch.Tensor], resolutions: Sequence[Tuple[int, int]], grad_db: bool = True,
                    prev: List[torch.Tensor] = None):
    """Rasterise every item of an uneven batch in one launch; returns (rasts, rast_dbs).

    rasts[i] is [1, H_i, W_i, 4] (item-local triangle ids), rast_dbs[i] the
    screen-space derivatives ([1, H_i, W_i, 0] when grad_db is False).
    prev: the previous layer's rasts (depth peeling), not differentiated.
    """
    items, pos, tri = _prepare(pos_list, tri_list, resolutions)
    key = batch_key(items)
    if key in ctx._tuned and ctx.tile is None and ctx.bin_capacity is None:
        tile, caps = ctx._tuned[key]
    else:
        tile = ctx.tile or default_batch_tile(items)
        caps = None if ctx.bin_capacity is None else [ctx.bin_capacity] * len(items)
    plan = BatchPlan(items, tile, pos.device, caps)
    prev_packed = None
    if prev is not None:
        prev_packed = torch.cat([p.detach().reshape(-1, 4) for p in prev]).contiguous()
        if prev_packed.shape[0] != plan.n_pix:
            raise ValueError('prev does not match the batch layout')
    rast, db = _RasterizeBatchFunc.apply(pos, tri, plan, prev_packed, bool(grad_db))
    rasts = _split(plan, rast, 4)
    dbs = _split(plan, db, 4) if grad_db else [torch.empty((1, H, W, 0), device=pos.device)
                                                for (_, _, H, W) in items]
    return rasts, dbs


class BatchDepthPeeler:
    """Depth peeling over an uneven batch (``dr.DepthPeeler`` per item, in one launch per layer)."""

    def __init__(self, ctx, pos_list, tri_list, resolutions, grad_db=True):
        self.ctx, self.pos_list, self.tri_list = ctx, list(pos_list), list(tri_list)
        self.resolutions, self.grad_db = list(resolutions), grad_db
        self.prev = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.prev = None

    def rasterize_next_layer(self):
        rasts, dbs = rasterize_batch(self.ctx, self.pos_list, self.tri_list, self.resolutions, self.grad_db,
                                     prev=self.prev)
        self.prev = [r.detach() for r in rasts]
        return rasts, dbs


def tune_batch(ctx, pos_list, tri_list, resolutions, tiles=((8, 8), (16, 8), (16, 16), (32, 8)), iters=3,
               max_cap=1 << 16):
    """Autotune the tile shape and the per-item bin capacities of an uneven
    batch and install them in ``ctx``.

    Per tile shape: one probe launch reads the exact per-tile bin counts back
    from the workspace; each item's capacity is its largest count + 15 %
    rounded up to a power of two (so no bin spills); then the launch is timed.
    Results never depend on the choice."""
    items, pos, tri = _prepare(pos_list, tri_list, resolutions)
    pos = pos.detach()
    results = {}
    for tile in tiles:
        plan = BatchPlan(items, tile, pos.device)
        _, _, ws = _launch(plan, pos, tri, None, False)
        counts = ws[:4 * plan.n_tiles].view(torch.int32).cpu().numpy()  # counts are exact whatever the capacity
        caps = []
        for i in range(len(plan)):
            t0, nt = int(plan.tbase[i]), int(plan.tiles[i])
            peak = int(counts[t0:t0 + nt].max()) if nt else 0
            cap = 32
            while cap < int(peak * 1.15) + 1 and cap < max_cap:
                cap *= 2
            caps.append(cap)
        plan = BatchPlan(items, tile, pos.device, caps)
        _launch(plan, pos, tri, None, False)
        torch.cuda.synchronize(pos.device)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        st = torch.cuda.current_stream(pos.device)
        e0.record(st)
        for _ in range(iters):
            _launch(plan, pos, tri, None, False)
        e1.record(st)
        torch.cuda.synchronize(pos.device)
        results[tile] = (e0.elapsed_time(e1) / iters, tuple(caps))
    best = min(results, key=lambda t: results[t][0])
    ctx._tuned[batch_key(items)] = (best, results[best][1])
    return {'tile': best, 'bin_capacities': list(results[best][1]),
            'candidates_ms': {f'{t[0]}x{t[1]}': round(v[0], 4) for t, v in results.items()}}
