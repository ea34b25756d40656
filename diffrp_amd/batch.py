"""
Uneven batches: many meshes, each with its own triangle count and resolution,
rendered together.

diffrp renders one mesh set per ``DepthPeeler`` call
(diffrp/rendering/surface_deferred.py:259-265) and interpolates / antialiases
per call (interpolator.py:83, surface_deferred.py:775); a training loop over
many objects repeats that per object (scripts/quick_render.py:85-93 loops over
views the same way).  Two forms:

``rasterize_batch`` runs the binning and the fine raster of all items in one
launch (one tile queue over all images, each tile reading its image's
descriptor) and returns one ``rast`` per item, exactly what ``nvdr.rasterize``
gives for that item alone (triangle ids are item-local), so the per-item
operators apply to it; its backward is one deterministic launch sequence over
all items.

    rasts, dbs = rasterize_batch(ctx, [pos_0, pos_1, ...], [tri_0, tri_1, ...],
                                 [(H_0, W_0), (H_1, W_1), ...])

``UnevenBatch`` keeps the items packed (pixels, vertices, triangles back to
back, triangle ids global) and runs rasterize, interpolate and antialias and
their backward passes as one launch sequence per operator over all items.  It
is built once per batch shape (plan, topology and incidence tables uploaded
then), so a training step over the batch can be captured in a HIP graph:

    ub = UnevenBatch(ctx, [tri_0, ...], [V_0, ...], [(H_0, W_0), ...])
    rast, boxes = ub.rasterize(pos)          # pos: the items' [V_i, 4] concatenated
    col = ub.interpolate(attr, rast, boxes)  # attr: [sum V_i, C]
    out = ub.antialias(col, rast, pos)       # [n_pix, C]; ub.split(out) gives per-item images

pos_i are [V_i, 4] clip-space vertices, tri_i [F_i, 3] int32 with item-local
indices.
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


def _launch(plan: BatchPlan, pos, tri, prev, want_db, global_ids=False, want_boxes=False):
    dev = pos.device
    L = _lib.lib()
    rast = torch.empty((plan.n_pix, 4), dtype=torch.float32, device=dev)
    db = torch.empty((plan.n_pix, 4), dtype=torch.float32, device=dev) if want_db else None
    boxes = torch.empty((max(plan.n_tri, 1), 3), dtype=torch.int32, device=dev) if want_boxes else None
    ws_bytes = L.drp_rasterize_batch_workspace_size(plan.host_ptr())
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    rc = L.drp_rasterize_batch_fwd(stream_of(pos), ptr(pos), ptr(tri), plan.host_ptr(), ptr(plan.dev), ptr(prev),
                                   ptr(rast), ptr(db), ptr(ws), ws_bytes, ptr(boxes), int(global_ids))
    check(rc, 'batch rasterize')
    return rast, db, ws, boxes


def _global_tri(plan: BatchPlan, tri: torch.Tensor) -> torch.Tensor:
    """The batch's triangles with batch (global) vertex indices: each item's
    offset by the vertices of the items before it."""
    F = torch.as_tensor(plan.items[:, 1].astype(np.int64))
    vb = torch.as_tensor(plan.vbase.astype(np.int64))
    off = torch.repeat_interleave(vb, F).to(tri.device)
    return (tri.long() + off[:, None]).int().contiguous()


def _rasterize_bwd(plan: BatchPlan, pos, tri_g, rast, local_ids, grast, gdb, boxes, inc):
    """Gradient of every item's rast (and rast_db) w.r.t. the packed positions:
    one deterministic launch sequence over all items."""
    gpos = torch.zeros_like(pos)
    if plan.n_tri == 0:
        return gpos
    L = _lib.lib()
    ws_bytes = L.drp_batch_rasterize_bwd_workspace_size(plan.host_ptr())
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=pos.device)
    check(L.drp_batch_rasterize_bwd(stream_of(pos), plan.host_ptr(), ptr(plan.dev), ptr(pos), ptr(tri_g), ptr(rast),
                                    int(local_ids), ptr(grast), ptr(gdb), ptr(gpos), ptr(boxes), ptr(inc[0]),
                                    ptr(inc[1]), ptr(ws), ws_bytes), 'batch rasterize backward')
    return gpos


class _RasterizeBatchFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pos, tri, plan, prev, want_db):
        rast, db, _, boxes = _launch(plan, pos, tri, prev, want_db, want_boxes=True)
        fctx.save_for_backward(pos, tri, rast, boxes)
        fctx.plan = plan
        fctx.want_db = want_db
        fctx.set_materialize_grads(False)  # a gradient nothing sends arrives as None, not as zeros
        if db is None:
            db = torch.empty((plan.n_pix, 0), dtype=torch.float32, device=pos.device)
        return rast, db

    @staticmethod
    def backward(fctx, grast, gdb):
        pos, tri, rast, boxes = fctx.saved_tensors
        plan = fctx.plan
        gdb = nvdr._as_f32(gdb, 'grad_db') if (fctx.want_db and gdb is not None and gdb.numel()) else None
        if grast is None and gdb is None:
            return None, None, None, None, None
        grast = torch.zeros_like(rast) if grast is None else nvdr._as_f32(grast, 'grad_rast')
        tri_g = _global_tri(plan, tri)
        inc = nvdr.incidence(tri_g, plan.n_vert)
        return _rasterize_bwd(plan, pos, tri_g, rast, True, grast, gdb, boxes, inc), None, None, None, None


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
        _, _, ws, _ = _launch(plan, pos, tri, None, False)
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
    with ctx._lock:
        ctx._remember(batch_key(items), (best, results[best][1]))
    return {'tile': best, 'bin_capacities': list(results[best][1]),
            'candidates_ms': {f'{t[0]}x{t[1]}': round(v[0], 4) for t, v in results.items()}}


# ----------------------------------------------------------------------------
# packed uneven batches: the whole fwd + bwd per operator in one launch sequence
# ----------------------------------------------------------------------------
class UnevenBatch:
    """Rasterize, interpolate and antialias (and their backward passes) over an
    uneven batch kept packed: item i's H_i x W_i pixels follow the items before
    it in one [n_pix, .] buffer, its vertices and triangles likewise, and rast
    ids are batch triangles + 1.

    Built once per batch shape: the plan (items, tile shape, per-item bin
    capacities -- the tuned ones when ``tune_batch`` installed them in ``ctx``)
    is uploaded, and the batch triangle table with global vertex indices, its
    topology table and its vertex incidence table are built, here.  The
    operators then only launch kernels (their scratch comes from torch's
    caching allocator), so a step can be captured in a HIP graph.

    tri_list: the items' [F_i, 3] int32 triangles (item-local vertex indices);
    n_verts: V_i per item; resolutions: (H_i, W_i) per item.
    """

    def __init__(self, ctx: 'nvdr.RasterizeCudaContext', tri_list: Sequence[torch.Tensor], n_verts: Sequence[int],
                 resolutions: Sequence[Tuple[int, int]]):
        if not (len(tri_list) == len(n_verts) == len(resolutions)) or not tri_list:
            raise ValueError('tri_list, n_verts and resolutions must be non-empty and of equal length')
        require_device(*tri_list)
        self.device = tri_list[0].device
        items = np.asarray([(int(v), nvdr._as_tri(t).shape[0], int(h), int(w))
                            for t, v, (h, w) in zip(tri_list, n_verts, resolutions)], np.int32)
        key = batch_key(items)
        tuned = ctx._tuned.get(key) if ctx.tile is None and ctx.bin_capacity is None else None
        if tuned is not None:
            tile, caps = tuned
        else:
            tile = ctx.tile or default_batch_tile(items)
            caps = None if ctx.bin_capacity is None else [ctx.bin_capacity] * len(items)
        self.plan = BatchPlan(items, tile, self.device, caps)
        self.tri_local = torch.cat([nvdr._as_tri(t) for t in tri_list]).contiguous()
        self.tri = _global_tri(self.plan, self.tri_local)
        self.opp = nvdr.antialias_construct_topology_hash(self.tri).opp
        self.inc = nvdr.incidence(self.tri, self.plan.n_vert)

    @property
    def n_pix(self):
        return self.plan.n_pix

    def split(self, packed: torch.Tensor) -> List[torch.Tensor]:
        """Per-item [1, H_i, W_i, C] views of a packed [n_pix, C] buffer."""
        return _split(self.plan, packed, packed.shape[-1])

    def rasterize(self, pos: torch.Tensor):
        """(rast [n_pix, 4], boxes) of every item; pos [sum V_i, 4] clip space.
        Differentiable w.r.t. pos; pass ``boxes`` on to ``interpolate``."""
        require_device(pos, names=('pos',))
        pos = nvdr._as_f32(pos, 'pos')
        if pos.shape != (self.plan.n_vert, 4):
            raise ValueError(f'pos must be [{self.plan.n_vert}, 4] (got {tuple(pos.shape)})')
        rast, boxes = _URasterize.apply(pos, self)
        if rast.requires_grad:
            # the positions it rasterised: an interpolate of this rast can run the
            # rasterize backward in its own pass over the pixels
            rast._drp_upos = (self, pos, rast._version)
        return rast, boxes

    def interpolate(self, attr: torch.Tensor, rast: torch.Tensor, boxes: torch.Tensor) -> torch.Tensor:
        """[n_pix, C] attributes; attr [sum V_i, C]."""
        require_device(attr, rast, names=('attr', 'rast'))
        attr = nvdr._as_f32(attr, 'attr')
        if attr.dim() != 2 or attr.shape[0] != self.plan.n_vert:
            raise ValueError(f'attr must be [{self.plan.n_vert}, C] (got {tuple(attr.shape)})')
        meta = getattr(rast, '_drp_upos', None)
        src_pos = meta[1] if (meta is not None and meta[0] is self and meta[2] == rast._version
                              and attr.shape[1] in _FUSED_C) else None
        return _UInterpolate.apply(attr, nvdr._as_f32(rast, 'rast'), boxes, self, src_pos)

    def antialias(self, color: torch.Tensor, rast: torch.Tensor, pos: torch.Tensor,
                  pos_gradient_boost: float = 1.0) -> torch.Tensor:
        """[n_pix, C] antialiased colours; differentiable w.r.t. color and pos."""
        require_device(color, rast, pos, names=('color', 'rast', 'pos'))
        color = nvdr._as_f32(color, 'color')
        if color.dim() != 2 or color.shape[0] != self.plan.n_pix:
            raise ValueError(f'color must be [{self.plan.n_pix}, C] (got {tuple(color.shape)})')
        return _UAntialias.apply(color, nvdr._as_f32(rast, 'rast'), nvdr._as_f32(pos, 'pos'), self,
                                 float(pos_gradient_boost))


class _URasterize(torch.autograd.Function):
    @staticmethod
    def forward(fctx, pos, ub):
        rast, _, _, boxes = _launch(ub.plan, pos, ub.tri_local, None, False, global_ids=True, want_boxes=True)
        fctx.save_for_backward(pos, rast, boxes)
        fctx.ub = ub
        fctx.mark_non_differentiable(boxes)
        # no gradient reaching rast (its consumers fused or non-differentiable)
        # arrives as None, not as zeros to run the backward over
        fctx.set_materialize_grads(False)
        return rast, boxes

    @staticmethod
    def backward(fctx, grast, _gboxes):
        pos, rast, boxes = fctx.saved_tensors
        ub = fctx.ub
        if grast is None:
            return None, None
        grast = nvdr._as_f32(grast, 'grad_rast')
        return _rasterize_bwd(ub.plan, pos, ub.tri, rast, False, grast, None, boxes, ub.inc), None


_FUSED_C = (4, 8)  # channel counts of the fused attribute + position backward (tests switch it off)


class _UInterpolate(torch.autograd.Function):
    @staticmethod
    def forward(fctx, attr, rast, boxes, ub, src_pos):
        C = attr.shape[1]
        out = torch.empty((ub.plan.n_pix, C), dtype=torch.float32, device=attr.device)
        check(_lib.lib().drp_batch_interpolate_fwd(stream_of(attr), ub.plan.host_ptr(), ptr(attr), C, ptr(rast),
                                                   ptr(ub.tri), ptr(out)), 'batch interpolate')
        fctx.save_for_backward(attr, rast, boxes, src_pos)
        fctx.ub = ub
        return out

    @staticmethod
    def backward(fctx, gout):
        attr, rast, boxes, src_pos = fctx.saved_tensors
        ub = fctx.ub
        C = attr.shape[1]
        gout = nvdr._as_f32(gout, 'grad_out')
        gattr = torch.zeros_like(attr)
        L = _lib.lib()
        if src_pos is not None and fctx.needs_input_grad[4]:
            # the rasterize backward of this interpolation's grad_rast in the same
            # pass, returned as the gradient of the rasterised positions (the
            # rasterize backward, linear in grad_rast, sees only rast's other consumers)
            gpos = torch.zeros_like(src_pos)
            ws_bytes = L.drp_batch_interpolate_rasterize_bwd_workspace_size(ub.plan.host_ptr(), C)
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=attr.device)
            check(L.drp_batch_interpolate_rasterize_bwd(stream_of(attr), ub.plan.host_ptr(), ptr(ub.plan.dev),
                                                        ptr(attr), C, ptr(src_pos), ptr(rast), ptr(ub.tri),
                                                        ptr(gout), ptr(gattr), ptr(gpos), ptr(boxes),
                                                        ptr(ub.inc[0]), ptr(ub.inc[1]), ptr(ws), ws_bytes),
                  'batch interpolate + rasterize backward')
            return gattr, None, None, None, gpos
        grast = torch.empty_like(rast)
        ws_bytes = L.drp_batch_interpolate_bwd_workspace_size(ub.plan.host_ptr(), C)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=attr.device)
        check(L.drp_batch_interpolate_bwd(stream_of(attr), ub.plan.host_ptr(), ptr(ub.plan.dev), ptr(attr), C,
                                          ptr(rast), ptr(ub.tri), ptr(gout), ptr(gattr), ptr(grast), ptr(boxes),
                                          ptr(ub.inc[0]), ptr(ub.inc[1]), ptr(ws), ws_bytes),
              'batch interpolate backward')
        return gattr, grast, None, None, None


class _UAntialias(torch.autograd.Function):
    @staticmethod
    def forward(fctx, color, rast, pos, ub, boost):
        C = color.shape[1]
        out = torch.empty_like(color)
        L = _lib.lib()
        ws_bytes = L.drp_batch_antialias_workspace_size(ub.plan.host_ptr())
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=color.device)
        check(L.drp_batch_antialias_fwd(stream_of(color), ub.plan.host_ptr(), ptr(ub.plan.dev), ptr(color), C,
                                        ptr(rast), ptr(pos), ptr(ub.tri), ptr(ub.opp), ptr(out), ptr(ws), ws_bytes),
              'batch antialias')
        fctx.save_for_backward(color, rast, pos)
        fctx.ub, fctx.boost, fctx.ws = ub, boost, ws
        return out

    @staticmethod
    def backward(fctx, gout):
        color, rast, pos = fctx.saved_tensors
        ub = fctx.ub
        C = color.shape[1]
        gout = nvdr._as_f32(gout, 'grad_out')
        gcol = torch.empty_like(color)
        gpos = torch.zeros_like(pos) if fctx.needs_input_grad[2] else None
        L = _lib.lib()
        ws, reuse = fctx.ws, 1
        if ws is None:  # a second backward: the pair decisions are recomputed
            ws_bytes = L.drp_batch_antialias_workspace_size(ub.plan.host_ptr())
            ws, reuse = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=color.device), 0
        check(L.drp_batch_antialias_bwd(stream_of(color), ub.plan.host_ptr(), ptr(ub.plan.dev), ptr(color), C,
                                        ptr(rast), ptr(pos), ptr(ub.tri), ptr(ub.opp), ptr(gout), ptr(gcol),
                                        ptr(gpos), float(fctx.boost), reuse, ptr(ws), ws.numel()),
              'batch antialias backward')
        fctx.ws = None
        return gcol, None, gpos, None, None
