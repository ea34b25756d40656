"""
nvdiffrast.torch-compatible operators on MI355X (gfx950).

This module is the drop-in for the surface diffrp calls
(``diffrp/rendering/surface_deferred.py:4,209,225,260-265,775``,
``diffrp/rendering/interpolator.py:12,83``):

    RasterizeCudaContext()                               surface_deferred.py:225
    RasterizeGLContext                                   surface_deferred.py:209 (type only)
    DepthPeeler(ctx, pos, tri, resolution, ranges)       surface_deferred.py:260
      .rasterize_next_layer() -> (rast, rast_db)         surface_deferred.py:262
    interpolate(attr, rast, tri) -> (out, out_da)        interpolator.py:83
    antialias(color, rast, pos, tri) -> out              surface_deferred.py:775

plus the rest of nvdiffrast's documented surface the configurations need:
``rasterize`` (range and instanced modes), ``interpolate(..., rast_db,
diff_attrs)``, ``antialias(..., topology_hash, pos_gradient_boost)`` and
``antialias_construct_topology_hash``.  Every operator runs hand-written HIP
kernels through the C-ABI of ``libdrp_hip.so``; gradients flow through
``torch.autograd``.  Inputs must be ROCm device tensors: there is no CPU path.
``install_as_nvdiffrast()`` registers this module as ``nvdiffrast.torch`` so
diffrp itself can import it unchanged.
"""
import ctypes
import sys
import threading
import types
import weakref
from collections import OrderedDict
from typing import Optional, Sequence, Tuple, Union

import numpy as np
import torch

from . import _lib
from ._lib import ptr, check, stream_of, require_device

__all__ = [
    'RasterizeCudaContext', 'RasterizeGLContext', 'RasterizeHipContext', 'DepthPeeler', 'rasterize',
    'interpolate', 'antialias', 'antialias_construct_topology_hash', 'TopologyHash', 'install_as_nvdiffrast',
    'get_log_level', 'set_log_level',
]

_TILES = {(8, 8), (16, 8), (16, 16), (32, 8)}


# ----------------------------------------------------------------------------
# context
# ----------------------------------------------------------------------------
class RasterizeCudaContext:
    """Rasteriser context (``dr.RasterizeCudaContext``, built at ``surface_deferred.py:225``).

    Holds the device and the binning configuration (tile shape, bin capacity).
    Scratch memory is taken from PyTorch's caching allocator per call, so one
    context may be shared by threads as diffrp's per-thread cache does.
    ``tile`` / ``bin_capacity`` = None lets the heuristic (or an autotuned
    table installed with ``set_tuning``) choose per call shape.  Both affect
    speed only; a capacity above four rounds of a tile's pixels (learnt for
    meshes whose tiles hold that many triangles) also selects fine-raster
    kernels that cull triangles behind a tile's current depths.

    ``prepare_antialias=True`` (an option of this implementation): every
    single-layer ``rasterize`` of this context also prepares what a following
    ``antialias`` of the same ``pos`` and ``tri`` computes per vertex and per
    triangle (pixel positions, screen orientations, silhouette flags), on the
    library's auxiliary stream beside the fine raster pass and joined before
    the rasterize call's work ends; ``antialias`` then starts at its pair
    search.  The results are
    the same; the preparation is wasted work when no such ``antialias`` follows.
    """

    def __init__(self, device: Optional[Union[int, str, torch.device]] = None, tile=None, bin_capacity=None,
                 prepare_antialias=False):
        if device is None:
            device = torch.cuda.current_device()
        self.device = torch.device('cuda', device) if isinstance(device, int) else torch.device(device)
        self.prepare_antialias = bool(prepare_antialias)
        if tile is not None and tuple(tile) not in _TILES:
            raise ValueError(f'tile must be one of {sorted(_TILES)}')
        self.tile = None if tile is None else tuple(tile)
        self.bin_capacity = bin_capacity
        self._tuned = OrderedDict()    # shape -> (tile, capacity), least recently used first
        self._pending = OrderedDict()  # shape -> (tile, capacity, pinned peak bin count, event, tiles, images)
        self._lock = threading.Lock()
        _lib.lib()

    MAX_TUNED = 1024    # remembered shapes (least recently used are forgotten)
    MAX_PENDING = 64    # shapes waiting for their first read-back

    def set_tuning(self, key, tile, bin_capacity):
        """Install a tuned (tile, bin capacity) for shapes keyed by ``tuning_key``."""
        with self._lock:
            self._pending.pop(key, None)  # an installed tuning replaces what the first call would teach
            self._remember(key, (tuple(tile), int(bin_capacity)))

    def _remember(self, key, val):
        self._tuned[key] = val
        self._tuned.move_to_end(key)
        while len(self._tuned) > self.MAX_TUNED:
            self._tuned.popitem(last=False)

    @staticmethod
    def tuning_key(n_img, H, W, F):
        return (int(n_img), int(H), int(W), int(F))

    def config(self, n_img, H, W, F) -> Tuple[Tuple[int, int], int]:
        key = self.tuning_key(n_img, H, W, F)
        if self.tile is None and self.bin_capacity is None:
            with self._lock:
                self._settle_all()
                ent = self._tuned.get(key)
                if ent is not None:
                    self._tuned.move_to_end(key)
                    return ent
        tile = self.tile or default_tile(H, W, F)
        cap = self.bin_capacity or default_bin_capacity(tile, H, W, F)
        return tile, cap

    # Capacity learnt from the first eager call of a shape: the per-tile bin
    # counts are exact whatever the capacity, so their maximum, read back
    # without a host wait at a later call, sizes the bins of the calls that
    # follow (a crumpled mesh can load its tiles far beyond the heuristic, and
    # spilled triangles are scanned by every tile of the image).  The learnt
    # capacity is bounded by a memory budget (learnt_bin_capacity).
    def _watch(self, key, tile, cap, counts: torch.Tensor, n_img, F):
        if (self.tile is not None or self.bin_capacity is not None or counts.numel() == 0
                or torch.cuda.is_current_stream_capturing()):
            return
        with self._lock:
            if key in self._tuned or key in self._pending:
                return
            while len(self._pending) >= self.MAX_PENDING:
                self._pending.popitem(last=False)  # the oldest shape waits for a later first call
            st = torch.cuda.current_stream(counts.device)
            with torch.cuda.stream(st):
                host = torch.empty(1, dtype=torch.int32, pin_memory=True)
                host.copy_(counts.max().view(1), non_blocking=True)
                ev = torch.cuda.Event()
                ev.record(st)
            self._pending[key] = (tile, cap, host, ev, counts.numel(), n_img * F)

    def _settle_all(self):
        """Settle every pending shape whose read-back has completed (caller holds the lock)."""
        if not self._pending or torch.cuda.is_current_stream_capturing():
            return
        for key in [k for k, ent in self._pending.items() if ent[3].query()]:
            tile, cap, host, _, n_tiles, items = self._pending.pop(key)
            if key not in self._tuned:  # a tuning installed meanwhile wins
                self._remember(key, (tile, learnt_bin_capacity(int(host[0]), cap, n_tiles, items)))


def learnt_bin_capacity(peak: int, cap: int, n_tiles: int, items: int) -> int:
    """Bin capacity for the calls after a first call with capacity ``cap`` whose
    busiest tile held ``peak`` entries: the next power of two above peak + 15 %,
    at most 65536, and at most what a budget of 256 MiB plus 64 bytes per
    (image, triangle) gives every tile (all tiles share one capacity, so one
    dense tile must not size the bins of a large image); never below ``cap``."""
    if peak <= cap:
        return cap
    want = 32
    while want < int(peak * 1.15) + 1 and want < (1 << 16):
        want *= 2
    budget = (256 << 20) + 64 * int(items)
    limit = 32
    while limit * 2 * 4 * max(int(n_tiles), 1) <= budget and limit < (1 << 16):
        limit *= 2
    return max(cap, min(want, limit))


RasterizeHipContext = RasterizeCudaContext


class RasterizeGLContext:
    """OpenGL contexts do not exist on this platform; the name is kept for type annotations
    (``surface_deferred.py:209``)."""

    def __init__(self, *args, **kwargs):
        raise RuntimeError('RasterizeGLContext is not available on MI355X; use RasterizeCudaContext')


def default_tile(H, W, F):
    return (16, 16) if H * W >= 256 * 256 else (8, 8)


def default_bin_capacity(tile, H, W, F):
    # a closed mesh seen from outside puts ~2 F (tiles per triangle) / tiles
    # triangles in a tile on average; silhouette tiles reach ~3x the mean.
    # Overflow only costs speed; the autotuner sizes this from measured loads.
    tiles = max(1, ((W + tile[0] - 1) // tile[0]) * ((H + tile[1] - 1) // tile[1]))
    est = 6.0 * F / tiles + 32
    cap = 32
    while cap < est and cap < 4096:
        cap *= 2
    return cap


_log_level = 1


def get_log_level():
    return _log_level


def set_log_level(level):
    global _log_level
    _log_level = level


# ----------------------------------------------------------------------------
# argument handling
# ----------------------------------------------------------------------------
def _as_f32(t, name):
    if t.dtype != torch.float32:
        raise TypeError(f'{name} must be float32 (got {t.dtype})')
    t = t.contiguous()
    if t.data_ptr() % 16:
        t = t.clone()
    return t


def _as_tri(tri):
    if tri.dtype != torch.int32:
        raise TypeError(f'tri must be int32 (got {tri.dtype})')
    if tri.dim() != 2 or tri.shape[1] != 3:
        raise ValueError(f'tri must have shape [F, 3] (got {tuple(tri.shape)})')
    return tri.contiguous()


class _Ranges:
    """Range-mode table uploaded to the device: (first, count, spill-list offset) per image."""

    def __init__(self, ranges, F, device):
        r = torch.as_tensor(ranges)
        if r.is_cuda:
            raise ValueError('ranges must be a CPU tensor (as in nvdiffrast)')
        r = r.to(torch.int64).reshape(-1, 2).numpy()
        start = np.clip(r[:, 0], 0, F)
        count = np.clip(r[:, 0] + r[:, 1], 0, F) - start
        count = np.maximum(count, 0)
        off = np.concatenate([[0], np.cumsum(count)[:-1]]) if len(count) else np.zeros(0, np.int64)
        tab = np.stack([r[:, 0], r[:, 1], off], 1).astype(np.int32)
        self.n = len(tab)
        self.list_total = int(count.sum())
        self.max_range = int(count.max()) if len(count) else 0
        self.dev = torch.from_numpy(tab).to(device)


# ----------------------------------------------------------------------------
# rasterize
# ----------------------------------------------------------------------------
def _rasterize_launch(ctx, pos, tri, H, W, rng, prev, want_db, aa=None):
    batched = pos.dim() == 3
    n = pos.shape[0] if batched else (rng.n if rng is not None else 1)
    V = pos.shape[-2]
    F = tri.shape[0]
    dev = pos.device
    rast = torch.empty((n, H, W, 4), dtype=torch.float32, device=dev)
    rast_db = torch.empty((n, H, W, 4), dtype=torch.float32, device=dev) if want_db else None
    boxes = None
    if n == 0:
        return rast, rast_db, boxes
    if rng is None and F > 0 and max(H, W) <= 65535:
        boxes = torch.empty((n, F, 3), dtype=torch.int32, device=dev)
    (tw, th), cap = ctx.config(n, H, W, F)
    list_total = rng.list_total if rng is not None else n * F
    max_range = rng.max_range if rng is not None else F
    L = _lib.lib()
    ws_bytes = L.drp_rasterize_workspace_size(n, H, W, list_total, tw, th, cap)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    if aa is not None:  # the antialias preparation (aa = (workspace, opp)) beside the fine raster
        rc = L.drp_rasterize_fwd_aa_prepare(stream_of(pos), ptr(pos), n, V, int(batched), ptr(tri), F, H, W,
                                            ptr(rast), ptr(rast_db), ptr(ws), ws_bytes, tw, th, cap, ptr(boxes),
                                            ptr(aa[1]), ptr(aa[0]), aa[0].numel())
    else:
        rc = L.drp_rasterize_fwd(stream_of(pos), ptr(pos), n, V, int(batched), ptr(tri), F,
                                 ptr(rng.dev) if rng is not None else None, H, W, ptr(prev), ptr(rast), ptr(rast_db),
                                 ptr(ws), ws_bytes, tw, th, cap, list_total, max_range, ptr(boxes))
    check(rc, 'rasterize')
    n_tiles = n * ((W + tw - 1) // tw) * ((H + th - 1) // th)
    ctx._watch(ctx.tuning_key(n, H, W, F), (tw, th), cap, ws[:4 * n_tiles].view(torch.int32), n, F)  # bin counts
    return rast, rast_db, boxes


# vertex -> incident corner tables, reused per triangle tensor (identity + version)
_inc_cache = {}


def incidence(tri: torch.Tensor, V: int):
    """CSR (inc_off [V+1], inc [3F]) of the corners incident to each vertex, built on the GPU."""
    key = (id(tri), int(V))
    ent = _inc_cache.get(key)
    if ent is not None and ent[0]() is tri and ent[1] == tri._version:
        return ent[2]
    F = tri.shape[0]
    off = torch.empty(V + 1, dtype=torch.int32, device=tri.device)
    inc = torch.empty(max(3 * F, 1), dtype=torch.int32, device=tri.device)
    L = _lib.lib()
    ws_bytes = L.drp_incidence_workspace_size(V, F)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=tri.device)
    check(L.drp_incidence_build(stream_of(tri), ptr(tri), F, V, ptr(off), ptr(inc), ptr(ws), ws_bytes), 'incidence')
    _inc_cache[key] = (weakref.ref(tri, lambda _r, k=key: _inc_cache.pop(k, None)), tri._version, (off, inc))
    return off, inc


class _RasterizeFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, raster_ctx, pos, tri, H, W, rng, prev, want_db, holder, aa):
        rast, rast_db, boxes = _rasterize_launch(raster_ctx, pos, tri, H, W, rng, prev, want_db, aa)
        fctx.save_for_backward(pos, tri, rast, boxes)
        fctx.rng = rng
        fctx.HW = (H, W)
        fctx.want_db = want_db
        fctx.set_materialize_grads(False)  # an unused rast_db brings no gradient
        if rast_db is None:
            rast_db = torch.empty((rast.shape[0], H, W, 0), dtype=torch.float32, device=pos.device)
        fctx.mark_non_differentiable(*(() if want_db else (rast_db,)))
        holder.append(boxes)
        return rast, rast_db

    @staticmethod
    def backward(fctx, dy, ddb):
        pos, tri, rast, boxes = fctx.saved_tensors
        H, W = fctx.HW
        rng = fctx.rng
        gdb_used = fctx.want_db and ddb is not None and ddb.numel() > 0
        if dy is None and not gdb_used:
            return None, None, None, None, None, None, None, None, None, None
        gpos = torch.zeros_like(pos)
        dy = _as_f32(dy, 'grad_rast') if dy is not None else torch.zeros_like(rast)
        gdb = _as_f32(ddb, 'grad_db') if gdb_used else None
        n, V, F = rast.shape[0], pos.shape[-2], tri.shape[0]
        L = _lib.lib()
        off = inc = ws = None
        ws_bytes = 0
        if boxes is not None:
            off, inc = incidence(tri, V)
            ws_bytes = L.drp_rasterize_bwd_workspace_size(n, H, W, F)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=pos.device)
        rc = L.drp_rasterize_bwd(stream_of(pos), ptr(pos), n, V, int(pos.dim() == 3), ptr(tri), F,
                                 ptr(rng.dev) if rng is not None else None, H, W, ptr(rast), ptr(dy), ptr(gdb),
                                 ptr(gpos), ptr(boxes), ptr(off), ptr(inc), ptr(ws), ws_bytes)
        check(rc, 'rasterize backward')
        return None, gpos, None, None, None, None, None, None, None, None


def _rasterize_impl(glctx, pos, tri, resolution, ranges, grad_db, prev):
    if not isinstance(glctx, RasterizeCudaContext):
        raise TypeError('glctx must be a RasterizeCudaContext')
    require_device(pos, tri, names=('pos', 'tri'))
    H, W = int(resolution[0]), int(resolution[1])
    pos = _as_f32(pos, 'pos')
    tri = _as_tri(tri)
    if pos.dim() not in (2, 3) or pos.shape[-1] != 4:
        raise ValueError(f'pos must have shape [V, 4] or [B, V, 4] (got {tuple(pos.shape)})')
    rng = None
    if pos.dim() == 2:
        if ranges is not None:
            r = torch.as_tensor(ranges).reshape(-1, 2)
            # one range over every triangle (diffrp's call, surface_deferred.py:259) is the plain case
            if not (r.shape[0] == 1 and int(r[0, 0]) == 0 and int(r[0, 1]) >= tri.shape[0]):
                rng = _Ranges(ranges, tri.shape[0], pos.device)
    elif ranges is not None:
        raise ValueError('ranges are only valid with 2-D pos (range mode)')
    holder = []
    prep = _aa_prepare_args(pos, tri, H, W) if glctx.prepare_antialias and rng is None and prev is None else None
    rast, rast_db = _RasterizeFunc.apply(glctx, pos, tri, H, W, rng, prev, bool(grad_db), holder,
                                         None if prep is None else prep[:2])
    if prep is not None:
        # the next antialias of this rast, positions and triangles starts at its pair search
        rast._drp_aa_prep = _AAPrep(prep[0], pos, tri, prep[1], prep[2], rast._version)
    if holder and holder[0] is not None:
        # pixel boxes of this rasterisation, for the triangle-centric interpolate
        # backward, and the positions and triangles it rasterised (when
        # differentiable: an interpolate of the same triangles can then run the
        # rasterize backward in the same pass)
        rast._drp_boxes = (holder[0], (tri.data_ptr(), tuple(tri.shape), tri._version), rast._version,
                           pos if rast.requires_grad else None)
    return rast, rast_db


def _aa_prepare_args(pos, tri, H, W):
    """The workspace, topology and shape of the antialias preparation that the
    rasterize call runs beside its fine raster (``drp_rasterize_fwd_aa_prepare``)."""
    n = pos.shape[0] if pos.dim() == 3 else 1
    batched = int(pos.dim() == 3)
    V, F = pos.shape[-2], tri.shape[0]
    if F == 0 or V == 0 or n == 0:
        return None
    opp = _cached_topology(tri).opp
    ws_bytes = _lib.lib().drp_antialias_workspace_size(n, H, W, V, F, batched)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=pos.device)
    return ws, opp, (n, H, W, V, F, batched)


class _AAPrep:
    """An antialias workspace prepared for (pos, tri, opp, shape) and the rast
    it belongs to."""
    __slots__ = ('ws', 'pos', 'pos_version', 'tri', 'opp', 'shape', 'rast_version')

    def __init__(self, ws, pos, tri, opp, shape, rast_version):
        self.ws, self.pos, self.pos_version = ws, pos, pos._version
        self.tri, self.opp, self.shape, self.rast_version = tri, opp, shape, rast_version


def rasterize(glctx, pos, tri, resolution, ranges=None, grad_db=True):
    """Rasterize triangles (``nvdiffrast.torch.rasterize``).

    pos [V,4] (range mode, ``ranges`` CPU int32 [n,2]) or [B,V,4] (instanced);
    tri [F,3] int32.  Returns ``(rast, rast_db)``: rast [n,H,W,4] =
    (u, v, z/w, triangle id + 1), bottom row first; rast_db the pixel-space
    barycentric derivatives when ``grad_db`` (else an empty [n,H,W,0] tensor).
    """
    return _rasterize_impl(glctx, pos, tri, resolution, ranges, grad_db, None)


class DepthPeeler:
    """Depth peeling (``nvdiffrast.torch.DepthPeeler``, used at ``surface_deferred.py:260-265``).

    Each ``rasterize_next_layer()`` returns the nearest fragments strictly
    behind the previous layer's (by the (z/w, triangle) order).
    """

    def __init__(self, glctx, pos, tri, resolution, ranges=None, grad_db=True):
        self.glctx, self.pos, self.tri = glctx, pos, tri
        self.resolution, self.ranges, self.grad_db = resolution, ranges, grad_db
        self.prev = None
        self.active = False

    def __enter__(self):
        if self.active:
            raise RuntimeError('DepthPeeler is already active')
        self.active = True
        self.prev = None
        return self

    def __exit__(self, *args):
        self.active = False
        self.prev = None
        return False

    def rasterize_next_layer(self):
        if not self.active:
            raise RuntimeError('DepthPeeler must be used as a context manager')
        prev = None if self.prev is None else self.prev.detach()
        rast, rast_db = _rasterize_impl(self.glctx, self.pos, self.tri, self.resolution, self.ranges, self.grad_db,
                                        prev)
        self.prev = rast
        return rast, rast_db


# ----------------------------------------------------------------------------
# interpolate
# ----------------------------------------------------------------------------
def _attr_layout(attr, rast):
    if attr.dim() == 2:
        return 0, attr.shape[0], attr.shape[1]
    if attr.dim() == 3:
        if attr.shape[0] != rast.shape[0]:
            raise ValueError('batched attr must have one attribute buffer per image')
        return 1, attr.shape[1], attr.shape[2]
    raise ValueError(f'attr must have shape [V, C] or [n, V, C] (got {tuple(attr.shape)})')


_TRI_BWD_C = (1, 2, 3, 4, 8, 16)


def _views_per_group(n, F):
    """Views folded into one partial of the triangle-centric attribute backward
    (a power of two up to 8: fewer partials to gather, blocks still fill the GPU)."""
    p = 1
    while p * 2 <= min(n, 8):
        p *= 2
    return p


_FUSED_C = (4, 8)


class _InterpolateFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, attr, rast, tri, rast_db, diff_idx, boxes, src_pos):
        batched, V, C = _attr_layout(attr, rast)
        n, H, W, _ = rast.shape
        out = torch.empty((n, H, W, C), dtype=torch.float32, device=attr.device)
        nd = 0 if diff_idx is None else diff_idx.numel()
        out_da = torch.empty((n, H, W, 2 * nd), dtype=torch.float32, device=attr.device)
        rc = _lib.lib().drp_interpolate_fwd(stream_of(attr), ptr(attr), n, V, batched, C, ptr(rast), H, W, ptr(tri),
                                            tri.shape[0], ptr(rast_db) if nd else None,
                                            ptr(diff_idx) if nd else None, nd, ptr(out), ptr(out_da) if nd else None)
        check(rc, 'interpolate')
        fused = (src_pos is not None and boxes is not None and nd == 0 and not batched and C in _FUSED_C
                 and src_pos.shape[-2] == V and tri.shape[0] > 0)
        fctx.save_for_backward(attr, rast, tri, rast_db if nd else None, diff_idx if nd else None, boxes,
                               src_pos if fused else None)
        fctx.nd = nd
        # an antialias of these positions downstream may leave its position
        # gradient here (_AntialiasFunc): this node's backward adds its own to it
        fctx._drp_fused_pos = src_pos if fused else None
        fctx._drp_aa_pos = None
        if not nd:
            fctx.mark_non_differentiable(out_da)
        return out, out_da

    @staticmethod
    def backward(fctx, dy, dda):
        attr, rast, tri, rast_db, diff_idx, boxes, src_pos = fctx.saved_tensors
        batched, V, C = _attr_layout(attr, rast)
        n, H, W, _ = rast.shape
        F = tri.shape[0]
        nd = fctx.nd
        dy = _as_f32(dy, 'grad_out')
        gattr = torch.zeros_like(attr)
        L = _lib.lib()
        aa_gpos = _take_aa_pos(fctx)
        if src_pos is not None and fctx.needs_input_grad[6]:
            # the rasterize backward of this interpolation's grad_rast in the same
            # pass, returned as the gradient of the rasterised positions; grad_rast
            # itself is neither written nor returned, so the rasterize backward
            # (linear in its grad_rast) sees only what rast's other consumers send
            gpos = aa_gpos if aa_gpos is not None else torch.zeros_like(src_pos)  # (accumulated into)
            off, inc = incidence(tri, V)
            vpg = _views_per_group(n, F)
            ws_bytes = L.drp_interpolate_rasterize_bwd_workspace_size(n, H, W, F, C, vpg)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=attr.device)
            rc = L.drp_interpolate_rasterize_bwd(stream_of(attr), ptr(attr), n, V, C, ptr(rast), H, W, ptr(tri), F,
                                                 ptr(src_pos), int(src_pos.dim() == 3), ptr(dy), ptr(gattr), None,
                                                 ptr(gpos), ptr(boxes), vpg, ptr(off), ptr(inc), ptr(ws), ws_bytes)
            check(rc, 'interpolate + rasterize backward')
            return gattr, None, None, None, None, None, gpos
        grast = torch.empty_like(rast)
        gdb = torch.empty_like(rast) if nd else None
        dda = _as_f32(dda, 'grad_da') if (nd and dda is not None) else None
        if nd and dda is None:
            dda = torch.zeros((n, H, W, 2 * nd), dtype=torch.float32, device=attr.device)
        off = inc = ws = None
        ws_bytes = vpg = 0
        if boxes is not None and nd == 0 and C in _TRI_BWD_C:
            off, inc = incidence(tri, V)
            vpg = _views_per_group(n, F)
            ws_bytes = L.drp_interpolate_bwd_workspace_size(n, H, W, F, C, batched, vpg)
            ws = torch.empty(ws_bytes, dtype=torch.uint8, device=attr.device)
        else:
            boxes = None
        rc = L.drp_interpolate_bwd(stream_of(attr), ptr(attr), n, V, batched, C, ptr(rast), H, W, ptr(tri), F,
                                   ptr(rast_db), ptr(diff_idx), nd, ptr(dy), ptr(dda), ptr(gattr), ptr(grast),
                                   ptr(gdb), ptr(boxes), vpg, ptr(off), ptr(inc), ptr(ws), ws_bytes)
        check(rc, 'interpolate backward')
        return gattr, grast, None, gdb, None, None, aa_gpos


def interpolate(attr, rast, tri, rast_db=None, diff_attrs=None):
    """Interpolate vertex attributes (``nvdiffrast.torch.interpolate``, ``interpolator.py:83``).

    attr [V,C] or [n,V,C]; rast [n,H,W,4]; tri [F,3] int32.  ``diff_attrs``
    ('all' or a list of channel indices) with ``rast_db`` adds out_da
    [n,H,W,2*len] = (dA/dX, dA/dY) per listed channel.  Returns ``(out, out_da)``.
    """
    require_device(attr, rast, tri, rast_db, names=('attr', 'rast', 'tri', 'rast_db'))
    attr = _as_f32(attr, 'attr')
    rast = _as_f32(rast, 'rast')
    tri = _as_tri(tri)
    if rast.dim() != 4 or rast.shape[-1] != 4:
        raise ValueError(f'rast must have shape [n, H, W, 4] (got {tuple(rast.shape)})')
    diff_idx = None
    if diff_attrs is not None:
        C = attr.shape[-1]
        idx = list(range(C)) if diff_attrs == 'all' else [int(i) for i in diff_attrs]
        if any(i < 0 or i >= C for i in idx):
            raise ValueError('diff_attrs index out of range')
        if rast_db is None or rast_db.numel() == 0:
            raise ValueError('diff_attrs requires rast_db (rasterize with grad_db=True)')
        if len(idx):
            diff_idx = torch.tensor(idx, dtype=torch.int32, device=attr.device)
            rast_db = _as_f32(rast_db, 'rast_db')
    boxes = src_pos = None
    meta = getattr(rast, '_drp_boxes', None)
    # boxes describe where each triangle id's pixels lie: valid while rast is
    # unmodified, for any triangle list of the same length; the positions'
    # chain only for the triangles that were rasterised
    if meta is not None and meta[1][1] == tuple(tri.shape) and meta[2] == rast._version and \
            meta[0].shape[0] == rast.shape[0]:
        boxes = meta[0]
        if meta[1] == (tri.data_ptr(), tuple(tri.shape), tri._version):
            src_pos = meta[3]
    return _InterpolateFunc.apply(attr, rast, tri, rast_db if diff_idx is not None else None, diff_idx, boxes,
                                  src_pos)


# ----------------------------------------------------------------------------
# antialias
# ----------------------------------------------------------------------------
class TopologyHash:
    """Edge -> opposite-triangle table of a triangle list (``antialias_construct_topology_hash``)."""

    def __init__(self, opp: torch.Tensor, F: int):
        self.opp = opp
        self.F = F


def _build_topology(tri):
    F = tri.shape[0]
    opp = torch.empty((F, 3), dtype=torch.int32, device=tri.device)
    L = _lib.lib()
    ws_bytes = L.drp_topology_workspace_size(F)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=tri.device)
    check(L.drp_topology_build(stream_of(tri), ptr(tri), F, ptr(opp), ptr(ws), ws_bytes), 'topology hash')
    return TopologyHash(opp, F)


def antialias_construct_topology_hash(tri):
    require_device(tri, names=('tri',))
    tri = _as_tri(tri)
    th = _build_topology(tri)
    _remember_topology(tri, th)  # antialias of this tri without a hash, and its preparation, reuse it
    return th


# per-tensor reuse of the topology table when antialias is called without one
# (diffrp never passes it, surface_deferred.py:775); keyed by tensor identity
# and its in-place version counter
_topo_cache = {}


def _remember_topology(tri: torch.Tensor, th):
    key = id(tri)
    _topo_cache[key] = (weakref.ref(tri, lambda _r, k=key: _topo_cache.pop(k, None)), tri._version, th)


def _cached_topology(tri: torch.Tensor):
    ent = _topo_cache.get(id(tri))
    if ent is not None and ent[0]() is tri and ent[1] == tri._version:
        return ent[2]
    th = _build_topology(tri)
    _remember_topology(tri, th)
    return th


# hand the antialias position gradient to a fused interpolation (tests switch it
# off); it needs autograd's id of the running backward pass to tell passes apart
_graph_task_id = getattr(torch._C, '_current_graph_task_id', None)
_HAND_AA_POS = _graph_task_id is not None


def _fused_host(color, pos, limit=64):
    """A fused interpolation among the autograd ancestors of ``color`` (at
    most ``limit`` nodes searched) whose rasterised positions are ``pos``: its
    backward runs after the antialias backward of ``color`` (the gradient it
    needs passes through there) and returns ``pos``'s gradient, so it can add
    the antialias's share to its own instead of autograd adding the two."""
    stack, seen = [color.grad_fn], set()
    while stack and len(seen) < limit:
        f = stack.pop()
        if f is None or id(f) in seen:
            continue
        seen.add(id(f))
        if getattr(f, '_drp_fused_pos', None) is pos:
            return f
        stack.extend(g for g, _ in f.next_functions)
    return None


def _take_aa_pos(fctx):
    """The antialias position gradient left on an interpolation node in this
    backward pass (None if none, or if it was left by an earlier pass that did
    not reach this node)."""
    dep, fctx._drp_aa_pos = fctx._drp_aa_pos, None
    return dep[0] if dep is not None and dep[1] == _graph_task_id() else None


class _AntialiasFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, color, rast, pos, tri, opp, boost):
        n, H, W, C = color.shape
        batched = int(pos.dim() == 3)
        V, F = pos.shape[-2], tri.shape[0]
        out = torch.empty_like(color)
        L = _lib.lib()
        ws_bytes = L.drp_antialias_workspace_size(n, H, W, V, F, batched)
        prep = getattr(rast, '_drp_aa_prep', None)
        if (prep is not None and prep.pos is pos and prep.pos_version == pos._version and prep.tri is tri
                and prep.opp is opp and prep.shape == (n, H, W, V, F, batched)
                and prep.rast_version == rast._version):
            # the rasterize call prepared this workspace (same positions, triangles and size)
            ws, fwd = prep.ws, L.drp_antialias_fwd_prepared
            rast._drp_aa_prep = None  # used once: the workspace now keeps this call's pair decisions
        else:
            ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=color.device)
            fwd = L.drp_antialias_fwd
        rc = fwd(stream_of(color), ptr(color), n, H, W, C, ptr(rast), ptr(pos), V, batched, ptr(tri), F, ptr(opp),
                 ptr(out), ptr(ws), ws_bytes)
        check(rc, 'antialias')
        fctx.save_for_backward(color, rast, pos, tri, opp)
        fctx.boost = boost
        # the workspace keeps the pair decisions; the backward reads them
        # (saved tensors cannot change in place before backward, torch checks)
        fctx.ws = ws
        fctx.host = _fused_host(color, pos) if _HAND_AA_POS and fctx.needs_input_grad[2] else None
        return out

    @staticmethod
    def backward(fctx, dy):
        color, rast, pos, tri, opp = fctx.saved_tensors
        n, H, W, C = color.shape
        batched = int(pos.dim() == 3)
        V, F = pos.shape[-2], tri.shape[0]
        dy = _as_f32(dy, 'grad_out')
        gcol = torch.empty_like(color)
        host, gpos = fctx.host, None
        if fctx.needs_input_grad[2]:
            # with a fused interpolation of pos upstream, accumulate into the
            # gradient already left there in this pass (another antialias), and
            # leave the sum for that interpolation's backward to add to
            task = _graph_task_id() if host is not None else None
            dep = host._drp_aa_pos if host is not None else None
            gpos = dep[0] if dep is not None and dep[1] == task else torch.zeros_like(pos)
        L = _lib.lib()
        ws, reuse = fctx.ws, 1
        if ws is None:
            # a second backward through the same graph: the forward's workspace was
            # released by the first one, so the pair decisions are recomputed
            ws_bytes = L.drp_antialias_workspace_size(n, H, W, V, F, batched)
            ws, reuse = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=color.device), 0
        rc = L.drp_antialias_bwd(stream_of(color), ptr(color), n, H, W, C, ptr(rast), ptr(pos), V, batched, ptr(tri),
                                 F, ptr(opp), ptr(dy), ptr(gcol), ptr(gpos), float(fctx.boost), reuse, ptr(ws),
                                 ws.numel())
        fctx.ws = None  # released after its one reuse
        check(rc, 'antialias backward')
        if host is not None and gpos is not None:
            host._drp_aa_pos = (gpos, task)
            return gcol, None, None, None, None, None
        return gcol, None, gpos, None, None, None


def antialias(color, rast, pos, tri, topology_hash=None, pos_gradient_boost=1.0):
    """Silhouette antialiasing with visibility gradients (``nvdiffrast.torch.antialias``,
    ``surface_deferred.py:775``).

    color [n,H,W,C]; rast [n,H,W,4] (a single layer); pos [V,4] or [n,V,4];
    tri [F,3].  Differentiable w.r.t. ``color`` and ``pos``.
    """
    require_device(color, rast, pos, tri, names=('color', 'rast', 'pos', 'tri'))
    color = _as_f32(color, 'color')
    rast = _as_f32(rast, 'rast')
    pos = _as_f32(pos, 'pos')
    tri = _as_tri(tri)
    if color.dim() != 4 or rast.shape[:3] != color.shape[:3]:
        raise ValueError('color must be [n, H, W, C] matching rast [n, H, W, 4]')
    if pos.dim() == 3 and pos.shape[0] != color.shape[0]:
        raise ValueError('batched pos must have one buffer per image')
    if topology_hash is None:
        topology_hash = _cached_topology(tri)
    elif not isinstance(topology_hash, TopologyHash) or topology_hash.F != tri.shape[0]:
        raise ValueError('topology_hash does not match tri')
    return _AntialiasFunc.apply(color, rast, pos, tri, topology_hash.opp, float(pos_gradient_boost))


# ----------------------------------------------------------------------------
def install_as_nvdiffrast():
    """Register this module as ``nvdiffrast`` / ``nvdiffrast.torch`` so that
    ``import nvdiffrast.torch as dr`` (diffrp's import) resolves here."""
    mod = sys.modules[__name__]
    pkg = sys.modules.get('nvdiffrast')
    if pkg is None:
        pkg = types.ModuleType('nvdiffrast')
        pkg.__path__ = []
        sys.modules['nvdiffrast'] = pkg
    pkg.torch = mod
    sys.modules['nvdiffrast.torch'] = mod
    return mod
