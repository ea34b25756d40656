"""
Depth-peeled layers composited front to back, with the option of recomputing
each layer in the backward pass instead of keeping its buffers.

diffrp's session peels up to ``max_layers`` layers (surface_deferred.py:240-266),
builds a G-buffer per layer (``gbuffer_collect``, :299-346), blends the shaded
layers (``compose_layers``, :348-395) and antialiases the result against the
first layer (``nvdr_antialias``, :758-775).  ``render_layers`` runs that chain
for one mesh with per-vertex attributes and a caller-provided shading
function.  With ``recompute=True`` each layer's rasterize + interpolate + shade
runs under ``torch.utils.checkpoint``: autograd keeps only the layer's inputs
(the previous layer's rast, 16 bytes per pixel) and re-runs the layer in the
backward pass, so the per-layer G-buffers (4 C bytes per pixel) and shaded
colours are never held for the whole graph.
"""
from typing import Callable, List, Optional, Tuple

import torch
import torch.utils.checkpoint as ckpt

from . import _lib, nvdr
from ._lib import check, ptr, stream_of


def count_covered(rast: torch.Tensor) -> torch.Tensor:
    """Covered pixels per image of rast [n, H, W, 4] as a device int32 [n]
    (no host synchronisation)."""
    if rast.dim() != 4 or rast.shape[-1] != 4 or rast.dtype != torch.float32 or not rast.is_contiguous():
        raise ValueError('rast must be a contiguous float32 [n, H, W, 4] tensor')
    n, H, W = rast.shape[0], rast.shape[1], rast.shape[2]
    counts = torch.empty(n, dtype=torch.int32, device=rast.device)
    check(_lib.lib().drp_count_covered(stream_of(rast), ptr(rast), n, H, W, ptr(counts)), 'count_covered')
    return counts


def peel_layers(glctx: 'nvdr.RasterizeCudaContext', pos: torch.Tensor, tri: torch.Tensor,
                resolution: Tuple[int, int], max_layers: int, ranges: Optional[torch.Tensor] = None,
                grad_db: bool = False) -> List[torch.Tensor]:
    """The layers of diffrp's peeling loop (``surface_deferred.py:258-266``):
    layer i is kept unless it is empty and i < max_layers - 1; the loop stops at
    the first such empty layer.  Emptiness is counted on the device
    (``count_covered``) and copied to pinned host memory; the host reads layer
    i's count only after layer i + 1 has been queued, so the GPU never waits for
    the host (the one speculative layer behind an empty one is dropped)."""
    layers: List[torch.Tensor] = []
    pending = None  # (rast, pinned counts or None, event)
    with nvdr.DepthPeeler(glctx, pos, tri, resolution, ranges, grad_db=grad_db) as dp:
        for i in range(max_layers):
            rast, _ = dp.rasterize_next_layer()
            host, ev = None, None
            if i < max_layers - 1:
                cnt = count_covered(rast)
                host = torch.empty(cnt.shape, dtype=cnt.dtype, pin_memory=True)
                host.copy_(cnt, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
            if pending is not None:
                prast, phost, pev = pending
                pev.synchronize()
                if int(phost.sum()) == 0:
                    return layers
                layers.append(prast)
            pending = (rast, host, ev)
    prast, phost, pev = pending
    if phost is not None:
        pev.synchronize()
        if int(phost.sum()) == 0:
            return layers
    layers.append(prast)
    return layers


def _layer(ctx, pos, tri, attr, resolution, prev, shade):
    rast, _ = nvdr._rasterize_impl(ctx, pos, tri, resolution, None, False, prev)
    g, _ = nvdr.interpolate(attr, rast, tri)
    rgba = shade(g, rast)
    covered = (rast[..., 3:4] > 0).to(rgba.dtype)
    return rast, rgba * covered


def render_layers(ctx: 'nvdr.RasterizeCudaContext', pos: torch.Tensor, tri: torch.Tensor, attr: torch.Tensor,
                  resolution: Tuple[int, int], n_layers: int,
                  shade: Callable[[torch.Tensor, torch.Tensor], torch.Tensor], recompute: bool = True,
                  antialias: bool = True):
    """RGB [1, H, W, 3] of ``n_layers`` peeled layers composited front to back
    (premultiplied 'over'), antialiased against the first layer.

    pos [V, 4] clip space, tri [F, 3] int32, attr [V, C]; shade(g, rast) maps a
    layer's interpolated attributes [1, H, W, C] to RGBA [1, H, W, 4]
    (straight alpha).  Differentiable w.r.t. pos and attr.
    """
    out = None
    trans = None
    prev = None
    first = None
    for k in range(n_layers):
        if recompute and torch.is_grad_enabled():
            rast, rgba = ckpt.checkpoint(_layer, ctx, pos, tri, attr, resolution, prev, shade, use_reentrant=False)
        else:
            rast, rgba = _layer(ctx, pos, tri, attr, resolution, prev, shade)
        rgb, a = rgba[..., :3], rgba[..., 3:4]
        if out is None:
            out, trans, first = rgb * a, 1.0 - a, rast
        else:
            out = out + trans * a * rgb
            trans = trans * (1.0 - a)
        prev = rast.detach()
    if antialias:
        out = nvdr.antialias(out.contiguous(), first, pos, tri)
    return out
