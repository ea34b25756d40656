"""
Edge-gradient visibility operator ("Rasterized Edge Gradients: Handling
Discontinuities Differentiably"), as diffrp applies it
(``diffrp/rendering/surface_deferred.py:120-169, 777-804``).

Forward is the identity on the image.  Backward passes the image gradient
through and, at horizontally / vertically adjacent pixels with different
triangle ids, adds a gradient on the per-pixel screen position of the
fragment; that position is the perspective-interpolated clip position of the
pixel's triangle, so the gradient reaches the vertices through the usual
autograd graph.  The backward is one HIP kernel (``drp_edge_grad_bwd``).
"""
import torch

from . import _lib
from ._lib import check, ptr, require_device, stream_of
from .deferred import float_to_triidx, polyfill_interpolate


class EdgeGradientFunc(torch.autograd.Function):
    """(frag_xy [H,W,2], tris_xy [F,3,2], rast [H,W,4], rgb [H,W,C]) -> rgb;
    all bottom row first."""

    @staticmethod
    def forward(ctx, frag_xy, tris_xy, rast, rgb):
        ctx.save_for_backward(frag_xy, tris_xy, rast, rgb)
        return rgb.view_as(rgb)

    @staticmethod
    def backward(ctx, grad_out):
        frag_xy, tris_xy, rast, rgb = ctx.saved_tensors
        H, W, C = rgb.shape
        grad_out = grad_out.contiguous()
        frag = frag_xy.detach().contiguous()
        txy = tris_xy.detach().contiguous()
        gfrag = torch.empty((H, W, 2), dtype=torch.float32, device=rgb.device)
        rc = _lib.lib().drp_edge_grad_bwd(stream_of(rgb), ptr(grad_out), ptr(rgb.detach().contiguous()), H, W, C,
                                          ptr(rast.detach().contiguous()), ptr(frag), ptr(txy), txy.shape[0],
                                          ptr(gfrag))
        check(rc, 'edge gradient backward')
        return gfrag, None, None, grad_out.clone()


def edge_gradient(rgb: torch.Tensor, rast: torch.Tensor, clip: torch.Tensor, tris: torch.Tensor) -> torch.Tensor:
    """Identity on ``rgb`` [H, W, C] (top row first) with edge visibility gradients into
    ``clip`` [V, 4] through the fragment positions of ``rast`` [1, H, W, 4]."""
    require_device(rgb, rast, clip, tris, names=('rgb', 'rast', 'clip', 'tris'))
    r = rast.squeeze(0)
    ids = float_to_triidx(r[..., 3])
    frag = polyfill_interpolate(clip, r, tris[ids - 1].long(), 1.0)
    frag_xy = frag[..., :2] / frag[..., 3:4]
    tris_xy = (clip[..., :2] / clip[..., 3:4])[tris.long()]
    return torch.flipud(EdgeGradientFunc.apply(frag_xy, tris_xy, r, torch.flipud(rgb).contiguous()))
