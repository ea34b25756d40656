"""
Image L2 loss of a fitting step, ``sum((a - b)^2)``, on the HIP kernels of
``drp_loss.hip``: the forward reads both images once (a deterministic sum,
accumulated in double).  When an input requires grad, the same pass also
writes ``2 (a - b)``, which the backward scales by the loss gradient g in
place (a pass over it only when g is not 1); a second backward through the
same graph reads both images again and writes ``2 g (a - b)``.  The same
value as ``((a - b) ** 2).sum()`` /
``torch.nn.functional.mse_loss(a, b, reduction='sum')`` without their
intermediate [a.shape] tensors.
"""
import torch

from . import _lib
from ._lib import check, ptr, stream_of, require_device


class _L2Sum(torch.autograd.Function):
    @staticmethod
    def forward(ctx, a, b):
        L = _lib.lib()
        loss = torch.empty((), dtype=torch.float32, device=a.device)
        ws_bytes = L.drp_l2_loss_workspace_size()
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device=a.device)
        ctx.unit_grad = None
        if ctx.needs_input_grad[0] or ctx.needs_input_grad[1]:
            # the gradient for g = 1, written in the same pass as the sum
            ctx.unit_grad = torch.empty_like(a)
            check(L.drp_l2_loss_fwd_grad(stream_of(a), ptr(a), ptr(b), a.numel(), ptr(loss), ptr(ctx.unit_grad),
                                         ptr(ws), ws_bytes), 'l2 loss')
        else:
            check(L.drp_l2_loss_fwd(stream_of(a), ptr(a), ptr(b), a.numel(), ptr(loss), ptr(ws), ws_bytes),
                  'l2 loss')
        ctx.save_for_backward(a, b)
        return loss

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors  # (raises if either changed in place since the forward)
        need_a, need_b = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        if not (need_a or need_b):
            return None, None
        g = g.float().contiguous()
        L = _lib.lib()
        ga = ctx.unit_grad
        ctx.unit_grad = None  # handed out once: a second backward recomputes
        if ga is not None:
            check(L.drp_l2_loss_scale(stream_of(a), ptr(ga), ga.numel(), ptr(g)), 'l2 loss backward')
            gb = torch.neg(ga) if need_b else None
            return (ga if need_a else None), gb
        ga = torch.empty_like(a) if need_a else None
        gb = torch.empty_like(b) if need_b else None
        check(L.drp_l2_loss_bwd(stream_of(a), ptr(a), ptr(b), a.numel(), ptr(g), ptr(ga), ptr(gb)),
              'l2 loss backward')
        return ga, gb


def l2_loss(a: torch.Tensor, b: torch.Tensor) -> torch.Tensor:
    """``sum((a - b)^2)`` over every element (float32 tensors of one shape, on the GPU)."""
    require_device(a, b, names=('a', 'b'))
    if a.shape != b.shape:
        raise ValueError(f'shapes differ: {tuple(a.shape)} vs {tuple(b.shape)}')
    if a.dtype != torch.float32 or b.dtype != torch.float32:
        raise TypeError('l2_loss takes float32 tensors')
    return _L2Sum.apply(a.contiguous(), b.contiguous())
