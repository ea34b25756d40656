"""
Image L2 loss of a fitting step, ``sum((a - b)^2)``, on the HIP kernels of
``drp_loss.hip``: the forward reads both images once (a deterministic sum,
accumulated in double), the backward reads them once more and writes
``2 g (a - b)``.  The same value as ``((a - b) ** 2).sum()`` /
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
        check(L.drp_l2_loss_fwd(stream_of(a), ptr(a), ptr(b), a.numel(), ptr(loss), ptr(ws), ws_bytes), 'l2 loss')
        ctx.save_for_backward(a, b)
        return loss

    @staticmethod
    def backward(ctx, g):
        a, b = ctx.saved_tensors
        g = g.float().contiguous()
        ga = torch.empty_like(a) if ctx.needs_input_grad[0] else None
        gb = torch.empty_like(b) if ctx.needs_input_grad[1] else None
        if ga is None and gb is None:
            return None, None
        check(_lib.lib().drp_l2_loss_bwd(stream_of(a), ptr(a), ptr(b), a.numel(), ptr(g), ptr(ga), ptr(gb)),
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
