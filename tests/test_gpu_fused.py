"""The interpolate backward that also runs the rasterize backward of its
grad_rast (one pass over the pixels) gives the same gradients, bit for bit, as
the two backward passes one after the other.  It returns the position
gradient itself; gradients that rast receives from other consumers still go
through the rasterize backward, and an interpolation of another triangle list
(same length) is not fused."""
import pytest
import torch

from diffrp_amd import _lib, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


class _Count:
    """Counts the calls of one C-ABI function of the loaded library."""

    def __init__(self, name):
        self.L, self.name = _lib.lib(), name
        self.fn = getattr(self.L, name)
        self.n = 0

    def __enter__(self):
        def wrapped(*a):
            self.n += 1
            return self.fn(*a)
        setattr(self.L, self.name, wrapped)
        return self

    def __exit__(self, *exc):
        setattr(self.L, self.name, self.fn)


def _grads(pos, tri, attr, go, res, fused, twice=False, db=False, depth_w=None, itri=None):
    saved = nvdr._FUSED_C
    nvdr._FUSED_C = saved if fused else ()
    try:
        p = pos.clone().requires_grad_(True)
        a = attr.clone().requires_grad_(True)
        rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tri, res, grad_db=db)
        out, _ = nvdr.interpolate(a, rast, tri if itri is None else itri)
        loss = (out * go).sum()
        if twice:
            loss = loss + (nvdr.interpolate(a * 0.5, rast, tri)[0] * go).sum()
        if depth_w is not None:  # a second consumer of rast: its depth channel
            loss = loss + (rast[..., 2] * depth_w).sum()
        with _Count('drp_rasterize_bwd') as c:
            loss.backward()
        return p.grad, a.grad, c.n
    finally:
        nvdr._FUSED_C = saved


@pytest.mark.parametrize('views,C,res', [(1, 8, (96, 128)), (5, 4, (64, 64)), (3, 8, (200, 150))])
def test_fused_backward_equals_separate_passes(cuda, views, C, res):
    pos, tri, _ = noisy_clip(4, 0.01, res[0], views=views)
    pos = (pos[0] if views == 1 else pos).cuda().contiguous()
    tri = tri.cuda()
    g = torch.Generator().manual_seed(90)
    attr = torch.rand(pos.shape[-2], C, generator=g).cuda()
    go = torch.randn(views, res[0], res[1], C, generator=g).cuda()
    gp_f, ga_f, calls_f = _grads(pos, tri, attr, go, res, True)
    gp_s, ga_s, calls_s = _grads(pos, tri, attr, go, res, False)
    assert calls_f == 0 and calls_s == 1   # the fused pass replaced the separate rasterize backward
    assert torch.equal(ga_f, ga_s)
    assert torch.equal(gp_f, gp_s)
    assert gp_f.abs().sum() > 0


def test_two_interpolations_of_one_rast_each_fuse(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 96)
    pos, tri = pos.cuda().contiguous(), tri.cuda()
    g = torch.Generator().manual_seed(91)
    attr = torch.rand(pos.shape[0], 8, generator=g).cuda()
    go = torch.randn(1, 96, 96, 8, generator=g).cuda()
    gp_f, ga_f, calls = _grads(pos, tri, attr, go, (96, 96), True, twice=True)
    gp_s, ga_s, _ = _grads(pos, tri, attr, go, (96, 96), False, twice=True)
    assert calls == 0   # each interpolation returns its own position gradient
    assert torch.equal(ga_f, ga_s)
    scale = float(gp_s.abs().max())
    assert float((gp_f - gp_s).abs().max()) <= 1e-6 * scale  # the same sums, summed by autograd in either order


def test_other_consumer_of_rast_goes_through_the_rasterize_backward(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 96)
    pos, tri = pos.cuda().contiguous(), tri.cuda()
    g = torch.Generator().manual_seed(93)
    attr = torch.rand(pos.shape[0], 8, generator=g).cuda()
    go = torch.randn(1, 96, 96, 8, generator=g).cuda()
    dw = torch.randn(1, 96, 96, generator=g).cuda()
    gp_f, ga_f, calls = _grads(pos, tri, attr, go, (96, 96), True, depth_w=dw)
    gp_s, ga_s, _ = _grads(pos, tri, attr, go, (96, 96), False, depth_w=dw)
    assert calls == 1   # the depth channel's gradient alone
    assert torch.equal(ga_f, ga_s)
    scale = float(gp_s.abs().max())
    assert float((gp_f - gp_s).abs().max()) <= 1e-6 * scale
    gp_i, _, _ = _grads(pos, tri, attr, go, (96, 96), True)
    assert float((gp_f - gp_i).abs().max()) > 0   # the depth term did reach the positions


def test_interpolation_of_another_triangle_list_is_not_fused(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 96)
    pos, tri = pos.cuda().contiguous(), tri.cuda()
    itri = tri[:, [1, 2, 0]].contiguous()  # same length, other corner order: an attribute index list
    g = torch.Generator().manual_seed(94)
    attr = torch.rand(pos.shape[0], 8, generator=g).cuda()
    go = torch.randn(1, 96, 96, 8, generator=g).cuda()
    gp_f, ga_f, calls = _grads(pos, tri, attr, go, (96, 96), True, itri=itri)
    gp_s, ga_s, _ = _grads(pos, tri, attr, go, (96, 96), False, itri=itri)
    assert calls == 1   # the positions' chain runs through the rasterize backward of the rasterised triangles
    assert torch.equal(ga_f, ga_s) and torch.equal(gp_f, gp_s)


def test_unused_rast_db_keeps_the_fused_pass(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 96)
    pos, tri = pos.cuda().contiguous(), tri.cuda()
    g = torch.Generator().manual_seed(92)
    attr = torch.rand(pos.shape[0], 4, generator=g).cuda()
    go = torch.randn(1, 96, 96, 4, generator=g).cuda()
    gp_f, ga_f, calls = _grads(pos, tri, attr, go, (96, 96), True, db=True)
    gp_s, ga_s, _ = _grads(pos, tri, attr, go, (96, 96), False, db=True)
    assert calls == 0
    assert torch.equal(gp_f, gp_s) and torch.equal(ga_f, ga_s)
