"""
Stencil-masked G-buffer path: the one-sort partition of a layer's pixels by
owning object equals diffrp's per-object ``(stencils[id] == k).nonzero()``
(order included), and the masked interpolator's HIP interpolation of the
gathered fragments equals diffrp's ``polyfill_interpolate``
(``interpolator.py:32-65``) forward (1e-6) and backward (1e-5 of the gradient
scale; the attribute scatter uses float atomics).
"""
import pytest
import torch

from diffrp_amd import deferred as dd, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _layer(res, views=None):
    pos, tri, _ = noisy_clip(5, 0.01, res)
    ctx = nvdr.RasterizeCudaContext()
    rast, _ = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), (res, res), grad_db=False)
    return pos, tri, rast


@pytest.mark.parametrize('n_obj', [1, 3, 40])
def test_stencil_partition_matches_nonzero(cuda, n_obj):
    pos, tri, rast = _layer(192)
    F = tri.shape[0]
    g = torch.Generator().manual_seed(30 + n_obj)
    stencils = torch.cat([torch.zeros(1, dtype=torch.int32),
                          torch.randint(1, n_obj + 1, (F,), generator=g, dtype=torch.int32)]).cuda()
    order, runs = dd.stencil_partition(rast, stencils, n_obj)
    owner = stencils[dd.float_to_triidx(rast[..., 3])]
    total = 0
    for k in range(n_obj + 1):
        s, e = runs[k]
        ref = (owner.reshape(-1) == k).nonzero().reshape(-1)
        assert torch.equal(order[s:e].long(), ref), f'owner {k}'
        total += e - s
    assert total == 192 * 192


def test_masked_interpolator_matches_polyfill(cuda):
    res = 256  # the sphere covers more than one 8192-fragment row
    pos, tri, _ = noisy_clip(5, 0.01, res)
    g = torch.Generator().manual_seed(40)
    attr = torch.rand(pos.shape[0], 5, generator=g)
    ctx = nvdr.RasterizeCudaContext()
    out = {}
    for impl in ('hip', 'polyfill'):
        p = pos.cuda().requires_grad_(True)
        a = attr.cuda().requires_grad_(True)
        rast, _ = nvdr.rasterize(ctx, p, tri.cuda(), (res, res), grad_db=False)
        mask = rast[..., 3] > 0
        mi = dd.MaskedSparseInterpolator(rast, tri.cuda(), mask)
        if impl == 'hip':
            y = mi.interpolate(a)
        else:
            y = dd.polyfill_interpolate(a, mi.vi_data, mi.tri_idx)
        go = torch.randn(y.shape, generator=torch.Generator().manual_seed(41)).cuda()
        (y * go).sum().backward()
        out[impl] = (y.detach(), p.grad.clone(), a.grad.clone())
    assert out['hip'][0].shape[0] > 8192
    for (x, y_), what, tol in zip(zip(out['hip'], out['polyfill']), ('value', 'grad pos', 'grad attr'),
                                  (1e-6, 1e-5, 1e-5)):
        scale = y_.abs().max()
        assert scale > 0, what
        assert (x - y_).abs().max() <= tol * scale, what
