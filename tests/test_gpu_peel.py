"""
Depth peeling with on-device emptiness detection (``layered.peel_layers``):
the layers, and the gradients of a front-to-back transparent composite built
on them, equal those of diffrp's synchronous loop (``surface_deferred.py:
258-266``: stop at the first empty layer before the last) bit for bit; the
device covered-pixel counts equal torch's.
"""
import pytest
import torch

from diffrp_amd import layered, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _sync_layers(ctx, pos, tri, res, max_layers):
    out = []
    with nvdr.DepthPeeler(ctx, pos, tri, res, grad_db=False) as dp:
        for i in range(max_layers):
            rast, _ = dp.rasterize_next_layer()
            if i < max_layers - 1 and not bool((rast[..., 3] > 0).any()):
                break
            out.append(rast)
    return out


def test_count_covered_matches_torch(cuda):
    pos, tri, _ = noisy_clip(4, 0.02, 96, views=3)
    rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), pos.cuda(), tri.cuda(), (96, 80), grad_db=False)
    got = layered.count_covered(rast)
    assert got.dtype == torch.int32 and got.device.type == 'cuda'
    assert torch.equal(got.long(), (rast[..., 3] > 0).sum((1, 2)))


@pytest.mark.parametrize('max_layers', [1, 2, 3, 5, 64])
def test_peel_layers_matches_synchronous_loop(cuda, max_layers):
    print(f'case: max_layers={max_layers}', flush=True)
    pos, tri, _ = noisy_clip(4, 0.05, 128, views=2)  # folded mesh: several layers, then empty
    ctx = nvdr.RasterizeCudaContext()
    p, t = pos.cuda(), tri.cuda()
    with torch.no_grad():
        got = layered.peel_layers(ctx, p, t, (128, 128), max_layers)
        ref = _sync_layers(ctx, p, t, (128, 128), max_layers)
    assert len(got) == len(ref)
    for a, b in zip(got, ref):
        assert torch.equal(a.view(torch.int32), b.view(torch.int32))
    print(f'{len(ref)} layers', flush=True)
    if max_layers == 64:
        assert 2 < len(ref) < 64  # stopped at an empty layer


def _composite(layers, attr, tri, pos):
    out, trans = None, None
    for rast in layers:
        col, _ = nvdr.interpolate(attr, rast, tri)
        a = 0.6 * (rast[..., 3:4] > 0).float()
        rgb = col[..., :3]
        if out is None:
            out, trans = rgb * a, 1.0 - a
        else:
            out, trans = out + trans * a * rgb, trans * (1.0 - a)
    return nvdr.antialias(out.contiguous(), layers[0], pos, tri)


def test_transparent_composite_gradients_match_synchronous_loop(cuda):
    pos, tri, _ = noisy_clip(4, 0.05, 96, views=2)
    g = torch.Generator().manual_seed(21)
    attr = torch.rand(pos.shape[1], 3, generator=g)
    target = torch.rand(2, 96, 96, 3, generator=g).cuda()
    ctx = nvdr.RasterizeCudaContext()
    res = []
    for peel in (layered.peel_layers, _sync_layers):
        p = pos.cuda().requires_grad_(True)
        a = attr.cuda().requires_grad_(True)
        layers = peel(ctx, p, tri.cuda(), (96, 96), 8)
        y = _composite(layers, a, tri.cuda(), p)
        loss = ((y - target) ** 2).sum()
        loss.backward()
        res.append((len(layers), float(loss), p.grad.clone(), a.grad.clone()))
    (n0, l0, gp0, ga0), (n1, l1, gp1, ga1) = res
    assert n0 == n1 >= 3 and l0 == l1
    assert gp0.abs().max() > 0 and ga0.abs().max() > 0
    # antialias position gradients use float atomics: equal up to summation order
    assert torch.equal(ga0, ga1)
    assert (gp0 - gp1).abs().max() <= 1e-5 * gp1.abs().max()
