"""
Edge cases through the whole operator chain (rasterize -> interpolate ->
antialias -> L2 -> backward): an empty mesh, every triangle behind the eye,
ragged and extreme resolutions (1 x 1, 1 x 37, 97 x 33, 16384 x 4), and
triangles far larger than the screen; ids, barycentrics and depth bit-exact
with the oracle, gradients within 1e-4 of the oracle's scale (zero where
nothing is visible).
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _chain(pos, tri, res, C=3, seed=0):
    g = torch.Generator().manual_seed(seed)
    V = pos.shape[-2]
    attr = torch.rand(V, C, generator=g)
    ctx = nvdr.RasterizeCudaContext()
    p = pos.cuda().requires_grad_(True)
    a = attr.cuda().requires_grad_(True)
    t = tri.cuda().int().contiguous()
    rast, _ = nvdr.rasterize(ctx, p, t, res, grad_db=False)
    col, _ = nvdr.interpolate(a, rast, t)
    y = nvdr.antialias(col, rast, p, t)
    target = torch.rand(y.shape, generator=g).cuda()
    ((y - target) ** 2).sum().backward()
    torch.cuda.synchronize()
    return rast.detach().cpu().numpy(), y.detach().cpu().numpy(), p.grad.cpu().numpy(), a.grad.cpu().numpy()


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def test_empty_mesh(cuda):
    pos = torch.rand(5, 4) + torch.tensor([0.0, 0.0, 0.0, 1.0])
    tri = torch.zeros(0, 3, dtype=torch.int32)
    rast, y, gp, ga = _chain(pos, tri, (16, 24))
    assert not rast.any() and not y.any()
    assert not gp.any() and not ga.any()


def test_everything_behind_the_eye(cuda):
    pos, tri, _ = noisy_clip(3, 0.01, 64)
    pos = pos.clone()
    pos[:, 3] = -pos[:, 3].abs() - 0.1  # w < 0 everywhere
    rast, y, gp, ga = _chain(pos, tri, (64, 64))
    assert not rast.any() and not y.any()
    assert not gp.any() and not ga.any()


@pytest.mark.parametrize('res', [(1, 1), (1, 37), (37, 1), (97, 33), (4, 16384)])
def test_ragged_resolutions_match_oracle(cuda, res):
    print(f'case: {res}', flush=True)
    pos, tri, _ = noisy_clip(4, 0.01, max(res))
    # squeeze the view to the resolution's aspect: the mesh covers a band of a thin image
    rast, y, gp, ga = _chain(pos, tri, res)
    ref = oracle.rasterize(pos.numpy(), tri.numpy(), res)
    assert np.array_equal(_bits(rast), _bits(ref))
    if res[0] * res[1] > 1:
        assert (ref[..., 3] > 0).any()
    assert np.isfinite(gp).all() and np.isfinite(ga).all()


def test_triangles_far_larger_than_the_screen(cuda):
    # a quad spanning +-50 in NDC and one crossing w = 0, over a small triangle
    pos = torch.tensor([[-50.0, -50.0, 0.5, 1.0], [50.0, -50.0, 0.5, 1.0], [50.0, 50.0, 0.5, 1.0],
                        [-50.0, 50.0, 0.5, 1.0], [-0.5, -0.5, 0.0, 1.0], [0.5, -0.4, 0.0, 1.0],
                        [0.0, 0.6, -0.9, -0.5]])
    tri = torch.tensor([[0, 1, 2], [0, 2, 3], [4, 5, 6]], dtype=torch.int32)
    for res in [(64, 48), (1, 1), (300, 17)]:
        rast, y, gp, ga = _chain(pos, tri, res, seed=1)
        ref = oracle.rasterize(pos.numpy(), tri.numpy(), res)
        assert np.array_equal(_bits(rast), _bits(ref)), res
        assert (ref[..., 3] > 0).all()  # the quad covers every pixel
        g = torch.Generator().manual_seed(3)
        gr = torch.randn(1, res[0], res[1], 4, generator=g)
        gr[..., 3] = 0
        p = pos.cuda().requires_grad_(True)
        r, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tri.cuda(), res, grad_db=False)
        r.backward(gr.cuda())
        o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), r.detach().cpu().numpy(), gr.numpy())
        scale = max(np.abs(o).max(), 1e-30)
        assert np.abs(p.grad.cpu().numpy() - o).max() <= 1e-4 * scale, res


def test_triangles_with_out_of_range_vertices_are_skipped(cuda):
    """Triangles naming a vertex outside [0, V) draw nothing and receive no
    gradient; the rest of the mesh renders and differentiates exactly as
    without them."""
    pos, tri, _ = noisy_clip(4, 0.01, 96, views=2)
    V = pos.shape[1]
    bad = torch.tensor([[0, 1, V], [V + 5, 2, 3], [-1, 4, 5], [6, 7, 1 << 30]], dtype=torch.int32)
    out = []
    for t in (tri.int(), torch.cat([tri.int(), bad])):
        g = torch.Generator().manual_seed(14)
        attr = torch.rand(V, 4, generator=g)
        go = torch.randn(2, 96, 96, 4, generator=g).cuda()
        p = pos.cuda().requires_grad_(True)
        a = attr.cuda().requires_grad_(True)
        tc = t.cuda().contiguous()
        rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tc, (96, 96), grad_db=False)
        col, _ = nvdr.interpolate(a, rast, tc)
        y = nvdr.antialias(col, rast, p, tc)
        (y * go).sum().backward()
        out.append((rast.detach(), p.grad.clone(), a.grad.clone()))
    (r0, gp0, ga0), (r1, gp1, ga1) = out
    assert torch.equal(r0.view(torch.int32), r1.view(torch.int32))
    assert torch.equal(ga0, ga1)
    assert (gp0 - gp1).abs().max() <= 1e-6 * gp0.abs().max()  # antialias: float atomics
