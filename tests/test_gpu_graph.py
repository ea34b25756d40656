"""
HIP-graph capture: the whole rasterize -> interpolate -> antialias -> L2 ->
backward step, the covered-pixel count and texture sampling are captured once
and replayed three times with eager work between the replays; every replay
reproduces the eager results (attribute gradients and counts bit for bit;
position and texture gradients, which use float atomics, to 1e-5 of scale).
"""
import pytest
import torch

from diffrp_amd import layered, nvdr, texture
from diffrp_amd.camera import clip_space
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _capture(fn):
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            fn()
    torch.cuda.current_stream().wait_stream(side)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = fn()
    return g, out


def _eager_noise():
    # unrelated eager work between replays (allocations, kernels, a host sync)
    x = torch.rand(1 << 20, device='cuda')
    return float((x * 2).sum())


def test_full_step_replays_match_eager(cuda):
    pos, tri, verts = noisy_clip(4, 0.01, 128, views=4)
    g = torch.Generator().manual_seed(50)
    attr = torch.rand(pos.shape[1], 4, generator=g).cuda().requires_grad_(True)
    target = torch.rand(4, 128, 128, 4, generator=g).cuda()
    p_static = pos.cuda().contiguous()
    t = tri.cuda().int().contiguous()
    ctx = nvdr.RasterizeCudaContext()
    topo = nvdr.antialias_construct_topology_hash(t)
    p_leaf = p_static.clone().requires_grad_(True)

    def step():
        p_leaf.grad = None
        attr.grad = None
        rast, _ = nvdr.rasterize(ctx, p_leaf, t, (128, 128), grad_db=False)
        col, _ = nvdr.interpolate(attr, rast, t)
        y = nvdr.antialias(col, rast, p_leaf, t, topology_hash=topo)
        loss = ((y - target) ** 2).sum()
        loss.backward()
        return loss.detach()

    l_ref = step()
    torch.cuda.synchronize()
    gp_ref, ga_ref = p_leaf.grad.clone(), attr.grad.clone()
    assert ga_ref.abs().max() > 0 and gp_ref.abs().max() > 0
    graph, loss_out = _capture(step)
    for i in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(loss_out, l_ref), i
        assert torch.equal(attr.grad, ga_ref), i
        assert (p_leaf.grad - gp_ref).abs().max() <= 1e-5 * gp_ref.abs().max(), i
        _eager_noise()


def test_count_and_texture_replays_match_eager(cuda):
    pos, tri, _ = noisy_clip(4, 0.02, 96, views=3)
    rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), pos.cuda(), tri.cuda(), (96, 96), grad_db=False)
    g = torch.Generator().manual_seed(51)
    tex = torch.rand(33, 47, 3, generator=g).cuda().requires_grad_(True)
    uv = (torch.rand(5000, 2, generator=g) * 1.4 - 0.2).cuda().requires_grad_(True)
    go = torch.randn(5000, 3, generator=g).cuda()

    def work():
        tex.grad = None
        uv.grad = None
        cnt = layered.count_covered(rast)
        y = texture.sample2d(tex, uv, 'reflection')
        (y * go).sum().backward()
        return cnt, y.detach()

    c_ref, y_ref = work()
    torch.cuda.synchronize()
    gt_ref, gu_ref = tex.grad.clone(), uv.grad.clone()
    graph, (c_out, y_out) = _capture(work)
    for i in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(c_out, c_ref) and torch.equal(y_out, y_ref), i
        assert (tex.grad - gt_ref).abs().max() <= 1e-5 * gt_ref.abs().max(), i
        assert torch.equal(uv.grad, gu_ref), i
        _eager_noise()
