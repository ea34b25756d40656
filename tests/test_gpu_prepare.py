"""Antialias prepared by the rasterize call (contexts with
prepare_antialias=True: the per-vertex and per-triangle state on an auxiliary
stream beside the rasterizer) gives the same image and gradients as the
unprepared call, eagerly and replayed from a captured graph; a second
antialias of the same rast, or one with other positions, runs unprepared."""
import pytest
import torch

from diffrp_amd import nvdr
from tests.scenes import noisy_clip
from tests.test_gpu_fused import _Count

pytestmark = pytest.mark.gpu


def _step(ctx, pos, tri, attr, target, res):
    p = pos.clone().requires_grad_(True)
    a = attr.clone().requires_grad_(True)
    rast, _ = nvdr.rasterize(ctx, p, tri, res, grad_db=False)
    col, _ = nvdr.interpolate(a, rast, tri)
    y = nvdr.antialias(col, rast, p, tri)
    ((y - target) ** 2).sum().backward()
    return y.detach(), p.grad, a.grad


@pytest.mark.parametrize('views,res', [(1, (96, 128)), (6, (160, 192))])
def test_prepared_antialias_equals_unprepared(cuda, views, res):
    print(f'case: views={views} res={res}', flush=True)
    pos, tri, _ = noisy_clip(4, 0.01, res[0], seed=5, views=views)
    pos, tri = pos.to(cuda).contiguous(), tri.to(cuda).int().contiguous()
    g = torch.Generator(device='cpu').manual_seed(6)
    attr = torch.rand((pos.shape[-2], 4), generator=g).to(cuda)
    target = torch.rand((views, res[0], res[1], 4), generator=g).to(cuda)
    if views == 1:
        pos = pos[0]
    y0, gp0, ga0 = _step(nvdr.RasterizeCudaContext(), pos, tri, attr, target, res)
    with _Count('drp_antialias_fwd_prepared') as c:
        y1, gp1, ga1 = _step(nvdr.RasterizeCudaContext(prepare_antialias=True), pos, tri, attr, target, res)
    torch.cuda.synchronize()
    assert c.n == 1
    assert torch.equal(y0, y1)
    assert torch.equal(ga0, ga1)
    # the antialias position gradient is summed with float atomics
    assert (gp0 - gp1).abs().max() <= 1e-5 * gp0.abs().max()


def test_prepared_antialias_is_used_once_and_only_for_its_positions(cuda):
    pos, tri, _ = noisy_clip(3, 0.01, 64, seed=7, views=2)
    pos, tri = pos.to(cuda).contiguous(), tri.to(cuda).int().contiguous()
    color = torch.rand((2, 64, 64, 3), device=cuda)
    ctx = nvdr.RasterizeCudaContext(prepare_antialias=True)
    rast, _ = nvdr.rasterize(ctx, pos, tri, (64, 64), grad_db=False)
    ref = nvdr.antialias(color, rast, pos.clone(), tri)  # other positions: unprepared
    with _Count('drp_antialias_fwd_prepared') as c:
        y1 = nvdr.antialias(color, rast, pos, tri)
        y2 = nvdr.antialias(color, rast, pos, tri)
    assert c.n == 1
    assert torch.equal(ref, y1) and torch.equal(y1, y2)


def test_prepared_antialias_replays_from_a_graph(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 128, seed=8, views=4)
    pos, tri = pos.to(cuda).contiguous(), tri.to(cuda).int().contiguous()
    g = torch.Generator(device='cpu').manual_seed(9)
    attr = torch.rand((pos.shape[-2], 8), generator=g).to(cuda)
    target = torch.rand((4, 128, 128, 8), generator=g).to(cuda)
    y_ref, gp_ref, ga_ref = _step(nvdr.RasterizeCudaContext(), pos, tri, attr, target, (128, 128))
    ctx = nvdr.RasterizeCudaContext(prepare_antialias=True)
    p = pos.clone().requires_grad_(True)
    a = attr.clone().requires_grad_(True)
    out = {}

    def step():
        p.grad = None
        a.grad = None
        rast, _ = nvdr.rasterize(ctx, p, tri, (128, 128), grad_db=False)
        col, _ = nvdr.interpolate(a, rast, tri)
        y = nvdr.antialias(col, rast, p, tri)
        ((y - target) ** 2).sum().backward()
        out['y'] = y.detach()

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out['y'], y_ref)
        assert torch.equal(a.grad, ga_ref)
        assert (p.grad - gp_ref).abs().max() <= 1e-5 * gp_ref.abs().max()
