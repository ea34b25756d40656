"""Antialias position gradients handed to the fused interpolation of the same
positions upstream (which adds them to its own in place of autograd's sum)
give the gradients of the ordinary path: with antialias straight after the
interpolation and after shading, with two antialias calls, for a partial
gradient request, after a pass that stopped short of the interpolation (its
left gradient is not reused), for a second backward, and replayed from a
captured graph; antialias of a colour that no interpolation made returns its
own position gradient."""
import pytest
import torch

from diffrp_amd import nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu


def _scene(cuda, views, res, seed):
    pos, tri, _ = noisy_clip(4, 0.01, res[0], seed=seed, views=views)
    pos, tri = pos.to(cuda).contiguous(), tri.to(cuda).int().contiguous()
    if views == 1:
        pos = pos[0]
    g = torch.Generator(device='cpu').manual_seed(seed + 1)
    attr = torch.rand((pos.shape[-2], 8), generator=g).to(cuda)
    target = torch.rand((views, res[0], res[1], 8), generator=g).to(cuda)
    return pos, tri, attr, target


def _graph(pos, tri, attr, target, res, shade=False, twice=False):
    p = pos.clone().requires_grad_(True)
    a = attr.clone().requires_grad_(True)
    rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tri, res, grad_db=False)
    col, _ = nvdr.interpolate(a, rast, tri)
    if shade:  # the antialiased colour descends from the interpolation through shading
        col = col * 0.5 + 0.25 * col.flip(-1)
    y = nvdr.antialias(col, rast, p, tri)
    loss = ((y - target) ** 2).sum()
    if twice:
        loss = loss + 0.1 * (nvdr.antialias(col * 2.0, rast, p, tri) ** 2).sum()
    return p, a, col, loss


def _with(hand, fn):
    saved = nvdr._HAND_AA_POS
    nvdr._HAND_AA_POS = hand
    try:
        out = fn()
        torch.cuda.synchronize()
        return out
    finally:
        nvdr._HAND_AA_POS = saved


def _close(x, y):
    # the antialias position gradient is summed with float atomics
    return float((x - y).abs().max()) <= 1e-5 * float(y.abs().max())


@pytest.mark.parametrize('views,res,shade,twice', [(1, (96, 128), False, False), (6, (128, 160), False, False),
                                                   (3, (96, 96), True, False), (2, (80, 112), False, True)])
def test_handoff_equals_autograd_sum(cuda, views, res, shade, twice):
    print(f'case: views={views} res={res} shade={shade} twice={twice}', flush=True)
    scene = _scene(cuda, views, res, 20 + views)

    def run():
        p, a, _, loss = _graph(*scene, res, shade, twice)
        loss.backward()
        return p.grad, a.grad
    gp1, ga1 = _with(True, run)
    gp0, ga0 = _with(False, run)
    assert torch.equal(ga1, ga0)
    assert _close(gp1, gp0) and float(gp0.abs().max()) > 0


def test_handoff_partial_requests_and_second_backward(cuda):
    res = (96, 96)
    scene = _scene(cuda, 2, res, 30)

    def only_attr():
        _, a, _, loss = _graph(*scene, res)
        return torch.autograd.grad(loss, [a])[0]
    assert torch.equal(_with(True, only_attr), _with(False, only_attr))

    def stopped_short_then_full():
        # the first pass ends at col: the antialias leaves its position
        # gradient, the interpolation never takes it; the second pass must not add it
        p, a, col, loss = _graph(*scene, res)
        torch.autograd.grad(loss, [col], retain_graph=True)
        loss.backward()
        return p.grad, a.grad
    gp1, ga1 = _with(True, stopped_short_then_full)
    gp0, ga0 = _with(False, stopped_short_then_full)
    assert torch.equal(ga1, ga0) and _close(gp1, gp0)

    def twice_backward():
        p, a, _, loss = _graph(*scene, res)
        loss.backward(retain_graph=True)
        loss.backward()
        return p.grad, a.grad
    gp1, ga1 = _with(True, twice_backward)
    gp0, ga0 = _with(False, twice_backward)
    assert torch.equal(ga1, ga0) and _close(gp1, gp0)


def test_colour_without_interpolation_keeps_its_gradient(cuda):
    res = (64, 64)
    pos, tri, _, _ = _scene(cuda, 1, res, 40)
    p = pos.clone().requires_grad_(True)
    color = torch.rand((1, 64, 64, 3), device=cuda, requires_grad=True)
    rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tri, res, grad_db=False)
    (nvdr.antialias(color, rast, p, tri) ** 2).sum().backward()
    assert p.grad is not None and float(p.grad.abs().max()) > 0 and color.grad is not None


def test_handoff_replays_from_a_graph(cuda):
    res = (128, 128)
    pos, tri, attr, target = _scene(cuda, 4, res, 50)

    def ref():
        p, a, _, loss = _graph(pos, tri, attr, target, res)
        loss.backward()
        return p.grad, a.grad
    gp_ref, ga_ref = _with(False, ref)
    ctx = nvdr.RasterizeCudaContext()
    p = pos.clone().requires_grad_(True)
    a = attr.clone().requires_grad_(True)

    def step():
        p.grad = None
        a.grad = None
        rast, _ = nvdr.rasterize(ctx, p, tri, res, grad_db=False)
        col, _ = nvdr.interpolate(a, rast, tri)
        y = nvdr.antialias(col, rast, p, tri)
        ((y - target) ** 2).sum().backward()

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
        assert torch.equal(a.grad, ga_ref)
        assert _close(p.grad, gp_ref)
        torch.rand(1 << 20, device=cuda).sum().item()  # eager work between replays
