"""The clip-space transform and the image L2 loss kernels against float64
torch references (GPU)."""
import pytest
import torch

from diffrp_amd.camera import PerspectiveCamera, clip_space
from diffrp_amd.geometry import fibonacci_directions
from diffrp_amd.loss import l2_loss
from tests.scenes import vp_from_direction

pytestmark = pytest.mark.gpu


def _vps(n, res=64):
    return torch.stack([vp_from_direction(d, res) for d in fibonacci_directions(n)])


@pytest.mark.parametrize('n', [1, 7])
def test_clip_space_forward_and_backward(cuda, n):
    g = torch.Generator().manual_seed(80)
    xyz = torch.randn(1000, 3, generator=g)
    VP = _vps(n)
    x = xyz.cuda().requires_grad_(True)
    pos = clip_space(x, VP.cuda())
    want = torch.matmul(torch.nn.functional.pad(xyz.double(), (0, 1), value=1.0), VP.double().transpose(-1, -2))
    assert pos.shape == (n, 1000, 4)
    torch.testing.assert_close(pos.double().cpu(), want, rtol=1e-6, atol=1e-6)
    go = torch.randn(n, 1000, 4, generator=g)
    pos.backward(go.cuda())
    gwant = torch.einsum('bvj,bjk->vk', go.double(), VP.double())[:, :3]
    torch.testing.assert_close(x.grad.double().cpu(), gwant, rtol=1e-5, atol=1e-5)
    # one [4, 4] matrix gives [V, 4]; a differentiable matrix gets its gradient too
    vp1 = VP[0].cuda().requires_grad_(True)
    p1 = clip_space(x.detach(), vp1)
    assert p1.shape == (1000, 4)
    torch.testing.assert_close(p1.double().cpu(), want[0], rtol=1e-6, atol=1e-6)
    p1.backward(go[0].cuda())
    h = torch.nn.functional.pad(xyz.double(), (0, 1), value=1.0)
    torch.testing.assert_close(vp1.grad.double().cpu(), go[0].double().T @ h, rtol=1e-5, atol=1e-4)


@pytest.mark.parametrize('n', [1, 3, 4 * 1000 + 3, 10_000_000])
def test_l2_loss_value_and_gradient(cuda, n):
    g = torch.Generator().manual_seed(81)
    a = torch.rand(n, generator=g).cuda().requires_grad_(True)
    b = torch.rand(n, generator=g).cuda()
    loss = l2_loss(a, b)
    want = ((a.detach().double() - b.double()) ** 2).sum()
    assert abs(float(loss) - float(want)) <= 1e-6 * float(want) + 1e-30
    loss.backward(torch.tensor(0.75, device='cuda'))
    assert torch.equal(a.grad, (2.0 * 0.75) * (a.detach() - b))


def test_l2_loss_unaligned_views_and_target_gradient(cuda):
    g = torch.Generator().manual_seed(82)
    base_a = torch.rand(4097, generator=g).cuda()
    base_b = torch.rand(4097, generator=g).cuda()
    a = base_a[1:].requires_grad_(True)      # 4-byte offset: the scalar path
    b = base_b[1:].clone().requires_grad_(True)
    loss = l2_loss(a, b)
    loss.backward()
    d = a.detach() - b.detach()
    assert torch.equal(a.grad, 2.0 * d) and torch.equal(b.grad, -(2.0 * d))
    assert abs(float(loss) - float((d.double() ** 2).sum())) <= 1e-6 * float((d.double() ** 2).sum())


def test_l2_loss_second_backward_recomputes(cuda):
    # the forward's unit gradient is handed out once; a second backward through
    # the same graph reads both images again
    g = torch.Generator().manual_seed(84)
    a = torch.rand(4 * 1000 + 1, generator=g).cuda().requires_grad_(True)
    b = torch.rand(4 * 1000 + 1, generator=g).cuda().requires_grad_(True)
    loss = l2_loss(a, b)
    loss.backward(torch.tensor(0.5, device='cuda'), retain_graph=True)
    ga1, gb1 = a.grad.clone(), b.grad.clone()
    loss.backward(torch.tensor(2.0, device='cuda'))
    d = a.detach() - b.detach()
    assert torch.equal(ga1, (2.0 * 0.5) * d) and torch.equal(gb1, -((2.0 * 0.5) * d))
    assert torch.equal(a.grad, ga1 + (2.0 * 2.0) * d) and torch.equal(b.grad, gb1 - (2.0 * 2.0) * d)


def test_l2_loss_without_gradients_writes_none(cuda):
    a = torch.rand(1000, device='cuda')
    b = torch.rand(1000, device='cuda')
    loss = l2_loss(a, b)
    assert not loss.requires_grad
    assert abs(float(loss) - float(((a.double() - b.double()) ** 2).sum())) <= 1e-6 * float(loss)


def test_l2_loss_is_deterministic_and_capturable(cuda):
    g = torch.Generator().manual_seed(83)
    a = torch.rand(64, 96, 96, 8, generator=g).cuda().requires_grad_(True)
    b = torch.rand(64, 96, 96, 8, generator=g).cuda()

    def step():
        a.grad = None
        loss = l2_loss(a, b)
        loss.backward()
        return loss.detach()
    ref = step()
    ref_g = a.grad.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        out = step()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(out, ref) and torch.equal(a.grad, ref_g)
        torch.rand(1 << 20, device='cuda').sum().item()
