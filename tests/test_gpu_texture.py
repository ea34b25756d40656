"""
Texture sampling (diffrp's sample2d / sample3d, shader_ops.py:198-310) on the
HIP kernels against the restatement in oracle.glue, i.e. torch's CPU
grid_sample in float64 (on the same float32 inputs) under diffrp's coordinate
mapping: values within 1e-5 (the float32 source-index arithmetic alone moves a
sample by ~1e-7 of the texture size) and gradients (texture: float atomics;
coordinates) within 1e-4 of their scale.  Pinned independently: texel centres return texel values, and (0, 0)
is the bottom-left corner.
"""
import numpy as np
import pytest
import torch

from oracle import glue
from diffrp_amd import texture

pytestmark = pytest.mark.gpu


def _close(a, b, tol, what):
    a = a.detach().double().cpu()
    b = b.detach().double().cpu()
    scale = max(float(b.abs().max()), 1e-12)
    err = float((a - b).abs().max()) / scale
    assert err <= tol, f'{what}: {err:.3e} > {tol}'


def _coords(n, dims, lo, hi, seed, texel_grid=None):
    g = torch.Generator().manual_seed(seed)
    if texel_grid is None:
        return torch.rand(n, dims, generator=g, dtype=torch.float64) * (hi - lo) + lo
    # nearest: texel centres + offsets within +-0.3 texel (no rounding ties)
    sizes = torch.tensor(texel_grid, dtype=torch.float64)
    idx = torch.floor(torch.rand(n, dims, generator=g, dtype=torch.float64) * sizes * (hi - lo) + lo * sizes)
    off = (torch.rand(n, dims, generator=g, dtype=torch.float64) - 0.5) * 0.6
    return (idx + 0.5 + off) / sizes


@pytest.mark.parametrize('C', [1, 3, 4, 6])
@pytest.mark.parametrize('wrap', ['border', 'reflection', 'cyclic', 'latlong'])
def test_sample2d_bilinear_matches_grid_sample(cuda, wrap, C):
    g = torch.Generator().manual_seed(C)
    tex = torch.rand(37, 53, C, generator=g).double()
    uv = _coords(4000, 2, -2.0 if wrap == 'cyclic' else -0.4, 3.0 if wrap == 'cyclic' else 1.4, 7).float().double()
    go = torch.randn(4000, C, generator=g, dtype=torch.float64)
    t_ref, uv_ref = tex.clone().requires_grad_(True), uv.clone().requires_grad_(True)
    ref = glue.sample2d(t_ref, uv_ref, wrap)
    (ref * go).sum().backward()
    t, u = tex.float().cuda().requires_grad_(True), uv.float().cuda().requires_grad_(True)
    out = texture.sample2d(t, u, wrap)
    (out * go.float().cuda()).sum().backward()
    _close(out, ref, 1e-5, 'value')
    _close(t.grad, t_ref.grad, 1e-4, 'grad texture')
    _close(u.grad, uv_ref.grad, 1e-4, 'grad uv')


@pytest.mark.parametrize('wrap', ['border', 'reflection', 'cyclic', 'latlong'])
def test_sample2d_nearest_matches_grid_sample(cuda, wrap):
    g = torch.Generator().manual_seed(3)
    tex = torch.rand(29, 41, 3, generator=g).double()
    uv = _coords(3000, 2, 0.0, 1.0, 8, texel_grid=[41, 29]).float().double()
    t_ref = tex.clone().requires_grad_(True)
    ref = glue.sample2d(t_ref, uv, wrap, 'nearest')
    go = torch.randn(3000, 3, generator=g, dtype=torch.float64)
    (ref * go).sum().backward()
    t = tex.float().cuda().requires_grad_(True)
    out = texture.sample2d(t, uv.float().cuda(), wrap, 'nearest')
    (out * go.float().cuda()).sum().backward()
    _close(out, ref, 1e-7, 'value')
    _close(t.grad, t_ref.grad, 1e-5, 'grad texture')


@pytest.mark.parametrize('wrap', ['border', 'reflection', 'cyclic', 'latlong'])
def test_sample2d_bicubic_matches_grid_sample(cuda, wrap):
    g = torch.Generator().manual_seed(4)
    tex = torch.rand(23, 31, 3, generator=g).double()
    uv = _coords(3000, 2, -2.0 if wrap == 'cyclic' else -0.3, 3.0 if wrap == 'cyclic' else 1.3, 10).float().double()
    go = torch.randn(3000, 3, generator=g, dtype=torch.float64)
    t_ref, uv_ref = tex.clone().requires_grad_(True), uv.clone().requires_grad_(True)
    ref = glue.sample2d(t_ref, uv_ref, wrap, 'bicubic')
    (ref * go).sum().backward()
    t, u = tex.float().cuda().requires_grad_(True), uv.float().cuda().requires_grad_(True)
    out = texture.sample2d(t, u, wrap, 'bicubic')
    (out * go.float().cuda()).sum().backward()
    _close(out, ref, 1e-5, 'value')
    _close(t.grad, t_ref.grad, 1e-4, 'grad texture')
    _close(u.grad, uv_ref.grad, 1e-4, 'grad uv')


@pytest.mark.parametrize('wrap', ['border', 'reflection', 'latlong'])
def test_sample3d_matches_grid_sample(cuda, wrap):
    g = torch.Generator().manual_seed(5)
    tex = torch.rand(9, 13, 17, 4, generator=g).double()
    uvw = _coords(3000, 3, -0.3, 1.3, 9).float().double()
    go = torch.randn(3000, 4, generator=g, dtype=torch.float64)
    t_ref, c_ref = tex.clone().requires_grad_(True), uvw.clone().requires_grad_(True)
    ref = glue.sample3d(t_ref, c_ref, wrap)
    (ref * go).sum().backward()
    t, c = tex.float().cuda().requires_grad_(True), uvw.float().cuda().requires_grad_(True)
    out = texture.sample3d(t, c, wrap)
    (out * go.float().cuda()).sum().backward()
    _close(out, ref, 1e-5, 'value')
    _close(t.grad, t_ref.grad, 1e-4, 'grad texture')
    _close(c.grad, c_ref.grad, 1e-4, 'grad uvw')


def test_texel_centres_and_orientation(cuda):
    H, W = 6, 10
    tex = torch.arange(H * W * 2, dtype=torch.float32).reshape(H, W, 2)
    ys, xs = torch.meshgrid(torch.arange(H), torch.arange(W), indexing='ij')
    # texel (y, x) sits at u = (x + 0.5) / W, v = 1 - (y + 0.5) / H: row 0 is the top
    uv = torch.stack([(xs + 0.5) / W, 1 - (ys + 0.5) / H], -1).float()
    for wrap in ('border', 'reflection', 'cyclic'):
        for mode in ('bilinear', 'nearest', 'bicubic'):
            out = texture.sample2d(tex.cuda(), uv.cuda(), wrap, mode).cpu()
            np.testing.assert_allclose(out.numpy(), tex.numpy(), rtol=0, atol=1e-4, err_msg=f'{wrap} {mode}')
    corner = texture.sample2d(tex.cuda(), torch.zeros(1, 2).cuda(), 'border').cpu()
    np.testing.assert_allclose(corner[0].numpy(), tex[H - 1, 0].numpy(), atol=1e-5)


def test_gltf_sampler_repeat(cuda):
    g = torch.Generator().manual_seed(6)
    img = torch.rand(16, 16, 4, generator=g)
    uv = torch.rand(500, 2, generator=g) * 4 - 2
    s = texture.GLTFSampler(img.cuda(), 'repeat', 'linear')
    ref = glue.sample2d(img.double(), uv.double().remainder(1.0), 'reflection')
    _close(s.sample(uv.cuda()), ref, 1e-5, 'repeat sampler')
