"""GPU parity of the rasteriser (HIP, through the C-ABI) against the C oracle.

Triangle ids and z/w must be bit-exact; (u, v) are produced by the same
float operation sequence and are compared bit-exact too.
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import nvdr
from tests.scenes import noisy_clip, random_soup, readme_clip

pytestmark = pytest.mark.gpu


def _gpu_rast(pos, tri, res, ranges=None, grad_db=False, ctx=None, layers=1):
    ctx = ctx or nvdr.RasterizeCudaContext()
    pos_d, tri_d = pos.cuda(), tri.cuda()
    out = []
    with nvdr.DepthPeeler(ctx, pos_d, tri_d, res, ranges=ranges, grad_db=grad_db) as dp:
        for _ in range(layers):
            r, db = dp.rasterize_next_layer()
            out.append((r.cpu().numpy(), db.cpu().numpy() if grad_db else None))
    torch.cuda.synchronize()
    return out


def _assert_rast_equal(g, o, what):
    gi, oi = oracle.float_to_triidx(g[..., 3]), oracle.float_to_triidx(o[..., 3])
    n_bad = int((gi != oi).sum())
    assert n_bad == 0, f'{what}: {n_bad} pixels with different triangle ids'
    np.testing.assert_array_equal(g[..., 2].view(np.uint32), o[..., 2].view(np.uint32), err_msg=f'{what}: z/w')
    np.testing.assert_array_equal(g[..., :2].view(np.uint32), o[..., :2].view(np.uint32), err_msg=f'{what}: u,v')


def test_readme_scene_256(cuda):
    pos, tri, _, _ = readme_clip(256)
    print('case: README icosphere s=3, 256x256', flush=True)
    (g, _), = _gpu_rast(pos, tri, (256, 256))
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (256, 256))
    _assert_rast_equal(g, o, 'readme 256')
    assert (g[..., 3] > 0).sum() > 40000


def test_readme_scene_512_matches_reference_mask(cuda):
    """Coverage of the README render equals the reference's documented 512^2 output."""
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'readme_normals_512.npz'))['rgba_bottom_first']
    pos, tri, _, _ = readme_clip(512)
    (g, _), = _gpu_rast(pos, tri, (512, 512))
    np.testing.assert_array_equal(g[0, ..., 3] > 0, fx[..., 3] > 0)


@pytest.mark.parametrize('tile', [(8, 8), (16, 8), (16, 16), (32, 8)])
def test_noisy_sphere_1024_all_tiles(cuda, tile):
    pos, tri, _ = noisy_clip(6, 0.01, 1024)
    print(f'case: noisy s=6 1024^2 tile={tile}', flush=True)
    ctx = nvdr.RasterizeCudaContext(tile=tile)
    (g, _), = _gpu_rast(pos, tri, (1024, 1024), ctx=ctx)
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (1024, 1024))
    _assert_rast_equal(g, o, f'noisy 1024 {tile}')


@pytest.mark.parametrize('cap', [1, 4, 32])
def test_bin_overflow_spills_without_changing_results(cuda, cap):
    pos, tri, _ = noisy_clip(5, 0.01, 256)
    print(f'case: noisy s=5 256^2 bin_cap={cap}', flush=True)
    ctx = nvdr.RasterizeCudaContext(tile=(16, 16), bin_capacity=cap)
    (g, _), = _gpu_rast(pos, tri, (256, 256), ctx=ctx)
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (256, 256))
    _assert_rast_equal(g, o, f'cap {cap}')


def test_instanced_fibonacci_views(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 192, views=6)
    print('case: instanced 6 views s=4 192^2', flush=True)
    (g, _), = _gpu_rast(pos, tri, (192, 160))
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (192, 160))
    _assert_rast_equal(g, o, 'instanced')


def test_range_mode(cuda):
    pos, tri = random_soup(300, seed=3)
    ranges = torch.tensor([[0, 100], [50, 200], [290, 40], [0, 0]], dtype=torch.int32)
    print('case: range mode 4 ranges, 300 random tris, 96x80', flush=True)
    (g, _), = _gpu_rast(pos, tri, (96, 80), ranges=ranges)
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (96, 80), ranges=ranges.numpy())
    assert g.shape == (4, 96, 80, 4)
    _assert_rast_equal(g, o, 'ranges')


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_random_overlapping_soup(cuda, seed):
    pos, tri = random_soup(2000, seed=seed)
    print(f'case: random soup seed={seed}', flush=True)
    (g, _), = _gpu_rast(pos, tri, (77, 131))
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (77, 131))
    _assert_rast_equal(g, o, f'soup {seed}')


def test_clipping_and_guard_band(cuda):
    """Triangles crossing w = 0, behind the camera, far outside the guard band, degenerate."""
    g_ = torch.Generator().manual_seed(7)
    n = 400
    pos = torch.randn(n * 3, 4, generator=g_) * torch.tensor([3.0, 3.0, 1.0, 1.0])
    pos[:, 3] = pos[:, 3] * 1.5 + 0.3          # many w <= 0
    pos[:, 2] = pos[:, 2].clamp(-0.9, 0.9) * pos[:, 3].abs()
    pos[::7, 0] *= 1e5                          # beyond the guard band
    pos[5:9] = pos[5:6]                         # degenerate triangles
    tri = torch.arange(n * 3, dtype=torch.int32).reshape(n, 3)
    print('case: clipping / guard band', flush=True)
    (g, _), = _gpu_rast(pos, tri, (64, 64))
    o = oracle.rasterize(pos.numpy(), tri.numpy(), (64, 64))
    _assert_rast_equal(g, o, 'clipping')
    assert (g[..., 3] > 0).sum() > 100


def test_depth_peeling_layers(cuda):
    pos, tri, _ = noisy_clip(4, 0.02, 128)
    print('case: 4 peel layers s=4 128^2', flush=True)
    layers = _gpu_rast(pos, tri, (128, 128), layers=4)
    prev = None
    for k, (g, _) in enumerate(layers):
        o = oracle.rasterize(pos.numpy(), tri.numpy(), (128, 128), prev=prev)
        _assert_rast_equal(g, o, f'layer {k}')
        prev = o
    assert (layers[1][0][..., 3] > 0).sum() > 0


def test_rast_db(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 160)
    print('case: rast_db s=4 160^2', flush=True)
    (g, gdb), = _gpu_rast(pos, tri, (160, 160), grad_db=True)
    o, odb = oracle.rasterize(pos.numpy(), tri.numpy(), (160, 160), with_db=True)
    _assert_rast_equal(g, o, 'db rast')
    np.testing.assert_array_equal(gdb.view(np.uint32), odb.view(np.uint32))


def test_deterministic_repeat(cuda):
    pos, tri, _ = noisy_clip(5, 0.01, 256, views=3)
    a = _gpu_rast(pos, tri, (256, 256))[0][0]
    b = _gpu_rast(pos, tri, (256, 256))[0][0]
    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32))


def test_rejects_cpu_and_bad_dtypes(cuda):
    pos, tri, _, _ = readme_clip(64)
    ctx = nvdr.RasterizeCudaContext()
    with pytest.raises(RuntimeError, match='ROCm device'):
        nvdr.rasterize(ctx, pos, tri, (64, 64))
    with pytest.raises(TypeError):
        nvdr.rasterize(ctx, pos.cuda(), tri.long().cuda(), (64, 64))
    with pytest.raises(RuntimeError, match='resolution'):
        nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), (0, 64))
