"""
Heavy screen tiles split across workgroups.  A dense mesh at a small
resolution puts thousands of triangles into every tile (81,920 triangles,
none culled, over 16-64 tiles: at least 1,280 entries per tile against a
split threshold of 8 rounds, 512-2,048 entries), so the fine raster runs each
tile's entry list in chunks on separate workgroups and merges their depth
keys with a device-wide 64-bit minimum.  The result must equal the
single-workgroup one: ids, barycentrics, depth and derivatives bit-exact with
the oracle for every tile shape and bin capacity (bins that overflow into the
spill list included), for peeled layers, inside uneven batches, and through
the backward, which relies on the per-triangle hit flags the last chunk sets.
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import batch, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu

TILES = [(8, 8), (16, 8), (16, 16), (32, 8)]
RES = (64, 64)


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def _assert_same(a, b, what):
    a, b = _bits(a), _bits(b)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f'{what}: {len(bad)} words differ, first at {bad[0].tolist()}'


@pytest.fixture(scope='module')
def dense():
    pos, tri, _ = noisy_clip(6, 0.02, RES[0], views=2)
    ref, ref_db = oracle.rasterize(pos.numpy(), tri.numpy(), RES, with_db=True)
    assert (ref[..., 3] > 0).mean() > 0.3
    return pos, tri, ref, ref_db


@pytest.mark.parametrize('tile', TILES)
@pytest.mark.parametrize('cap', [None, 65536])
def test_split_tiles_match_oracle(cuda, dense, tile, cap):
    print(f'case: tile {tile} bin capacity {cap}', flush=True)
    pos, tri, ref, ref_db = dense
    ctx = nvdr.RasterizeCudaContext(tile=tile, bin_capacity=cap)
    rast, db = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), RES, grad_db=True)
    _assert_same(rast.cpu().numpy(), ref, 'rast')
    _assert_same(db.cpu().numpy(), ref_db, 'rast_db')
    # the same call again: the split bookkeeping starts from zero every call
    rast2, _ = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), RES, grad_db=False)
    _assert_same(rast2.cpu().numpy(), ref, 'rast (second call)')


@pytest.mark.parametrize('tile', [(16, 16), (8, 8)])
def test_split_tiles_depth_peeling(cuda, dense, tile):
    print(f'case: peel, tile {tile}', flush=True)
    pos, tri, _, _ = dense
    ctx = nvdr.RasterizeCudaContext(tile=tile)
    with nvdr.DepthPeeler(ctx, pos.cuda(), tri.cuda(), RES, grad_db=False) as dp:
        layers = [dp.rasterize_next_layer()[0].cpu().numpy() for _ in range(4)]
    prev = None
    for k in range(4):
        o = oracle.rasterize(pos.numpy(), tri.numpy(), RES, prev=prev)
        _assert_same(layers[k], o, f'layer {k}')
        assert (o[..., 3] > 0).any()
        prev = o


def test_split_tiles_backward(cuda, dense):
    print('case: backward through split tiles', flush=True)
    pos, tri, ref, _ = dense
    g = torch.Generator().manual_seed(21)
    gr = torch.randn(2, RES[0], RES[1], 4, generator=g)
    gr[..., 3] = 0
    p = pos.cuda().requires_grad_(True)
    r, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(tile=(16, 16)), p, tri.cuda(), RES, grad_db=False)
    _assert_same(r.detach().cpu().numpy(), ref, 'rast')
    r.backward(gr.cuda())
    o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), r.detach().cpu().numpy(), gr.numpy())
    scale = max(np.abs(o).max(), 1e-30)
    err = np.abs(p.grad.cpu().numpy() - o).max()
    assert err <= 1e-4 * scale, f'max |diff| {err:.3e}, scale {scale:.3e}'
    assert scale > 0


def test_split_tiles_in_uneven_batch(cuda):
    print('case: uneven batch, dense items between light ones', flush=True)
    subdivs = [2, 6, 3, 5]
    res = [(96, 80), (64, 64), (128, 128), (40, 56)]
    pos, tri = [], []
    for i, (s, r) in enumerate(zip(subdivs, res)):
        p, t, _ = noisy_clip(s, 0.02, r[0], seed=i)
        pos.append(p)
        tri.append(t.int())
    for tile in [(16, 16), (32, 8)]:
        ctx = nvdr.RasterizeCudaContext(tile=tile)
        rasts, dbs = batch.rasterize_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], res)
        for i in range(len(res)):
            o, odb = oracle.rasterize(pos[i].numpy(), tri[i].numpy(), res[i], with_db=True)
            _assert_same(rasts[i].cpu().numpy(), o, f'tile {tile} item {i} rast')
            _assert_same(dbs[i].cpu().numpy(), odb, f'tile {tile} item {i} rast_db')


def test_more_heavy_tiles_than_split_slots(cuda):
    """Two crumpled s=8 views at 512² with 8x8 tiles: ~6,500 of the 8,192
    tiles exceed the split threshold while at most 4,096 tiles are split per
    call; the rest run whole in one block.  Both kinds must match the oracle."""
    print('case: s=8 (1.3M triangles) x 2 views, 512^2, 8x8 tiles, more heavy tiles than split slots', flush=True)
    res = (512, 512)
    pos, tri, _ = noisy_clip(8, 0.02, res[0], views=2)
    ref = oracle.rasterize(pos.numpy(), tri.numpy(), res)
    ctx = nvdr.RasterizeCudaContext(tile=(8, 8), bin_capacity=8192)
    rast, _ = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), res, grad_db=False)
    _assert_same(rast.cpu().numpy(), ref, 'rast')
    assert (ref[..., 3] > 0).mean() > 0.3
