"""
The fine raster's depth cull: the kernels that skip, after a tile's first
round of entries, every direct-class triangle whose lowest possible depth in
its box lies beyond the tile's farthest current depth (chosen when bins hold
more than four rounds of a tile's lanes).  On crumpled meshes (vertex noise
several edge lengths), where tiles hold many rounds of entries and most
triangles lie behind the tile's current depths, ids, barycentrics, depth and
derivatives must be bit-exact with the kernels without the cull (a capacity of
exactly four rounds) and with the oracle: single layers for every tile shape,
peeled layers, and an uneven batch.  The bound behind the cull is checked
against brute force in test_depth_cull_math.py.
"""
import numpy as np
import pytest

import oracle
from diffrp_amd import batch, nvdr
from tests.scenes import noisy_clip

pytestmark = pytest.mark.gpu

TILES = [(8, 8), (16, 8), (16, 16), (32, 8)]


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def _assert_same(a, b, what):
    a, b = _bits(a), _bits(b)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f'{what}: {len(bad)} words differ, first at {bad[0].tolist()}'


@pytest.fixture(scope='module')
def crumpled():
    res = (128, 128)
    pos, tri, _ = noisy_clip(7, 0.02, res[0], views=2)  # 327,680 triangles, noise ~5 edge lengths
    ref, ref_db = oracle.rasterize(pos.numpy(), tri.numpy(), res, with_db=True)
    assert (ref[..., 3] > 0).mean() > 0.3
    return pos, tri, res, ref, ref_db


@pytest.mark.parametrize('tile', TILES)
def test_culled_raster_matches_unculled_and_oracle(cuda, crumpled, tile):
    print(f'case: crumpled s=7 x 2 views at 128^2, tile {tile}', flush=True)
    pos, tri, res, ref, ref_db = crumpled
    plain = nvdr.RasterizeCudaContext(tile=tile, bin_capacity=4 * tile[0] * tile[1])
    culled = nvdr.RasterizeCudaContext(tile=tile, bin_capacity=65536)
    r0, d0 = nvdr.rasterize(plain, pos.cuda(), tri.cuda(), res, grad_db=True)
    r1, d1 = nvdr.rasterize(culled, pos.cuda(), tri.cuda(), res, grad_db=True)
    _assert_same(r1.cpu().numpy(), r0.cpu().numpy(), 'rast, culled vs not')
    _assert_same(d1.cpu().numpy(), d0.cpu().numpy(), 'rast_db, culled vs not')
    _assert_same(r1.cpu().numpy(), ref, 'rast vs oracle')
    _assert_same(d1.cpu().numpy(), ref_db, 'rast_db vs oracle')


def test_culled_depth_peeling_matches_oracle(cuda, crumpled):
    print('case: crumpled s=7, 4 peeled layers, culling kernels', flush=True)
    pos, tri, res, _, _ = crumpled
    ctx = nvdr.RasterizeCudaContext(tile=(16, 16), bin_capacity=65536)
    with nvdr.DepthPeeler(ctx, pos.cuda(), tri.cuda(), res, grad_db=False) as dp:
        layers = [dp.rasterize_next_layer()[0].cpu().numpy() for _ in range(4)]
    prev = None
    for k in range(4):
        o = oracle.rasterize(pos.numpy(), tri.numpy(), res, prev=prev)
        _assert_same(layers[k], o, f'layer {k}')
        assert (o[..., 3] > 0).any()
        prev = o


def test_culled_uneven_batch_matches_oracle(cuda):
    print('case: uneven batch with crumpled items, culling kernels', flush=True)
    subdivs = [7, 3, 6, 5]
    res = [(128, 96), (64, 64), (160, 160), (48, 80)]
    pos, tri = [], []
    for i, (s, r) in enumerate(zip(subdivs, res)):
        p, t, _ = noisy_clip(s, 0.02, r[0], seed=10 + i)
        pos.append(p)
        tri.append(t.int())
    ctx = nvdr.RasterizeCudaContext(tile=(32, 8), bin_capacity=65536)
    rasts, dbs = batch.rasterize_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], res)
    for i in range(len(res)):
        o, odb = oracle.rasterize(pos[i].numpy(), tri[i].numpy(), res[i], with_db=True)
        _assert_same(rasts[i].cpu().numpy(), o, f'item {i} rast')
        _assert_same(dbs[i].cpu().numpy(), odb, f'item {i} rast_db')
