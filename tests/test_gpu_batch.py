"""
Uneven batches (many meshes, each at its own resolution, one rasterizer launch).

Parity: every item's rast / rast_db is bit-exact with the oracle rasterizing
that item alone (small batch), and with ``nvdr.rasterize`` of the item alone
(full-size draw: 32 meshes of subdivision 3-8 at 256^2-2048^2, the sizes where
the oracle takes minutes); peeled layers match per-item peeling; gradients of
an interpolate + antialias loss through the batch equal those through per-item
rasterization (float atomics in both: 1e-5 of the gradient scale).
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import batch, nvdr
from diffrp_amd.camera import PerspectiveCamera, clip_space
from diffrp_amd.geometry import fibonacci_directions, noisy_icosphere

pytestmark = pytest.mark.gpu


def _scene(subdivs, resolutions, sigma, seed):
    rng = np.random.default_rng(seed)
    dirs = fibonacci_directions(len(subdivs))
    pos, tri = [], []
    for i, (s, (H, W)) in enumerate(zip(subdivs, resolutions)):
        v, f = noisy_icosphere(int(s), 0.8, sigma, seed=int(rng.integers(1 << 30)))
        cam = PerspectiveCamera.looking_at_origin_from(np.asarray(dirs[i]) * 3.2, h=H, w=W)
        pos.append(clip_space(v, cam.P('cpu') @ cam.V('cpu')).contiguous())
        tri.append(f.int().contiguous())
    return pos, tri


def _bits(x):
    return np.ascontiguousarray(x, np.float32).view(np.uint32)


def _assert_same(a, b, what):
    a, b = _bits(a), _bits(b)
    bad = np.argwhere(a != b)
    assert bad.size == 0, f'{what}: {len(bad)} words differ, first at {bad[0].tolist()}'


SMALL_S = [2, 3, 4, 2, 5, 3]
SMALL_RES = [(64, 64), (96, 160), (128, 128), (200, 120), (256, 256), (48, 80)]


def test_batch_matches_oracle_per_item(cuda):
    print('case: 6 meshes s=2..5, resolutions 48x80..256x256', flush=True)
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 1)
    ctx = nvdr.RasterizeCudaContext()
    rasts, dbs = batch.rasterize_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], SMALL_RES)
    for i, (p, t, res) in enumerate(zip(pos, tri, SMALL_RES)):
        o, odb = oracle.rasterize(p.numpy(), t.numpy(), res, with_db=True)
        _assert_same(rasts[i].cpu().numpy(), o, f'item {i} rast')
        _assert_same(dbs[i].cpu().numpy(), odb, f'item {i} rast_db')
        assert (rasts[i][..., 3] > 0).any()


@pytest.mark.parametrize('tile', [(8, 8), (16, 8), (16, 16), (32, 8)])
def test_batch_tiles_and_spills(cuda, tile):
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 2)
    ref = None
    for cap in (1, 64, 1024):  # cap 1: nearly everything spills
        ctx = nvdr.RasterizeCudaContext(tile=tile, bin_capacity=cap)
        rasts, _ = batch.rasterize_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], SMALL_RES,
                                         grad_db=False)
        got = [r.cpu().numpy() for r in rasts]
        if ref is None:
            ref = [oracle.rasterize(p.numpy(), t.numpy(), res) for p, t, res in zip(pos, tri, SMALL_RES)]
        for i in range(len(got)):
            _assert_same(got[i], ref[i], f'tile {tile} cap {cap} item {i}')


def test_batch_depth_peeling(cuda):
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 3)
    ctx = nvdr.RasterizeCudaContext()
    peeler = batch.BatchDepthPeeler(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], SMALL_RES, grad_db=False)
    layers = [peeler.rasterize_next_layer()[0] for _ in range(3)]
    for i, (p, t, res) in enumerate(zip(pos, tri, SMALL_RES)):
        prev = None
        for k in range(3):
            o = oracle.rasterize(p.numpy(), t.numpy(), res, prev=prev)
            _assert_same(layers[k][i].cpu().numpy(), o, f'item {i} layer {k}')
            prev = o


def _loss_grads(pos, tri, res, attrs, targets, batched, ctx):
    ps = [p.cuda().requires_grad_(True) for p in pos]
    ts = [t.cuda() for t in tri]
    if batched:
        rasts, _ = batch.rasterize_batch(ctx, ps, ts, res, grad_db=False)
    else:
        rasts = [nvdr.rasterize(ctx, p, t, r, grad_db=False)[0] for p, t, r in zip(ps, ts, res)]
    loss = 0.0
    for p, t, r, a, tg in zip(ps, ts, rasts, attrs, targets):
        col, _ = nvdr.interpolate(a, r, t)
        y = nvdr.antialias(col, r, p, t)
        loss = loss + ((y - tg) ** 2).sum()
    loss.backward()
    return [p.grad.cpu().numpy() for p in ps], float(loss)


def test_batch_gradients_match_per_item(cuda):
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 4)
    g = torch.Generator().manual_seed(5)
    attrs = [torch.rand(p.shape[0], 4, generator=g).cuda() for p in pos]
    targets = [torch.rand(1, H, W, 4, generator=g).cuda() for H, W in SMALL_RES]
    ctx = nvdr.RasterizeCudaContext()
    gb, lb = _loss_grads(pos, tri, SMALL_RES, attrs, targets, True, ctx)
    gp, lp = _loss_grads(pos, tri, SMALL_RES, attrs, targets, False, ctx)
    assert abs(lb - lp) <= 1e-6 * abs(lp)
    for i, (a, b) in enumerate(zip(gb, gp)):
        scale = max(np.abs(b).max(), 1e-30)
        err = np.abs(a - b).max()
        assert err <= 1e-5 * scale, f'item {i}: max |diff| {err:.3e}, scale {scale:.3e}'
        assert np.abs(b).max() > 0


def test_uneven_batch_full_size(cuda):
    """32 meshes, subdivision uniform in 3..8, resolution in {256, 512, 1024, 2048}^2,
    N(0, 0.02) vertex noise, 4 attribute channels (draws from seed 3): the
    batch, autotuned, against per-item rasterization, forward and backward."""
    rng = np.random.default_rng(3)
    subdivs = rng.integers(3, 9, 32)
    sizes = rng.choice([256, 512, 1024, 2048], 32)
    res = [(int(s), int(s)) for s in sizes]
    print(f'case: 32 meshes, faces {int(20 * 4 ** subdivs.min())}..{int(20 * 4 ** subdivs.max())}, '
          f'{int(sum(s * s for s in sizes))} pixels', flush=True)
    pos, tri = _scene(subdivs, res, 0.02, 3)
    ctx = nvdr.RasterizeCudaContext()
    tuned = batch.tune_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], res)
    print('autotune', tuned, flush=True)
    pc, tc = [p.cuda() for p in pos], [t.cuda() for t in tri]
    rasts, _ = batch.rasterize_batch(ctx, pc, tc, res, grad_db=False)
    for i, (p, t, r) in enumerate(zip(pc, tc, res)):
        single, _ = nvdr.rasterize(ctx, p, t, r, grad_db=False)
        assert torch.equal(rasts[i].view(torch.int32), single.view(torch.int32)), f'item {i}'
    del rasts
    # throughput of the one-launch batch against a loop of per-item calls (forward rasterisation)
    import time
    with torch.no_grad():
        def timed(fn):
            fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(5):
                fn()
            torch.cuda.synchronize()
            return (time.perf_counter() - t0) / 5
        tb = timed(lambda: batch.rasterize_batch(ctx, pc, tc, res, grad_db=False))
        tl = timed(lambda: [nvdr.rasterize(ctx, p, t, r, grad_db=False) for p, t, r in zip(pc, tc, res)])
    npx = sum(h * w for h, w in res)
    print(f'forward: batch {tb * 1e3:.2f} ms ({npx / tb / 1e6:.0f} Mpixel/s), per-item loop {tl * 1e3:.2f} ms',
          flush=True)
    g = torch.Generator().manual_seed(6)
    attrs = [torch.rand(p.shape[0], 4, generator=g).cuda() for p in pos]
    targets = [torch.rand(1, H, W, 4, generator=g).cuda() for H, W in res]
    gb, lb = _loss_grads(pos, tri, res, attrs, targets, True, ctx)
    gp, lp = _loss_grads(pos, tri, res, attrs, targets, False, ctx)
    assert abs(lb - lp) <= 1e-6 * abs(lp)
    for i, (a, b) in enumerate(zip(gb, gp)):
        err, scale = np.abs(a - b).max(), max(np.abs(b).max(), 1e-30)
        assert err <= 1e-5 * scale, f'item {i}: max |diff| {err:.3e}, scale {scale:.3e}'
