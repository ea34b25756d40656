"""
Gradients of the whole GPU pipeline against float64 central differences.

Noisy icosphere s=6 (40,962 vertices, 81,920 triangles), one view at
1024 x 1024, 8 attribute channels: world vertices -> clip (f32 matmul on the
GPU) -> rasterize -> interpolate -> antialias -> L2 loss, backward on the GPU.
The reference derivative is the float64 oracle's loss at the GPU's own
visibility (ids frozen), by central differences with h = 1e-6 (and h/2 to
skip coordinates where the piecewise-smooth antialias has a kink within h;
at this resolution a world-space step of 1e-4 moves edges by ~0.06 pixel and
crosses such a kink for most vertices, while at 1e-6 float64 rounding is
still ~1e-8 of the derivatives),
evaluated on the window of pixels that the coordinate can affect (equal to
the full-image difference, see test_window_fd_equals_full_image_fd).
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import nvdr
from diffrp_amd.geometry import noisy_icosphere
from tests.fd import central_diff, ids_f64, vertex_window, window_loss_fn
from tests.scenes import vp_from_direction

pytestmark = pytest.mark.gpu


def _gpu_grads(verts, attr, tri, VP, target, H, W):
    dev = torch.device('cuda')
    v = verts.to(dev).requires_grad_(True)
    a = torch.from_numpy(attr).to(dev).requires_grad_(True)
    t = tri.to(dev).int().contiguous()
    clip = torch.nn.functional.pad(v, (0, 1), value=1.0) @ VP.to(dev).T
    ctx = nvdr.RasterizeCudaContext(dev)
    rast, _ = nvdr.rasterize(ctx, clip[None].contiguous(), t, (H, W))
    col, _ = nvdr.interpolate(a, rast, t)
    y = nvdr.antialias(col, rast, clip[None].contiguous(), t)
    loss = ((y - torch.from_numpy(target).to(dev)) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    return (v.grad.double().cpu().numpy(), a.grad.double().cpu().numpy(), rast.cpu().numpy())


def _check(x, grad, coords, per_vertex, ids, tri, make_loss, rtol, floor, h=1e-6):
    flat, gflat = x.reshape(-1), grad.reshape(-1)
    checked, worst = 0, 0.0
    for i in coords:
        win = vertex_window(ids, tri, int(i) // per_vertex)
        assert win is not None, f'coordinate {i}: gradient without a visible incident triangle'
        f = make_loss(win)
        d1 = central_diff(f, flat, i, h)
        d2 = central_diff(f, flat, i, h / 2)
        if abs(d1 - d2) > 0.25 * rtol * max(abs(d2), floor):
            continue  # antialias kink within h
        err = abs(gflat[i] - d2) / max(abs(d2), floor)
        worst = max(worst, err)
        checked += 1
        assert err <= rtol, f'coordinate {i}: GPU {gflat[i]:.9g} vs FD {d2:.9g} (rel {err:.2e})'
    return checked, worst


def test_pipeline_gradients_match_fd_1024(cuda):
    print('case: noisy s=6, 1 view, 1024^2, C=8, 64 coordinates', flush=True)
    verts_t, tri_t = noisy_icosphere(6, 0.8, 0.01, seed=0)
    H = W = 1024
    VP = vp_from_direction([0.3, 0.5, 0.8], H)
    rng = np.random.default_rng(21)
    attr = rng.random((len(verts_t), 8)).astype(np.float32)
    target = rng.random((1, H, W, 8)).astype(np.float32)
    gv, ga, rast = _gpu_grads(verts_t, attr, tri_t, VP, target, H, W)
    ids = ids_f64(rast)
    tri = tri_t.numpy()
    opp = oracle.topology(tri)
    verts = verts_t.double().numpy()
    attr64 = attr.astype(np.float64)
    VP64 = VP.double().numpy()[None]
    target64 = target.astype(np.float64)

    def make_loss(win):
        return window_loss_fn(verts, attr64, tri, VP64, ids, target64, opp, win)

    gmax = np.abs(gv).max()
    nz = np.flatnonzero(np.abs(gv) > 1e-3 * gmax)
    coords = np.random.default_rng(22).choice(nz, 48, replace=False)
    n, worst = _check(verts, gv, coords, 3, ids[0, ..., 3], tri, make_loss, rtol=1e-4, floor=1e-3 * gmax)
    print(f'vertex coordinates checked {n}/48, worst rel err {worst:.2e}', flush=True)
    assert n >= 36
    amax = np.abs(ga).max()
    nz = np.flatnonzero(np.abs(ga) > 1e-3 * amax)
    coords = np.random.default_rng(23).choice(nz, 16, replace=False)
    n, worst = _check(attr64, ga, coords, 8, ids[0, ..., 3], tri, make_loss, rtol=1e-4, floor=1e-3 * amax)
    print(f'attribute coordinates checked {n}/16, worst rel err {worst:.2e}', flush=True)
    assert n >= 12
