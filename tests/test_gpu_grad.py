"""
Gradients of the whole GPU pipeline against float64 central differences.

Noisy icosphere s=6 (40,962 vertices, 81,920 triangles), one view at
1024 x 1024, 8 attribute channels: world vertices -> clip (float32 matmul on
the GPU) -> rasterize -> interpolate -> antialias -> L2 loss, backward on the
GPU, gradients w.r.t. world vertices and attributes.

Reference: the float64 oracle's loss as a function of the clip coordinates the
GPU received, with visibility frozen at the GPU's ids.  A world coordinate
(vertex k, axis j) moves clip vertex k along column j of the view-projection
matrix.  Differentiating at the float32 clip values (rather than at float64
clip values computed from the world vertices) keeps the rounding of the
inputs out of the comparison: on this mesh the silhouette slivers make the
gradient move by ~6e-4 under a relative input change of 6e-8.

Central differences use h = 1e-6 and h/2; where the two disagree (a kink of
the piecewise-smooth antialias within h) the step shrinks by 4 twice (down to
6.25e-8) before the coordinate is given up.  At this resolution a step of
1e-4 moves edges by ~0.06 pixel and crosses such a kink for most vertices; at
1e-6 float64 rounding is ~1e-8 of the derivatives.  Every coordinate given up
is printed, and at most one per set may be (none is, on this scene).  The loss is evaluated on the
window of pixels the coordinate can affect, which equals the full-image
difference (test_window_fd_equals_full_image_fd).
"""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import nvdr
from diffrp_amd.geometry import noisy_icosphere
from tests.fd import crop_vp, ids_f64, vertex_window
from tests.scenes import vp_from_direction

pytestmark = pytest.mark.gpu

H = W = 1024
C = 8


def _gpu(verts, attr, tri, VP, target):
    dev = torch.device('cuda')
    v = verts.to(dev).requires_grad_(True)
    a = torch.from_numpy(attr).to(dev).requires_grad_(True)
    t = tri.to(dev).int().contiguous()
    clip = (torch.nn.functional.pad(v, (0, 1), value=1.0) @ VP.to(dev).T)[None].contiguous()
    ctx = nvdr.RasterizeCudaContext(dev)
    rast, _ = nvdr.rasterize(ctx, clip, t, (H, W))
    col, _ = nvdr.interpolate(a, rast, t)
    y = nvdr.antialias(col, rast, clip, t)
    loss = ((y - torch.from_numpy(target).to(dev)) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    return (v.grad.double().cpu().numpy(), a.grad.double().cpu().numpy(), rast.detach().cpu().numpy(),
            clip.detach().double().cpu().numpy())


class _WindowLoss:
    """float64 loss on a pixel window, with clip vertex k or attribute row k moved."""

    def __init__(self, clip, attr, tri, ids, target, opp, win):
        r0, r1, c0, c1 = win
        self.M = crop_vp(np.eye(4), H, W, r0, c0, r1 - r0, c1 - c0)
        self.ids = np.ascontiguousarray(ids[:, r0:r1, c0:c1])
        self.target = np.ascontiguousarray(target[:, r0:r1, c0:c1])
        self.clip, self.attr, self.tri, self.opp = clip, attr, tri, opp

    def __call__(self, clip, attr):
        cw = (clip @ self.M.T)[None]
        rast = oracle.rasterize_fixed(cw, self.tri, self.ids, dtype=np.float64)
        col = oracle.interpolate(attr, rast, self.tri, dtype=np.float64)
        y = oracle.antialias(col, rast, cw, self.tri, self.opp, dtype=np.float64)
        return float(((y - self.target) ** 2).sum())

    def d_vertex(self, k, direction, h):
        def at(s):
            c = self.clip.copy()
            c[k] += s * direction
            return self(c, self.attr)
        return (at(h) - at(-h)) / (2 * h)

    def d_attr(self, k, j, h):
        def at(s):
            a = self.attr.copy()
            a[k, j] += s
            return self(self.clip, a)
        return (at(h) - at(-h)) / (2 * h)


def _compare(grad, coords, deriv, rtol, floor):
    checked, worst = 0, 0.0
    for i in coords:
        d2 = None
        for h in (1e-6, 2.5e-7, 6.25e-8):
            d1, dh = deriv(int(i), h), deriv(int(i), h / 2)
            if abs(d1 - dh) <= 0.25 * rtol * max(abs(dh), floor):
                d2 = dh
                break
        if d2 is None:  # an antialias kink within the smallest step
            print(f'  coordinate {i} given up: FD h / h/2 disagree down to h = 6.25e-8', flush=True)
            continue
        err = abs(grad[i] - d2) / max(abs(d2), floor)
        worst = max(worst, err)
        checked += 1
        assert err <= rtol, f'coordinate {i}: GPU {grad[i]:.9g} vs FD {d2:.9g} (rel {err:.2e})'
    return checked, worst


def test_pipeline_gradients_match_fd_1024(cuda):
    print(f'case: noisy s=6, 1 view, {H}x{W}, C={C}, 48 vertex + 16 attribute coordinates', flush=True)
    verts_t, tri_t = noisy_icosphere(6, 0.8, 0.01, seed=0)
    VP = vp_from_direction([0.3, 0.5, 0.8], H)
    rng = np.random.default_rng(21)
    attr = rng.random((len(verts_t), C)).astype(np.float32)
    target = rng.random((1, H, W, C)).astype(np.float32)
    gv, ga, rast, clip = _gpu(verts_t, attr, tri_t, VP, target)
    ids = ids_f64(rast)
    tri = tri_t.numpy()
    opp = oracle.topology(tri)
    VP64 = VP.double().numpy()
    attr64, target64 = attr.astype(np.float64), target.astype(np.float64)

    def loss_for(k):
        win = vertex_window(ids[0, ..., 3], tri, k)
        assert win is not None, f'vertex {k}: gradient without a visible incident triangle'
        return _WindowLoss(clip[0], attr64, tri, ids, target64, opp, win)

    gflat = gv.reshape(-1)
    gmax = np.abs(gv).max()
    nz = np.flatnonzero(np.abs(gflat) > 1e-3 * gmax)
    coords = np.random.default_rng(22).choice(nz, 48, replace=False)
    n, worst = _compare(gflat, coords, lambda i, h: loss_for(i // 3).d_vertex(i // 3, VP64[:, i % 3], h),
                        rtol=1e-4, floor=1e-3 * gmax)
    print(f'vertex coordinates checked {n}/48, worst rel err {worst:.2e}', flush=True)
    assert n >= 47
    aflat = ga.reshape(-1)
    amax = np.abs(ga).max()
    nz = np.flatnonzero(np.abs(aflat) > 1e-3 * amax)
    coords = np.random.default_rng(23).choice(nz, 16, replace=False)
    n, worst = _compare(aflat, coords, lambda i, h: loss_for(i // C).d_attr(i // C, i % C, h),
                        rtol=1e-4, floor=1e-3 * amax)
    print(f'attribute coordinates checked {n}/16, worst rel err {worst:.2e}', flush=True)
    assert n >= 15
