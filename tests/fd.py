"""Finite-difference harness on the float64 oracle (visibility frozen at a given id buffer).

The analytic gradients of rasterize / interpolate / antialias are derivatives
of the output with the triangle-id buffer held fixed; ``pipeline_f64``
re-evaluates the whole chain for perturbed inputs with the ids of ``rast_ids``
and is differentiated by central differences.
"""
import numpy as np

import oracle


def ids_f64(rast_f32):
    """Id buffer of a float32 rast in the float64 oracle's convention (plain triangle + 1)."""
    r = np.zeros(rast_f32.shape, np.float64)
    r[..., 3] = oracle.float_to_triidx(rast_f32[..., 3])
    return r


def clip_f64(verts, VP):
    h = np.concatenate([verts, np.ones((len(verts), 1))], 1)
    return np.einsum('vk,bjk->bvj', h, VP)          # [B, V, 4]


def pipeline_f64(verts, attr, tri, VP, rast_ids, target, opp):
    """Loss sum((AA(interp(attr, rast(pos))) - target)^2) at frozen visibility; float64."""
    clip = clip_f64(verts, VP)
    rast = oracle.rasterize_fixed(clip, tri, rast_ids, dtype=np.float64)
    col = oracle.interpolate(attr, rast, tri, dtype=np.float64)
    y = oracle.antialias(col, rast, clip, tri, opp, dtype=np.float64)
    return float(((y - target) ** 2).sum())


def pipeline_grad_f64(verts, attr, tri, VP, rast_ids, target, opp):
    """Analytic gradient of ``pipeline_f64`` w.r.t. (verts, attr), float64 oracle backward."""
    clip = clip_f64(verts, VP)
    rast = oracle.rasterize_fixed(clip, tri, rast_ids, dtype=np.float64)
    col = oracle.interpolate(attr, rast, tri, dtype=np.float64)
    y = oracle.antialias(col, rast, clip, tri, opp, dtype=np.float64)
    gy = 2.0 * (y - target)
    gcol, gclip_aa = oracle.antialias_bwd(col, rast, clip, tri, gy, opp, dtype=np.float64)
    gattr, grast = oracle.interpolate_bwd(attr, rast, tri, gcol, dtype=np.float64)
    gclip = oracle.rasterize_bwd(clip, tri, rast, grast, dtype=np.float64) + gclip_aa
    gverts = np.einsum('bvj,bjk->vk', gclip, VP)[:, :3]
    return gverts, gattr


def crop_vp(VP, H, W, r0, c0, h, w):
    """Projection of the h x w window at (r0, c0) of an H x W image, as an image of
    its own: pixel (r, c) of the window has the NDC centre of pixel (r0 + r, c0 + c).
    x' = (W/w) x + ((W - 2 c0 - w)/w) w_clip (y alike); z and w are unchanged, so
    barycentrics, depth and screen orientation are those of the full image."""
    M = np.eye(4)
    M[0, 0], M[0, 3] = W / w, (W - 2 * c0 - w) / w
    M[1, 1], M[1, 3] = H / h, (H - 2 * r0 - h) / h
    return M @ VP


def vertex_window(ids, tri, k, margin=2):
    """Pixel window [r0, r1) x [c0, c1) outside which the loss does not depend on
    vertex k at frozen visibility: the pixels that show a triangle incident to k,
    grown by ``margin`` (antialias pairs reach one pixel further).  ids: [H, W]
    plain triangle + 1.  None when no such pixel exists."""
    inc = np.flatnonzero((tri == k).any(1)) + 1
    rr, cc = np.nonzero(np.isin(ids, inc))
    if rr.size == 0:
        return None
    H, W = ids.shape
    return (max(rr.min() - margin, 0), min(rr.max() + 1 + margin, H),
            max(cc.min() - margin, 0), min(cc.max() + 1 + margin, W))


def window_loss_fn(verts, attr, tri, VP, rast_ids, target, opp, win):
    """Closure: ``pipeline_f64`` restricted to a window of a single view; it
    differs from the full-image loss by a constant in the window's vertex."""
    r0, r1, c0, c1 = win
    H, W = rast_ids.shape[1:3]
    VPw = crop_vp(VP[0], H, W, r0, c0, r1 - r0, c1 - c0)[None]
    ids_w = np.ascontiguousarray(rast_ids[:, r0:r1, c0:c1])
    tgt_w = np.ascontiguousarray(target[:, r0:r1, c0:c1])
    return lambda: pipeline_f64(verts, attr, tri, VPw, ids_w, tgt_w, opp)  # noqa: E731


def central_diff(f, x, idx, h):
    x0 = x[idx]
    x[idx] = x0 + h
    fp = f()
    x[idx] = x0 - h
    fm = f()
    x[idx] = x0
    return (fp - fm) / (2 * h)


def fd_check(f, x, grad, coords, h, rtol, floor):
    """Central differences at ``coords`` (flat indices of x).  A coordinate is
    compared only where differences with h and h/2 agree to a quarter of the
    tolerance (no kink of the piecewise-smooth antialias within h); returns
    (n_checked, worst rel err)."""
    worst, checked = 0.0, 0
    flat = x.reshape(-1)
    gflat = grad.reshape(-1)
    for i in coords:
        d1 = central_diff(f, flat, i, h)
        d2 = central_diff(f, flat, i, h / 2)
        if abs(d1 - d2) > 0.25 * rtol * max(abs(d2), floor):
            continue
        err = abs(gflat[i] - d2) / max(abs(d2), floor)
        worst = max(worst, err)
        checked += 1
        assert err <= rtol, f'coordinate {i}: analytic {gflat[i]:.9g} vs FD {d2:.9g} (rel {err:.2e})'
    return checked, worst
