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
