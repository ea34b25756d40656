"""Seeded scenes shared by the tests (CPU tensors; the GPU tests move them)."""
import numpy as np
import torch

from diffrp_amd.camera import PerspectiveCamera, clip_space
from diffrp_amd.geometry import fibonacci_directions, icosphere, noisy_icosphere


def readme_clip(res=512, subdiv=3, radius=0.8, flat=True):
    """README scene geometry (icosphere r=0.8, default camera) in clip space (float32, CPU)."""
    v, f = icosphere(subdiv, radius)
    verts = torch.from_numpy(v).float()
    tris = torch.from_numpy(f).int()
    cam = PerspectiveCamera(h=res, w=res)
    VP = cam.P('cpu') @ cam.V('cpu')
    return clip_space(verts, VP), tris, verts, cam


def noisy_clip(subdiv=6, sigma=0.01, res=1024, radius=0.8, seed=0, views=None):
    """Noisy icosphere seen by the default camera ([V,4]) or by ``views`` Fibonacci cameras ([B,V,4])."""
    verts, tris = noisy_icosphere(subdiv, radius, sigma, seed)
    if views is None:
        cam = PerspectiveCamera(h=res, w=res)
        return clip_space(verts, cam.P('cpu') @ cam.V('cpu')), tris, verts
    VP = torch.stack([vp_from_direction(d, res) for d in fibonacci_directions(views)])
    return clip_space(verts, VP), tris, verts


def vp_from_direction(d, res, dist=3.2):
    cam = PerspectiveCamera.looking_at_origin_from(np.asarray(d) * dist, h=res, w=res)
    return cam.P('cpu') @ cam.V('cpu')


def random_soup(n_tri, seed, spread=1.2, w_range=(0.5, 2.0)):
    """Random clip-space triangles (w > 0) as an indexed soup."""
    g = torch.Generator().manual_seed(seed)
    xyz = (torch.rand(n_tri * 3, 3, generator=g) * 2 - 1) * spread
    w = torch.rand(n_tri * 3, 1, generator=g) * (w_range[1] - w_range[0]) + w_range[0]
    pos = torch.cat([xyz[:, :2] * w, xyz[:, 2:3] * 0.9 * w, w], 1)
    tri = torch.arange(n_tri * 3, dtype=torch.int32).reshape(n_tri, 3)
    return pos.contiguous(), tri
