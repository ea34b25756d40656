"""
Cameras of the rasterisation path: GL projection and view matrices.

Closed forms of diffrp's camera classes (``diffrp/rendering/camera.py:10-137``),
which build ``P`` with ``calibur.projection_gl_persp`` for the centred case and
``V`` as the inverse of the camera pose (trimesh transformations).  Matrices are
computed in float64 on the host and returned as float32 tensors, like
``gpu_f32`` does in the reference (``diffrp/utils/shader_ops.py:28-38``).
"""
import math
from typing import Sequence, Tuple, Union

import numpy as np
import torch


def _translation(p) -> np.ndarray:
    t = np.eye(4)
    t[:3, 3] = p
    return t


def _normalized(x: np.ndarray) -> np.ndarray:
    return x / np.linalg.norm(x)


def perspective_gl(fov_deg: float, h: int, w: int, near: float, far: float) -> np.ndarray:
    """GL projection for a centred pinhole camera, focal from the vertical FOV.

    f = h / (2 tan(fov/2)) (fx = fy); P = [[2f/w,0,0,0],[0,2f/h,0,0],
    [0,0,-(far+near)/(far-near),-2 far near/(far-near)],[0,0,-1,0]]
    (``diffrp/rendering/camera.py:80-88``).
    """
    f = h / (2.0 * math.tan(math.radians(fov_deg) / 2.0))
    p = np.zeros((4, 4))
    p[0, 0] = 2.0 * f / w
    p[1, 1] = 2.0 * f / h
    p[2, 2] = -(far + near) / (far - near)
    p[2, 3] = -2.0 * far * near / (far - near)
    p[3, 2] = -1.0
    return p


class Camera:
    """Camera pose ``t`` (camera-to-world); default at (0, 0, 3.2) looking down -Z
    (``diffrp/rendering/camera.py:18-32``)."""

    def __init__(self) -> None:
        self.t = _translation([0.0, 0.0, 3.2])

    def V(self, device: Union[str, torch.device] = 'cuda') -> torch.Tensor:
        return torch.tensor(np.linalg.inv(self.t).astype(np.float32), device=device)

    def P(self, device: Union[str, torch.device] = 'cuda') -> torch.Tensor:
        raise NotImplementedError

    def resolution(self) -> Tuple[int, int]:
        raise NotImplementedError


class RawCamera(Camera):
    """View/projection given directly as tensors (``diffrp/rendering/camera.py:49-66``)."""

    def __init__(self, h: int, w: int, v: torch.Tensor, p: torch.Tensor) -> None:
        self.h, self.w, self.v, self.p = h, w, v, p

    def V(self, device=None):
        return self.v

    def P(self, device=None):
        return self.p

    def resolution(self):
        return self.h, self.w


class PerspectiveCamera(Camera):
    """Perspective camera, angles in degrees (``diffrp/rendering/camera.py:68-137``)."""

    def __init__(self, fov=30, h=512, w=512, near=0.1, far=10.0) -> None:
        super().__init__()
        self.fov, self.h, self.w, self.near, self.far = fov, h, w, near, far

    def P(self, device: Union[str, torch.device] = 'cuda') -> torch.Tensor:
        return torch.tensor(
            perspective_gl(self.fov, self.h, self.w, self.near, self.far).astype(np.float32), device=device)

    def resolution(self):
        return self.h, self.w

    def set_transform(self, tr: np.ndarray):
        self.t = tr

    def lookat(self, point: Sequence[float]):
        """Aim -Z at ``point`` with +Y up: z = -fwd, x = fwd x up, y = x x fwd
        (``diffrp/rendering/camera.py:96-108``)."""
        point = np.asarray(point, dtype=np.float64)
        pos = self.t[:3, 3].copy()
        fwd = point - pos
        x = np.cross(fwd, [0.0, 1.0, 0.0])
        y = np.cross(x, fwd)
        self.t[:3, 0] = _normalized(x)
        self.t[:3, 1] = _normalized(y)
        self.t[:3, 2] = _normalized(-fwd)
        return self

    @classmethod
    def from_orbit(cls, h, w, radius, azim, elev, origin, fov=30, near=0.1, far=10.0):
        """Orbit camera: azimuth from +Z, elevation in degrees (``diffrp/rendering/camera.py:110-137``)."""
        cam = cls(h=h, w=w, fov=fov, near=near, far=far)
        theta, phi = math.radians(azim), math.radians(elev)
        z = radius * math.cos(theta) * math.cos(phi) + origin[2]
        x = radius * math.sin(theta) * math.cos(phi) + origin[0]
        y = radius * math.sin(phi) + origin[1]
        cam.t = _translation([x, y, z])
        cam.lookat(origin)
        return cam

    @classmethod
    def looking_at_origin_from(cls, position: Sequence[float], h=512, w=512, fov=30, near=0.1, far=10.0):
        cam = cls(fov=fov, h=h, w=w, near=near, far=far)
        cam.t = _translation(position)
        return cam.lookat([0.0, 0.0, 0.0])


class _ClipSpace(torch.autograd.Function):
    """pos [n,V,4] = VP_b [x, y, z, 1]^T on the HIP kernels of ``drp_clip.hip``;
    the backward sums the views into the world vertices in view order."""

    @staticmethod
    def forward(ctx, xyz, VP):
        from . import _lib
        from ._lib import check, ptr, stream_of
        n, V = VP.shape[0], xyz.shape[0]
        pos = torch.empty((n, V, 4), dtype=torch.float32, device=xyz.device)
        check(_lib.lib().drp_clip_transform_fwd(stream_of(xyz), ptr(xyz), V, ptr(VP), n, ptr(pos)), 'clip transform')
        ctx.save_for_backward(xyz, VP)
        return pos

    @staticmethod
    def backward(ctx, g):
        from . import _lib
        from ._lib import check, ptr, stream_of
        xyz, VP = ctx.saved_tensors
        g = g.float().contiguous()
        gx = gvp = None
        if ctx.needs_input_grad[0]:
            gx = torch.empty_like(xyz)
            check(_lib.lib().drp_clip_transform_bwd(stream_of(g), ptr(g), xyz.shape[0], ptr(VP), VP.shape[0],
                                                    ptr(gx)), 'clip transform backward')
        if ctx.needs_input_grad[1]:
            h = torch.nn.functional.pad(xyz, (0, 1), value=1.0)
            gvp = torch.matmul(g.transpose(-1, -2), h)
        return gx, gvp


def clip_space(world_pos: torch.Tensor, VP: torch.Tensor) -> torch.Tensor:
    """[x, y, z, 1] @ VP^T (``diffrp/utils/shader_ops.py:123-136``, ``surface_deferred.py:229-238``).

    ``VP`` may be [4,4] (one view, result [V,4]) or [B,4,4] (one clip-space
    buffer per view, [B,V,4]).  Device tensors run the HIP transform (one fma
    chain per row); host tensors, which only prepare inputs, use torch.
    """
    if not world_pos.is_cuda:
        h = torch.nn.functional.pad(world_pos, (0, 1), value=1.0)
        return torch.matmul(h, VP.transpose(-1, -2)).contiguous()
    if world_pos.dim() != 2 or world_pos.shape[-1] != 3 or VP.shape[-2:] != (4, 4) or VP.dim() not in (2, 3):
        raise ValueError('clip_space takes world_pos [V, 3] and VP [4, 4] or [B, 4, 4]')
    xyz = world_pos.float().contiguous()
    vp = VP.float().to(xyz.device).reshape(-1, 4, 4).contiguous()
    pos = _ClipSpace.apply(xyz, vp)
    return pos[0] if VP.dim() == 2 else pos
