"""
diffrp's PyTorch glue on the rasterisation path, restated on torch-CPU -- TEST
INFRASTRUCTURE ONLY (the oracle of the interpolation / shading side).

Each function follows the cited reference lines; they run on CPU tensors
(float32 like the reference, or float64 for finite differences).
"""
from typing import Callable, List, Optional

import numpy as np
import torch
import torch.nn.functional as F


def float_to_triidx(x: torch.Tensor) -> torch.Tensor:
    """diffrp/rendering/interpolator.py:24-25."""
    return torch.where(x <= 16777216, x.int(), x.view(torch.int32) - 0x4a800000)


def triidx_to_float(x: torch.Tensor) -> torch.Tensor:
    """diffrp/rendering/interpolator.py:28-29."""
    return torch.where(x <= 0x01000000, x.float(), (0x4a800000 + x).view(torch.float32))


def polyfill_interpolate(vertex_buffer, vi_data, tri_idx, empty_region: Optional[float] = None):
    """diffrp/rendering/interpolator.py:32-65: fma(v1 - v3, u, fma(v2 - v3, v, v3))
    with fma(x, y, b) = torch.addcmul(b, x, y) (diffrp/utils/shader_ops.py:77-81)."""
    v123 = vertex_buffer[tri_idx]
    v1, v2, v3 = v123.select(-2, 0), v123.select(-2, 1), v123.select(-2, 2)
    u, v = vi_data[..., 0:1], vi_data[..., 1:2]
    result = torch.addcmul(torch.addcmul(v3, v2 - v3, v), v1 - v3, u)
    if empty_region is not None:
        result = torch.where(vi_data[..., -1:] > 0, result, empty_region)
    return result


class MaskedSparseInterpolator:
    """diffrp/rendering/interpolator.py:87-97."""

    def __init__(self, vi_data, tris, mask):
        self.tris = tris
        self.indices = mask.nonzero(as_tuple=True)
        self.vi_data = vi_data[self.indices]
        self.tri_idx = tris[float_to_triidx(self.vi_data[..., -1]) - 1].long()

    def interpolate(self, vertex_buffer):
        return polyfill_interpolate(vertex_buffer, self.vi_data, self.tri_idx)


def normalized(x):
    """diffrp/utils/shader_ops.py:353-360."""
    return F.normalize(x, dim=-1)


def transform_point(xyz, matrix):
    """diffrp/utils/shader_ops.py:123-136."""
    return torch.matmul(F.pad(xyz, (0, 1), value=1), matrix.T)


def transform_point4x3(xyz, matrix):
    """diffrp/utils/shader_ops.py:139-152."""
    return torch.addmm(matrix[:-1, -1], xyz, matrix[:-1, :-1].T)


def transform_vector3x3(xyz, matrix):
    """diffrp/utils/shader_ops.py:171-185."""
    return torch.matmul(xyz, matrix[:-1, :-1].T)


def face_soup_flat(verts, tris):
    """MeshObject.preprocess with normals='flat' (diffrp/scene/objects.py:87-97,
    diffrp/utils/geometry.py:5-41): vertex 3f+i = corner i of face f."""
    fl = tris.long()
    fn = normalized(torch.linalg.cross(verts[fl[:, 1]] - verts[fl[:, 0]], verts[fl[:, 2]] - verts[fl[:, 0]]))
    sv = verts[fl].reshape(-1, 3)
    st = torch.arange(fl.numel(), dtype=torch.int32).reshape(fl.shape)
    sn = torch.stack([fn] * 3, 1).reshape(-1, 3)
    return sv, st, sn


def gbuffer_collect_single(rast, tris, stencils, values_fn: Callable, default: List[float]):
    """One-object stencil-masked G-buffer of one layer (surface_deferred.py:283-290,
    :324-332; mixin.py:130-155): default-filled [1,H,W,C], then index_put of the
    material's [P,C] values at the object's pixels."""
    vi_idx = float_to_triidx(rast[..., -1])
    stencil_buf = stencils[vi_idx]
    interp = MaskedSparseInterpolator(rast, tris, stencil_buf == 1)
    H, W = rast.shape[1:3]
    buf = torch.full([1, H, W, len(default)], default[0], dtype=rast.dtype) if default == [default[0]] * len(default) \
        else torch.tensor(default, dtype=rast.dtype).expand(1, H, W, len(default))
    if len(interp.indices[0]) == 0:
        return buf, interp
    return buf.index_put(interp.indices, values_fn(interp)), interp


def vertex_array_object(objects):
    """diffrp/rendering/mixin.py:74-113 for objects given as dicts with 'verts'
    [V,3], 'tris' [F,3], 'M' [4,4] and per-vertex attributes: world positions
    (transform_point4x3 with each object's M), triangles offset by the vertices
    of the objects before, stencils [0] ++ (object index + 1 per triangle), and
    every other attribute concatenated."""
    world, tris, sts, offset = [], [], [torch.zeros(1, dtype=torch.int64)], 0
    for s, o in enumerate(objects):
        world.append(transform_point4x3(o['verts'], o['M']))
        tris.append(o['tris'].long() + offset)
        sts.append(torch.full([len(o['tris'])], s + 1, dtype=torch.int64))
        offset += len(o['verts'])
    keys = [k for k in objects[0] if k not in ('verts', 'tris', 'M')]
    out = {'world_pos': torch.cat(world), 'tris': torch.cat(tris).int(), 'stencils': torch.cat(sts).int(),
           'verts': torch.cat([o['verts'] for o in objects])}
    for k in keys:
        out[k] = torch.cat([o[k] for o in objects])
    return out


def gbuffer_collect_layer(rast, tris, stencils, n_obj, operator: Callable, default: List[float]):
    """One layer of the stencil-masked G-buffer path with any number of objects
    (surface_deferred.py:283-293 and :324-332, mixin.py:130-155): per object an
    interpolator over the pixels whose stencil is its index + 1, skipped when
    empty; ``operator(object index, interpolator)`` gives its [P,C] values or
    None; the default-filled [1,H,W,C] buffer receives one index_put (the
    objects' indices and values concatenated when more than one returns values)."""
    stencil_buf = stencils[float_to_triidx(rast[..., -1])]
    H, W = rast.shape[1:3]
    buf = torch.tensor(default, dtype=rast.dtype).expand(1, H, W, len(default))
    indices, values, nind = [], [], 0
    for pi in range(n_obj):
        interp = MaskedSparseInterpolator(rast, tris, stencil_buf == pi + 1)
        if len(interp.indices[0]) == 0:
            continue
        op = operator(pi, interp)
        if op is not None:
            nind = len(interp.indices)
            indices.append(interp.indices)
            values.append(op)
    if len(indices) == 1:
        buf = buf.index_put(indices[0], values[0])
    elif nind != 0:
        buf = buf.index_put(tuple(torch.cat([x[i] for x in indices]) for i in range(nind)), torch.cat(values))
    return buf


def alpha_blend(bgc, bga, fgc, fga):
    """diffrp/utils/composite.py:28-29."""
    return torch.lerp(bgc, fgc, fga), torch.addcmul(fga, bga, 1 - fga)


def compose_layers(colors, alphas, return_alpha=True):
    """surface_deferred.py:348-384 with alpha_blend; result is top row first."""
    fb, fa = colors[-1], alphas[-1]
    for g, a in zip(reversed(colors[:-1]), reversed(alphas[:-1])):
        fb, fa = alpha_blend(fb, fa, g, a)
    out = torch.cat([fb, fa], -1) if return_alpha else fb
    return torch.flipud(out.squeeze(0))


def false_color_camera_space_normal(rast, verts, tris, normals, M, V):
    """The README scene's output (surface_deferred.py:531-543 via :505-516,
    :422-428, :418-420, mixin.py:115-128, base_material.py:183-210), opaque,
    one object, on a rasterised layer ``rast`` [1,H,W,4]."""
    stencils = torch.tensor([0] + [1] * len(tris), dtype=torch.int32)
    wn_vert = normalized(transform_vector3x3(normals, M))
    nrm, _ = gbuffer_collect_single(rast, tris, stencils, lambda it: normalized(it.interpolate(wn_vert)), [0., 0., 0.])
    alpha, _ = gbuffer_collect_single(rast, tris, stencils, lambda it: torch.ones_like(it.vi_data[..., :1]), [0.])
    cam = normalized(transform_vector3x3(nrm, V))
    n = compose_layers([cam], [alpha])
    return torch.cat([n[..., :3] * 0.5 + 0.5, n[..., 3:]], -1)


def sign2d(p1, p2, p3):
    """diffrp/utils/geometry.py:68-72."""
    return (p1[..., 0:1] - p3[..., 0:1]) * (p2[..., 1:2] - p3[..., 1:2]) - \
        (p2[..., 0:1] - p3[..., 0:1]) * (p1[..., 1:2] - p3[..., 1:2])


def point_in_tri2d(pt, v1, v2, v3):
    """diffrp/utils/geometry.py:75-91: inside unless the three signs are mixed."""
    d = [sign2d(pt, v1, v2), sign2d(pt, v2, v3), sign2d(pt, v3, v1)]
    neg = (d[0] < 0) | (d[1] < 0) | (d[2] < 0)
    pos = (d[0] > 0) | (d[1] > 0) | (d[2] > 0)
    return ~(neg & pos)


def edge_gradient_backward(grad_output, frag_xy, flat_tris_xy, rast, rgb):
    """Backward of _EdgeGradient (surface_deferred.py:127-169), bottom-row-first
    [H, W, .] tensors: returns (grad_frag_xy [H,W,2], grad_rgb).  Pairs along one
    axis give e = sign * 0.5 <g_a + g_b, c_a - c_b> to a where b is empty or a's
    fragment lies in b's triangle, and to b symmetrically; the zero padding of
    the one-shorter pair arrays (torch_redstone.supercat in the reference) is
    written as explicit pads."""
    H, W = frag_xy.shape[:2]
    tid = float_to_triidx(rast[..., 3:4])
    empty = tid == 0
    scr = flat_tris_xy[(tid.squeeze(-1) - 1).clamp(min=0).long()]          # [H, W, 3, 2]

    def axis_grad(dim, sign):
        n = frag_xy.shape[dim]
        A = lambda t: t.narrow(dim, 0, n - 1)  # noqa: E731
        B = lambda t: t.narrow(dim, 1, n - 1)  # noqa: E731
        e = sign * 0.5 * ((A(grad_output) + B(grad_output)) * (A(rgb) - B(rgb))).sum(-1, keepdim=True)
        edge = A(tid) != B(tid)
        sb, sa = B(scr), A(scr)
        a_in_b = B(empty) | point_in_tri2d(A(frag_xy), sb[..., 0, :], sb[..., 1, :], sb[..., 2, :])
        b_in_a = A(empty) | point_in_tri2d(B(frag_xy), sa[..., 0, :], sa[..., 1, :], sa[..., 2, :])
        ga = (edge & a_in_b).float() * e
        gb = (edge & b_in_a).float() * e
        zero = torch.zeros_like(ga.narrow(dim, 0, 1))
        return torch.cat([ga, zero], dim) + torch.cat([zero, gb], dim)

    gy = axis_grad(0, -1.0)
    gx = axis_grad(1, 1.0)
    return torch.cat([gx * (W / 2), gy * (H / 2)], -1), grad_output.clone()


def sample2d(texture2d, texcoords, wrap: str = "border", mode: str = "bilinear"):
    """diffrp/utils/shader_ops.py:197-252: grid_sample on the (C, H, W) view of an
    (H, W, C) texture at x = 2u - 1, y = -(2v - 1); 'latlong' = reflection with
    align_corners; 'cyclic' = the 2 x 2 tiled texture at (uv mod 1) / 2 moved
    into [1/4, 3/4), reflection padding."""
    shape = texcoords.shape
    tc = texcoords.reshape(1, 1, -1, shape[-1])
    align = False
    if wrap == "latlong":
        align, wrap = True, "reflection"
    if wrap == "cyclic":
        texture2d = texture2d.tile(2, 2, 1)
        tc = (tc % 1.0) * 0.5
        tc = torch.where(tc >= 0.25, tc, tc + 0.5)
        wrap = "reflection"
    tc = (tc * 2 - 1) * torch.tensor([1.0, -1.0], dtype=tc.dtype)
    out = F.grid_sample(texture2d[None].permute(0, 3, 1, 2), tc, padding_mode=wrap, mode=mode, align_corners=align)
    return out.reshape(texture2d.shape[-1], -1).T.reshape(*shape[:-1], texture2d.shape[-1])


def sample3d(texture3d, texcoords, wrap: str = "border", mode: str = "bilinear"):
    """diffrp/utils/shader_ops.py:255-310: grid_sample on the (C, D, H, W) view
    of a (D, H, W, C) texture at (2u - 1, -(2v - 1), 2w - 1)."""
    shape = texcoords.shape
    tc = texcoords.reshape(1, 1, 1, -1, shape[-1])
    tc = (tc * 2 - 1) * torch.tensor([1.0, -1.0, 1.0], dtype=tc.dtype)
    align = False
    if wrap == "latlong":
        align, wrap = True, "reflection"
    out = F.grid_sample(texture3d[None].permute(0, 4, 1, 2, 3), tc, padding_mode=wrap, mode=mode,
                        align_corners=align)
    return out.reshape(texture3d.shape[-1], -1).T.reshape(*shape[:-1], texture3d.shape[-1])


def to_uint8(x: torch.Tensor) -> np.ndarray:
    """diffrp/utils/exchange.py:7-18: (saturate(x) * 255).byte()  (truncation)."""
    return (torch.clamp(x.float(), 0.0, 1.0) * 255).byte().numpy()
