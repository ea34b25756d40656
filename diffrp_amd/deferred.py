"""
Deferred render session on the MI355X rasteriser: the caller side of the path.

Mirrors the surface of diffrp's ``SurfaceDeferredRenderSession`` and its inputs
(``diffrp/rendering/surface_deferred.py:172-804``, ``mixin.py:12-155``,
``interpolator.py:24-97``, ``materials/base_material.py:12-382``,
``scene/objects.py:11-98``, ``scene/scene.py:10-31``) so that scenes, materials
and G-buffer queries written against diffrp run unchanged on top of
``diffrp_amd.nvdr``.  Shading stays PyTorch (as in diffrp); rasterisation,
interpolation and antialiasing are the HIP kernels.  Texture sampling for
materials is ``diffrp_amd.texture``; image-based lighting is not built.

Conventions kept from diffrp: G-buffers are produced bottom row first and the
composed images are flipped to top row first; background takes the G-buffer
default value with alpha 0; a session caches every query (one camera, one
scene snapshot).
"""
import functools
import threading
from dataclasses import dataclass, field
from typing import Callable, Dict, List, Optional, Sequence, Tuple, Union

import torch
import torch.nn.functional as F

from . import layered, nvdr
from .camera import Camera
from .geometry import compute_face_normals, compute_vertex_normals, make_face_soup

# ----------------------------------------------------------------------------
# small tensor helpers (GLSL-style channel access is written out explicitly)
# ----------------------------------------------------------------------------


def normalized(x: torch.Tensor) -> torch.Tensor:
    """Unit vectors along the last axis; zero vectors stay zero (eps 1e-12)."""
    return F.normalize(x, dim=-1)


def xform_point(xyz: torch.Tensor, m: torch.Tensor) -> torch.Tensor:
    """[x, y, z, 1] @ m^T -> homogeneous [..., 4]."""
    return torch.matmul(F.pad(xyz, (0, 1), value=1.0), m.transpose(-1, -2))


def xform_affine(xyz: torch.Tensor, m: torch.Tensor) -> torch.Tensor:
    """Affine point transform keeping 3 components: xyz @ m[:3,:3]^T + m[:3,3]."""
    return torch.addmm(m[:3, 3], xyz, m[:3, :3].t())


def xform_dir(xyz: torch.Tensor, m: torch.Tensor) -> torch.Tensor:
    """Direction transform by the linear part of an affine matrix."""
    return torch.matmul(xyz, m[:3, :3].t())


def float_to_triidx(x: torch.Tensor) -> torch.Tensor:
    """Decode the id channel of rast (triangle + 1, 0 = empty); values above 2^24
    are stored as the float with bits 0x4a800000 + n."""
    return torch.where(x <= 16777216, x.int(), x.view(torch.int32) - 0x4a800000)


def triidx_to_float(n: torch.Tensor) -> torch.Tensor:
    return torch.where(n <= 0x01000000, n.float(), (n + 0x4a800000).view(torch.float32))


def polyfill_interpolate(values: torch.Tensor, frag: torch.Tensor, tri_idx: torch.Tensor,
                         empty_region: Optional[float] = None) -> torch.Tensor:
    """Barycentric blend a0 u + a1 v + a2 (1 - u - v) of the vertex rows picked by
    ``tri_idx`` (``frag`` holds (u, v, z/w, id)), written as
    (a0 - a2) u + ((a1 - a2) v + a2) like the reference; optional fill for empty
    fragments."""
    corners = values[tri_idx]
    a0, a1, a2 = corners.unbind(-2)
    u, v = frag[..., 0:1], frag[..., 1:2]
    out = torch.addcmul(torch.addcmul(a2, a1 - a2, v), a0 - a2, u)
    if empty_region is not None:
        out = torch.where(frag[..., 3:4] > 0, out, empty_region)
    return out


def linear_to_srgb(rgb: torch.Tensor) -> torch.Tensor:
    """Linear -> sRGB (``diffrp/utils/colors.py:33-42``): 12.92 x below 0.0031308,
    else 1.055 x^(1/2.4) - 0.055 (as torch.add(-0.055, ., alpha=1.055));
    clamped at 1e-5 from below when differentiated."""
    if rgb.requires_grad:
        rgb = torch.clamp_min(rgb, 1e-5)
    return torch.where(rgb < 0.0031308, 12.92 * rgb, torch.add(-0.055, rgb ** (1 / 2.4), alpha=1.055))


def _memo(fn):
    """Per-instance memo keyed by the method's qualified name (one value per session)."""
    key = fn.__qualname__

    @functools.wraps(fn)
    def wrapper(self):
        store = self.__dict__.setdefault('_memo', {})
        if key not in store:
            store[key] = fn(self)
        return store[key]
    return wrapper


# ----------------------------------------------------------------------------
# interpolators (who sees which pixels)
# ----------------------------------------------------------------------------
class FullScreenInterpolator:
    """Every pixel of the layer; attributes come from the HIP interpolate kernel."""

    def __init__(self, rast: torch.Tensor, tris: torch.Tensor):
        self.vi_data = rast.contiguous()
        self.tris = tris.contiguous()

    def interpolate(self, values: torch.Tensor) -> torch.Tensor:
        return nvdr.interpolate(values.contiguous(), self.vi_data, self.tris)[0]


_FRAG_ROW = 8192  # fragments per row of the 2-D image the gathered fragments are laid out as


class MaskedSparseInterpolator:
    """Only the pixels where ``mask`` holds, as a flat batch [P, C]
    (``interpolator.py:87-97``).  ``indices`` are nonzero()'s row-major index
    tuple, ``vi_data`` the gathered fragments [P, 4]; ``interpolate`` runs the
    HIP interpolation kernel over the fragments as one 1 x P image (the same
    (a0 - a2) u + ((a1 - a2) v + a2) blend as diffrp's polyfill, without
    materialising the [P, 3, C] corner rows)."""

    def __init__(self, rast: torch.Tensor, tris: torch.Tensor, mask: Optional[torch.Tensor] = None,
                 pixels: Optional[torch.Tensor] = None):
        self.tris = tris
        if pixels is None:
            self.indices = mask.nonzero(as_tuple=True)
            self.vi_data = rast[self.indices]
        else:  # flat row-major pixel ids of a [1, H, W, 4] layer (from stencil_partition)
            W = rast.shape[2]
            p = pixels.long()
            self.indices = (torch.zeros_like(p), p // W, p % W)
            self.vi_data = rast.reshape(-1, 4).index_select(0, p)
        self._tri_idx = None

    @property
    def tri_idx(self) -> torch.Tensor:
        if self._tri_idx is None:
            self._tri_idx = self.tris[float_to_triidx(self.vi_data[..., 3]) - 1].long()
        return self._tri_idx

    def interpolate(self, values: torch.Tensor) -> torch.Tensor:
        P, C = self.vi_data.shape[0], values.shape[-1]
        if P == 0:
            return values.new_zeros(0, C)
        frag = self.vi_data.contiguous()
        rows = -(-P // _FRAG_ROW)
        if rows > 1:  # an image row holds at most _FRAG_ROW fragments; the padding is empty (id 0)
            frag = torch.cat([frag, frag.new_zeros(rows * _FRAG_ROW - P, 4)])
        frag = frag.view(1, rows, -1, 4)
        out, _ = nvdr.interpolate(values.float().contiguous(), frag, self.tris)
        return out.view(-1, C)[:P]


def stencil_partition(rast: torch.Tensor, stencils: torch.Tensor, n_owner: int):
    """Pixels of one layer [1, H, W, 4] grouped by owner = stencils[id]
    (``surface_deferred.py:283-284``): returns (order [H*W] int32 on the device,
    [(start, end)] per owner 0..n_owner on the host), with one host read for
    the whole layer.  Within an owner the pixels are in row-major order."""
    from . import _lib
    from ._lib import check, ptr, stream_of
    H, W = rast.shape[1], rast.shape[2]
    rast = rast.contiguous()
    st = stencils.int().contiguous()
    L = _lib.lib()
    order = torch.empty(H * W, dtype=torch.int32, device=rast.device)
    se = torch.empty(2, n_owner + 1, dtype=torch.int32, device=rast.device)
    ws_bytes = L.drp_stencil_partition_workspace_size(H, W, n_owner)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=rast.device)
    check(L.drp_stencil_partition(stream_of(rast), ptr(rast), H, W, ptr(st), st.numel() - 1, n_owner, ptr(order),
                                  ptr(se[0]), ptr(se[1]), ptr(ws), ws_bytes), 'stencil partition')
    return order, se.cpu().t().tolist()


# ----------------------------------------------------------------------------
# scene and materials
# ----------------------------------------------------------------------------
@dataclass
class MeshObject:
    """Triangle mesh with material, pose and per-vertex attributes
    (``diffrp/scene/objects.py:11-98``).  ``normals`` may be a tensor, 'flat'
    (face soup with face normals) or 'smooth' (normalised sums of face normals)."""
    material: 'SurfaceMaterial'
    verts: torch.Tensor
    tris: torch.Tensor
    normals: Union[str, torch.Tensor] = 'flat'
    M: Optional[torch.Tensor] = None
    color: Optional[torch.Tensor] = None
    uv: Optional[torch.Tensor] = None
    tangents: Optional[torch.Tensor] = None
    custom_attrs: Optional[Dict[str, torch.Tensor]] = None
    metadata: Optional[Dict] = None

    def preprocess(self) -> 'MeshObject':
        v = self.verts

        def filled(x, c, value):
            return x if x is not None else v.new_full((len(v), c), value)
        self.M = self.M if self.M is not None else torch.eye(4, device=v.device, dtype=v.dtype)
        self.color = filled(self.color, 4, 1.0)
        self.uv = filled(self.uv, 2, 0.0)
        self.tangents = filled(self.tangents, 4, 0.0)
        self.custom_attrs = dict(self.custom_attrs or {})
        self.metadata = dict(self.metadata or {})
        if isinstance(self.normals, str):
            if self.normals == 'smooth':
                self.normals = compute_vertex_normals(self.verts, self.tris)
            elif self.normals == 'flat':
                corners = self.tris.long()
                fn = compute_face_normals(self.verts, self.tris, normalize=True)
                self.verts, self.tris, self.normals = make_face_soup(self.verts, self.tris, fn)
                expand = lambda a: a[corners].flatten(0, 1)  # noqa: E731
                self.color, self.uv, self.tangents = expand(self.color), expand(self.uv), expand(self.tangents)
                self.custom_attrs = {k: expand(a) for k, a in self.custom_attrs.items()}
            else:
                raise ValueError(f"normals must be a tensor, 'flat' or 'smooth' (got {self.normals!r})")
        return self


class Scene:
    """Objects (and, in diffrp, lights) of a render (``diffrp/scene/scene.py:10-31``)."""

    def __init__(self) -> None:
        self.objects: List[MeshObject] = []
        self.lights: list = []
        self.metadata: dict = {}

    def add_mesh_object(self, obj: MeshObject) -> 'Scene':
        self.objects.append(obj.preprocess())
        return self

    def add_light(self, light) -> 'Scene':
        self.lights.append(light)
        return self


@dataclass
class VertexArrayObject:
    """All objects' vertex and index buffers concatenated (``base_material.py:12-34``)."""
    verts: torch.Tensor
    normals: torch.Tensor
    world_pos: torch.Tensor
    tris: torch.Tensor
    stencils: torch.Tensor
    color: torch.Tensor
    uv: torch.Tensor
    tangents: torch.Tensor
    custom_attrs: Dict[str, torch.Tensor]


@dataclass
class SurfaceUniform:
    """Per-object uniforms: model, view, projection (``base_material.py:37-76``)."""
    M: torch.Tensor
    V: torch.Tensor
    P: torch.Tensor

    @property
    def camera_matrix(self) -> torch.Tensor:
        return torch.linalg.inv(self.V)

    @property
    def camera_position(self) -> torch.Tensor:
        return self.camera_matrix[:3, 3]


class SurfaceInput:
    """Fragment attributes a material reads (``base_material.py:79-276``); every
    property is interpolated lazily and memoised."""

    def __init__(self, uniforms: SurfaceUniform, vao: VertexArrayObject, interpolator):
        self.uniforms = uniforms
        self.vertex_buffers = vao
        self.interpolator = interpolator
        self._cache: dict = {}

    def _cached(self, key, make):
        if key not in self._cache:
            self._cache[key] = make()
        return self._cache[key]

    def interpolate_ex(self, values: torch.Tensor, world_transform: str = 'none') -> torch.Tensor:
        M = self.uniforms.M
        if world_transform == 'point':
            values = xform_affine(values, M)
        elif world_transform == 'vector':
            values = xform_dir(values, M)
        elif world_transform == 'vectornor':
            values = normalized(xform_dir(values, M))
        elif world_transform == 'vector3ex1':
            values = torch.cat([xform_dir(values[..., :3], M), values[..., 3:4]], -1)
        elif world_transform == 'vector3norex1':
            values = torch.cat([normalized(xform_dir(values[..., :3], M)), values[..., 3:4]], -1)
        elif world_transform != 'none':
            raise ValueError(f'unknown world_transform {world_transform!r}')
        return self.interpolator.interpolate(values)

    @property
    def world_pos(self):
        return self._cached('world_pos', lambda: self.interpolate_ex(self.vertex_buffers.world_pos))

    @property
    def local_pos(self):
        return self._cached('local_pos', lambda: self.interpolate_ex(self.vertex_buffers.verts))

    @property
    def world_normal_unnormalized(self):
        return self._cached('wnu', lambda: self.interpolate_ex(self.vertex_buffers.normals, 'vectornor'))

    @property
    def world_normal(self):
        return self._cached('wn', lambda: normalized(self.world_normal_unnormalized))

    @property
    def view_dir(self):
        return self._cached('view_dir', lambda: normalized(self.world_pos - self.uniforms.camera_position))

    @property
    def color(self):
        return self._cached('color', lambda: self.interpolate_ex(self.vertex_buffers.color))

    @property
    def uv(self):
        return self._cached('uv', lambda: self.interpolate_ex(self.vertex_buffers.uv))

    @property
    def world_tangent(self):
        return self._cached('tangent', lambda: self.interpolate_ex(self.vertex_buffers.tangents, 'vector3norex1'))

    @property
    def custom_attrs(self):
        vb = self.vertex_buffers.custom_attrs
        return {k: self._cached(('custom', k), lambda k=k: self.interpolate_ex(vb[k])) for k in vb}


@dataclass
class SurfaceOutputStandard:
    """Material outputs collected into G-buffers (``base_material.py:300-363``)."""
    albedo: Optional[torch.Tensor] = None
    normal: Optional[torch.Tensor] = None
    emission: Optional[torch.Tensor] = None
    metallic: Optional[torch.Tensor] = None
    smoothness: Optional[torch.Tensor] = None
    occlusion: Optional[torch.Tensor] = None
    alpha: Optional[torch.Tensor] = None
    aovs: Optional[Dict[str, torch.Tensor]] = None
    normal_space: str = 'tangent'


class SurfaceMaterial:
    """A material maps fragment inputs to ``SurfaceOutputStandard`` (``base_material.py:366-382``)."""

    def shade(self, su: SurfaceUniform, si: SurfaceInput) -> SurfaceOutputStandard:
        raise NotImplementedError


class DefaultMaterial(SurfaceMaterial):
    """Albedo = interpolated vertex RGB, times an optional tint (``default_material.py:6-24``)."""

    def __init__(self, tint: Optional[torch.Tensor] = None) -> None:
        self.tint = tint

    def shade(self, su, si):
        rgb = si.color[..., :3]
        return SurfaceOutputStandard(rgb * self.tint if self.tint is not None else rgb)


# ----------------------------------------------------------------------------
# compositing
# ----------------------------------------------------------------------------
def alpha_blend(bgc, bga, fgc, fga):
    """Over: colour lerp(bg, fg, a_fg), alpha a_bg (1 - a_fg) + a_fg (``composite.py:28-29``)."""
    return torch.lerp(bgc, fgc, fga), torch.addcmul(fga, bga, 1 - fga)


def alpha_additive(bgc, bga, fgc, fga):
    """Additive emission: bg + fg a_fg, alpha clamped (``composite.py:36-37``)."""
    return bgc + fgc * fga, torch.clamp(bga + fga, 0.0, 1.0)


# ----------------------------------------------------------------------------
# the session
# ----------------------------------------------------------------------------
@dataclass
class SurfaceDeferredRenderSessionOptions:
    """``surface_deferred.py:26-117``: ``max_layers`` <= 0 means unlimited (capped
    at 32767); ``interpolator_impl`` is 'stencil_masked' (flat [P, C] batches per
    object) or 'full_screen' ([1, H, W, C] per object)."""
    max_layers: int = 0
    interpolator_impl: str = 'stencil_masked'
    deterministic: bool = True

    def __post_init__(self):
        if self.max_layers <= 0:
            self.max_layers = 32767


_tls = threading.local()


def _default_context():
    dev = torch.cuda.current_device()
    cache = getattr(_tls, 'ctx', None)
    if cache is None:
        cache = _tls.ctx = {}
    if dev not in cache:
        cache[dev] = nvdr.RasterizeCudaContext(dev)
    return cache[dev]


class SurfaceDeferredRenderSession:
    """Rasterise a scene once, shade every visible fragment per object, and
    compose G-buffers (``surface_deferred.py:172-804``)."""

    def __init__(self, scene: Scene, camera: Camera, opaque_only: bool = True,
                 options: Optional[SurfaceDeferredRenderSessionOptions] = None,
                 ctx: Optional[nvdr.RasterizeCudaContext] = None) -> None:
        self.scene = scene
        self.camera = camera
        self.opaque_only = opaque_only
        self.options = options if options is not None else SurfaceDeferredRenderSessionOptions()
        if opaque_only:
            self.options.max_layers = 1
        self.ctx = ctx if ctx is not None else _default_context()

    # -- camera ---------------------------------------------------------------
    @_memo
    def camera_V(self):
        return self.camera.V()

    @_memo
    def camera_P(self):
        return self.camera.P()

    @_memo
    def camera_VP(self):
        return torch.mm(self.camera_P(), self.camera_V())

    # -- geometry -------------------------------------------------------------
    @_memo
    def vertex_array_object(self) -> VertexArrayObject:
        objs = self.scene.objects
        dev = objs[0].verts.device
        keys = sorted(set().union(*(o.custom_attrs.keys() for o in objs)))
        widths = {}
        for k in keys:
            w = {o.custom_attrs[k].shape[-1] for o in objs if k in o.custom_attrs}
            if len(w) != 1:
                raise ValueError(f'custom attribute {k!r} has different widths across objects')
            widths[k] = w.pop()
        base, tris, stencil = 0, [], [torch.zeros(1, dtype=torch.int64, device=dev)]
        for i, o in enumerate(objs):
            if o.tris.shape[-1] != 3:
                raise ValueError('triangles must have 3 vertex indices')
            tris.append(o.tris + base)
            stencil.append(torch.full((len(o.tris),), i + 1, dtype=torch.int64, device=dev))
            base += len(o.verts)

        def cat(name, width=None):
            parts = []
            for o in objs:
                a = getattr(o, name)
                if a.shape[0] != o.verts.shape[0] or (width is not None and a.shape[-1] != width):
                    raise ValueError(f'vertex attribute {name!r} does not match the vertices')
                parts.append(a)
            return parts[0] if len(parts) == 1 else torch.cat(parts)

        custom = {}
        for k in keys:
            custom[k] = torch.cat([o.custom_attrs[k] if k in o.custom_attrs
                                   else o.verts.new_zeros(len(o.verts), widths[k]) for o in objs])
        world = [xform_affine(o.verts, o.M) for o in objs]
        return VertexArrayObject(
            verts=cat('verts', 3), normals=cat('normals', 3),
            world_pos=world[0] if len(world) == 1 else torch.cat(world),
            tris=(tris[0] if len(tris) == 1 else torch.cat(tris)).int().contiguous(),
            stencils=torch.cat(stencil).int().contiguous(),
            color=cat('color'), uv=cat('uv', 2), tangents=cat('tangents', 4), custom_attrs=custom)

    @_memo
    def clip_space(self) -> torch.Tensor:
        """World positions in GL clip space, [V, 4]."""
        return xform_point(self.vertex_array_object().world_pos, self.camera_VP()).contiguous()

    @_memo
    def rasterize(self) -> List[torch.Tensor]:
        """Depth layers [1, H, W, 4] = (u, v, z/w, id), nearest first; stops at the
        first empty layer when peeling."""
        vao = self.vertex_array_object()
        h, w = self.camera.resolution()
        ranges = torch.tensor([[0, len(vao.tris)]], dtype=torch.int32)
        # the emptiness test runs on the device, read back one layer late
        return layered.peel_layers(self.ctx, self.clip_space(), vao.tris, (h, w), self.options.max_layers, ranges)

    # -- materials --------------------------------------------------------------
    @_memo
    def layer_material(self) -> List[List[Tuple[SurfaceInput, SurfaceOutputStandard]]]:
        vao = self.vertex_array_object()
        V, P = self.camera_V(), self.camera_P()
        result = []
        n_obj = len(self.scene.objects)
        for rast in self.rasterize():
            if self.options.interpolator_impl == 'stencil_masked':
                order, runs = stencil_partition(rast, vao.stencils, n_obj)
            per_obj = []
            for i, obj in enumerate(self.scene.objects):
                su = SurfaceUniform(obj.M, V, P)
                if self.options.interpolator_impl == 'full_screen':
                    si = SurfaceInput(su, vao, FullScreenInterpolator(rast, vao.tris))
                elif self.options.interpolator_impl == 'stencil_masked':
                    s, e = runs[i + 1]
                    si = SurfaceInput(su, vao, MaskedSparseInterpolator(rast, vao.tris, pixels=order[s:e]))
                    if len(si.interpolator.indices[0]) == 0:
                        per_obj.append((si, SurfaceOutputStandard()))
                        continue
                else:
                    raise ValueError(f'unknown interpolator_impl {self.options.interpolator_impl!r}')
                per_obj.append((si, obj.material.shade(su, si)))
            result.append(per_obj)
        return result

    def gbuffer_collect(self, operator: Callable, default: Sequence[float]) -> List[torch.Tensor]:
        """One [1, H, W, C] buffer per layer: ``operator(si, so)`` at each object's
        fragments, ``default`` where it returns None and on empty pixels."""
        h, w = self.camera.resolution()
        dev = self.vertex_array_object().tris.device
        default = list(default)
        C = len(default)
        out = []
        if self.options.interpolator_impl == 'stencil_masked':
            fill = torch.tensor(default, dtype=torch.float32, device=dev).expand(1, h, w, C)
            for mats in self.layer_material():
                idx, vals = [], []
                for si, so in mats:
                    if len(si.interpolator.indices[0]) == 0:
                        continue
                    v = operator(si, so)
                    if v is not None:
                        idx.append(si.interpolator.indices)
                        vals.append(v)
                if not idx:
                    out.append(fill)
                elif len(idx) == 1:
                    out.append(fill.index_put(idx[0], vals[0]))
                else:
                    cat_idx = tuple(torch.cat([ix[d] for ix in idx]) for d in range(len(idx[0])))
                    out.append(fill.index_put(cat_idx, torch.cat(vals)))
            return out
        vao = self.vertex_array_object()
        for rast, mats in zip(self.rasterize(), self.layer_material()):
            buffers = [torch.tensor(default, dtype=torch.float32, device=dev).expand(1, h, w, C)]
            lookup = [0]
            for si, so in mats:
                v = operator(si, so)
                lookup.append(0 if v is None else len(buffers))
                if v is not None:
                    buffers.append(v)
            which = vao.stencils.new_tensor(lookup)[vao.stencils][float_to_triidx(rast[..., 3])]
            out.append(torch.gather(torch.cat(buffers), 0, which.long()[..., None].expand(1, h, w, C)))
        return out

    def compose_layers(self, colors: list, alphas: Optional[list] = None, blend_fn: Optional[Callable] = None,
                       return_alpha: bool = True) -> torch.Tensor:
        """Blend layers back to front; returns [H, W, C(+1)] top row first."""
        if alphas is None:
            alphas = self.alpha_layered()
        blend_fn = blend_fn or alpha_blend
        if len(colors) == 0:
            return torch.zeros(*self.camera.resolution(), 4, dtype=torch.float32, device='cuda')
        acc_c, acc_a = colors[-1], alphas[-1]
        for c, a in zip(colors[-2::-1], alphas[-2::-1]):
            acc_c, acc_a = blend_fn(acc_c, acc_a, c, a)
        img = torch.cat([acc_c, acc_a], -1) if return_alpha else acc_c
        return torch.flipud(img.squeeze(0))

    # -- G-buffers --------------------------------------------------------------
    @_memo
    def alpha_layered(self):
        def alpha(si, so):
            ones = si.interpolator.vi_data.new_ones(*si.interpolator.vi_data.shape[:-1], 1)
            if self.opaque_only or so.alpha is None:
                return ones
            return so.alpha
        return self.gbuffer_collect(alpha, [0.0])

    @_memo
    def albedo_layered(self):
        return self.gbuffer_collect(lambda si, so: so.albedo, [1.0, 0.0, 1.0])

    @_memo
    def emission_layered(self):
        return self.gbuffer_collect(lambda si, so: so.emission, [0.0, 0.0, 0.0])

    @_memo
    def mso_layered(self):
        def mso(si, so):
            if so.metallic is None and so.smoothness is None and so.occlusion is None:
                return None
            n = si.interpolator.vi_data.shape[:-1]
            dev = si.interpolator.vi_data.device

            def ch(x, d):
                return x if x is not None else torch.full((*n, 1), d, device=dev)
            return torch.cat([ch(so.metallic, 0.0), ch(so.smoothness, 0.5), ch(so.occlusion, 1.0)], -1)
        return self.gbuffer_collect(mso, [0.0, 0.5, 1.0])

    def _world_normal(self, si: SurfaceInput, so: SurfaceOutputStandard):
        """Geometry normal, or the material's normal in its declared space (``mixin.py:115-128``)."""
        if so.normal is None:
            return si.world_normal
        if so.normal_space == 'tangent':
            n = si.world_normal_unnormalized
            t, sign = si.world_tangent[..., :3], si.world_tangent[..., 3:4]
            b = sign * torch.linalg.cross(n, t)
            m = so.normal
            return normalized(torch.addcmul(torch.addcmul(m[..., 2:3] * n, m[..., 1:2], b), m[..., 0:1], t))
        if so.normal_space == 'object':
            return normalized(xform_dir(so.normal, si.uniforms.M))
        if so.normal_space == 'world':
            return normalized(so.normal)
        raise ValueError(f'unknown normal space {so.normal_space!r}')

    @_memo
    def world_space_normal_layered(self):
        return self.gbuffer_collect(self._world_normal, [0.0, 0.0, 0.0])

    @_memo
    def camera_space_normal_layered(self):
        V = self.camera.V()
        return [normalized(xform_dir(n, V)) for n in self.world_space_normal_layered()]

    @_memo
    def depth_layered(self):
        V = self.camera.V()
        return self.gbuffer_collect(lambda si, so: -xform_affine(si.world_pos, V)[..., 2:3], [0.0])

    @_memo
    def distance_layered(self):
        return self.gbuffer_collect(
            lambda si, so: torch.norm(si.world_pos - si.uniforms.camera_position, dim=-1, keepdim=True), [0.0])

    @_memo
    def world_position_layered(self):
        return self.gbuffer_collect(lambda si, so: si.world_pos, [0.0, 0.0, 0.0])

    @_memo
    def local_position_layered(self):
        return self.gbuffer_collect(lambda si, so: si.local_pos, [0.0, 0.0, 0.0])

    # -- composed images -----------------------------------------------------
    @_memo
    def albedo(self):
        return self.compose_layers(self.albedo_layered())

    @_memo
    def albedo_srgb(self):
        """Albedo with RGB in sRGB (``surface_deferred.py:457-467``)."""
        a = self.albedo()
        return torch.cat([linear_to_srgb(a[..., :3]), a[..., 3:]], -1)

    @_memo
    def emission(self):
        return self.compose_layers(self.emission_layered(), blend_fn=alpha_additive, return_alpha=False)

    @_memo
    def false_color_mask_mso(self):
        return self.compose_layers(self.mso_layered())

    @_memo
    def world_space_normal(self):
        return self.compose_layers(self.world_space_normal_layered())

    @_memo
    def camera_space_normal(self):
        return self.compose_layers(self.camera_space_normal_layered())

    @staticmethod
    def _false_color(img):
        return torch.cat([img[..., :3] * 0.5 + 0.5, img[..., 3:4]], -1)

    @_memo
    def false_color_world_space_normal(self):
        return self._false_color(self.world_space_normal())

    @_memo
    def false_color_camera_space_normal(self):
        return self._false_color(self.camera_space_normal())

    @_memo
    def local_position(self):
        return self.compose_layers(self.local_position_layered())

    @_memo
    def world_position(self):
        return self.compose_layers(self.world_position_layered())

    @_memo
    def false_color_nocs(self):
        return self._false_color(self.local_position())

    @_memo
    def depth(self):
        return self.compose_layers(self.depth_layered(), return_alpha=False)

    @_memo
    def distance(self):
        return self.compose_layers(self.distance_layered(), return_alpha=False)

    def aov(self, key: str, bg_value: Sequence[float]):
        return self.compose_layers(self.gbuffer_collect(lambda si, so: so.aovs and so.aovs.get(key), bg_value))

    # -- visibility gradients ----------------------------------------------------
    def nvdr_antialias(self, rgb: torch.Tensor) -> torch.Tensor:
        """Silhouette antialiasing of an [H, W, C] image rendered by this session
        (first layer only), with gradients into the vertex positions."""
        img = torch.flipud(rgb)[None].contiguous()
        rast = self.rasterize()[0].contiguous()
        out = nvdr.antialias(img, rast, self.clip_space(), self.vertex_array_object().tris)
        return torch.flipud(out.squeeze(0))

    def edge_gradient(self, rgb: torch.Tensor) -> torch.Tensor:
        """Identity on the image whose backward adds edge-based visibility
        gradients (``surface_deferred.py:777-804``)."""
        from .edge_grad import edge_gradient
        rast = self.rasterize()[0]
        return edge_gradient(rgb, rast, self.clip_space(), self.vertex_array_object().tris)
