"""The drop-in as diffrp uses it, on the GPU: nvdiffrast's module name and call
sequence, multi-object scenes through the session, nvdr_antialias, and a
second backward through one antialias."""
import math
import sys

import numpy as np
import pytest
import torch

import oracle
from oracle import glue
from diffrp_amd import deferred as dd
from diffrp_amd import nvdr
from diffrp_amd.camera import PerspectiveCamera
from diffrp_amd.geometry import icosphere, noisy_icosphere

pytestmark = pytest.mark.gpu


def _u32(t):
    return t.detach().cpu().numpy().view(np.uint32)


# ---------------------------------------------------------------------------
# import nvdiffrast.torch as dr, then diffrp's exact sequence
# ---------------------------------------------------------------------------
@pytest.fixture
def dr():
    saved = {k: sys.modules.get(k) for k in ('nvdiffrast', 'nvdiffrast.torch')}
    nvdr.install_as_nvdiffrast()
    import nvdiffrast.torch as mod
    yield mod
    for k, v in saved.items():
        if v is None:
            sys.modules.pop(k, None)
        else:
            sys.modules[k] = v


def test_install_as_nvdiffrast_runs_diffrps_call_sequence(cuda, dr):
    assert dr is nvdr
    verts, tris = noisy_icosphere(3, 0.8, 0.01, seed=4)
    h, w = 80, 96
    cam = PerspectiveCamera(h=h, w=w)
    VP = cam.P('cpu') @ cam.V('cpu')
    clip = torch.nn.functional.pad(verts, (0, 1), value=1.0) @ VP.T
    pos = clip.cuda().contiguous().requires_grad_(True)
    tri = tris.cuda().int().contiguous()
    attr = torch.rand(len(verts), 4, generator=torch.Generator().manual_seed(5)).cuda().requires_grad_(True)
    # surface_deferred.py:219-226: one context per (thread, device), no arguments
    ctx = dr.RasterizeCudaContext()
    # surface_deferred.py:259-265: CPU int32 ranges [[0, F]], default grad_db, a layer loop
    rng = torch.tensor([[0, len(tri)]], dtype=torch.int32)
    layers, dbs = [], []
    with dr.DepthPeeler(ctx, pos, tri, (h, w), rng) as dp:
        for i in range(4):
            rast, rast_db = dp.rasterize_next_layer()
            if i < 3 and (rast[..., 3] <= 0).all():
                break
            layers.append(rast)
            dbs.append(rast_db)
    assert len(layers) >= 2
    # the layers and derivatives equal the oracle's peeling, bit for bit
    prev = None
    for rast, db in zip(layers, dbs):
        r_o, db_o = oracle.rasterize(clip.numpy(), tris.numpy(), (h, w), prev=prev, with_db=True)
        np.testing.assert_array_equal(_u32(rast), r_o.view(np.uint32))
        np.testing.assert_array_equal(_u32(db), db_o.view(np.uint32))
        prev = r_o
    # interpolator.py:83: dr.interpolate(...)[0]; surface_deferred.py:775: dr.antialias
    col = dr.interpolate(attr, layers[0], tri)[0]
    out = dr.antialias(col, layers[0], pos, tri)
    # the same through the module's direct names
    r_d, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), pos, tri, (h, w), grad_db=False)
    c_d = nvdr.interpolate(attr, r_d, tri)[0]
    o_d = nvdr.antialias(c_d, r_d, pos, tri)
    np.testing.assert_array_equal(_u32(out), _u32(o_d))
    out.square().sum().backward()
    g_pos, g_attr = pos.grad.clone(), attr.grad.clone()
    pos.grad = attr.grad = None
    o_d.square().sum().backward()
    torch.testing.assert_close(g_pos, pos.grad, rtol=1e-5, atol=1e-6 * float(pos.grad.abs().max()))
    torch.testing.assert_close(g_attr, attr.grad, rtol=1e-5, atol=1e-6 * float(attr.grad.abs().max()))
    with pytest.raises(RuntimeError):
        dr.RasterizeGLContext()


# ---------------------------------------------------------------------------
# nvdr_antialias (surface_deferred.py:758-775)
# ---------------------------------------------------------------------------
def test_nvdr_antialias_equals_oracle_on_flipped_image(cuda):
    verts, tris = noisy_icosphere(3, 0.7, 0.01, seed=6)
    scene = dd.Scene().add_mesh_object(dd.MeshObject(dd.DefaultMaterial(), verts.cuda(), tris.cuda(),
                                                     normals='smooth',
                                                     color=torch.rand(len(verts), 4, device='cuda')))
    rp = dd.SurfaceDeferredRenderSession(scene, PerspectiveCamera(h=72, w=88))
    rgb = rp.albedo()[..., :3]
    got = rp.nvdr_antialias(rgb)
    clip = rp.clip_space().cpu().numpy()
    tri = rp.vertex_array_object().tris.cpu().numpy()
    r0 = oracle.rasterize(clip, tri, (72, 88))
    np.testing.assert_array_equal(_u32(rp.rasterize()[0]), r0.view(np.uint32))
    img = np.ascontiguousarray(np.flipud(rgb.cpu().numpy())[None])      # top row first -> bottom first
    want = np.flipud(oracle.antialias(img, r0, clip, tri)[0])
    np.testing.assert_array_equal(got.cpu().numpy().view(np.uint32), np.ascontiguousarray(want).view(np.uint32))
    assert (got != rgb).any()


# ---------------------------------------------------------------------------
# a second backward through one antialias (retain_graph)
# ---------------------------------------------------------------------------
def test_antialias_backward_twice_with_retained_graph(cuda):
    verts, tris = noisy_icosphere(3, 0.8, 0.01, seed=7)
    cam = PerspectiveCamera(h=64, w=64)
    clip = torch.nn.functional.pad(verts, (0, 1), value=1.0) @ (cam.P('cpu') @ cam.V('cpu')).T
    pos = clip.cuda().contiguous().requires_grad_(True)
    tri = tris.cuda()
    ctx = nvdr.RasterizeCudaContext()
    rast, _ = nvdr.rasterize(ctx, pos, tri, (64, 64), grad_db=False)
    col = torch.rand(1, 64, 64, 3, device='cuda', requires_grad=True)
    loss = nvdr.antialias(col, rast, pos, tri).square().sum()
    g1 = torch.autograd.grad(loss, (col, pos), retain_graph=True)
    g2 = torch.autograd.grad(loss, (col, pos), retain_graph=True)   # decisions recomputed
    loss.backward()
    # colour gradients are summed in a fixed order; position gradients are
    # flushed with one float atomic per vertex component and tile
    assert torch.equal(g1[0], g2[0]) and torch.equal(col.grad, g1[0])
    scale = float(g1[1].abs().max())
    assert scale > 0
    for other in (g2[1], pos.grad):
        assert float((other - g1[1]).abs().max()) <= 1e-6 * scale


# ---------------------------------------------------------------------------
# a scene of three objects (mixin.py:74-113, surface_deferred.py:268-346, mixin.py:130-155)
# ---------------------------------------------------------------------------
def _rot(axis, deg):
    c, s = math.cos(math.radians(deg)), math.sin(math.radians(deg))
    m = torch.eye(4)
    i, j = {'x': (1, 2), 'y': (2, 0), 'z': (0, 1)}[axis]
    m[i, i], m[i, j], m[j, i], m[j, j] = c, -s, s, c
    return m


def _trans(x, y, z):
    m = torch.eye(4)
    m[:3, 3] = torch.tensor([x, y, z])
    return m


class _Procedural(dd.SurfaceMaterial):
    def shade(self, su, si):
        p = si.world_pos
        return dd.SurfaceOutputStandard(albedo=torch.cat([torch.sin(3 * p[..., 0:1]) * 0.5 + 0.5,
                                                          torch.cos(2 * p[..., 1:2]) * 0.5 + 0.5,
                                                          p[..., 2:3] * 0.3 + 0.5], -1))


def _procedural_cpu(p):
    return torch.cat([torch.sin(3 * p[..., 0:1]) * 0.5 + 0.5, torch.cos(2 * p[..., 1:2]) * 0.5 + 0.5,
                      p[..., 2:3] * 0.3 + 0.5], -1)


def _three_objects():
    g = torch.Generator().manual_seed(8)
    v0, f0 = icosphere(2, 0.45)
    v1, f1 = icosphere(3, 0.4)
    quad_v = torch.tensor([[-0.9, -0.7, 0.0], [0.9, -0.7, 0.0], [0.9, 0.7, 0.0], [-0.9, 0.7, 0.0]])
    quad_f = torch.tensor([[0, 1, 2], [0, 2, 3]], dtype=torch.int32)
    tint = torch.tensor([0.5, 1.0, 0.25], device='cuda')
    objs = [
        dd.MeshObject(dd.DefaultMaterial(), torch.tensor(v0, dtype=torch.float32, device='cuda'),
                      torch.tensor(f0, dtype=torch.int32, device='cuda'), normals='flat',
                      M=_trans(-0.5, -0.1, 0.1).cuda(), color=torch.rand(len(v0), 4, generator=g).cuda()),
        dd.MeshObject(_Procedural(), torch.tensor(v1, dtype=torch.float32, device='cuda'),
                      torch.tensor(f1, dtype=torch.int32, device='cuda'), normals='smooth',
                      M=(_trans(0.45, 0.2, 0.3) @ _rot('y', 35.0)).cuda()),
        dd.MeshObject(dd.DefaultMaterial(tint), quad_v.cuda(), quad_f.cuda(),
                      normals=torch.tensor([[0.0, 0.0, 1.0]] * 4, device='cuda'),
                      M=(_trans(0.0, 0.0, -0.6) @ _rot('x', 20.0)).cuda(),
                      color=torch.rand(4, 4, generator=g).cuda()),
    ]
    scene = dd.Scene()
    for o in objs:
        scene.add_mesh_object(o)
    return scene, tint.cpu()


@pytest.mark.parametrize('impl', ['stencil_masked', 'full_screen'])
def test_three_object_scene_matches_multi_object_restatement(cuda, impl):
    scene, tint = _three_objects()
    H, W = 96, 128
    opts = dd.SurfaceDeferredRenderSessionOptions(interpolator_impl=impl)
    rp = dd.SurfaceDeferredRenderSession(scene, PerspectiveCamera(h=H, w=W), options=opts)
    cpu_objs = [{'verts': o.verts.cpu(), 'tris': o.tris.cpu(), 'M': o.M.cpu(), 'normals': o.normals.cpu(),
                 'color': o.color.cpu()} for o in scene.objects]
    ref_vao = glue.vertex_array_object(cpu_objs)
    vao = rp.vertex_array_object()
    assert torch.equal(vao.tris.cpu(), ref_vao['tris'])
    assert torch.equal(vao.stencils.cpu(), ref_vao['stencils'])
    torch.testing.assert_close(vao.world_pos.cpu(), ref_vao['world_pos'], rtol=1e-6, atol=1e-6)
    # rasterised layer: bit-exact with the oracle on the session's clip-space positions
    clip = rp.clip_space().cpu().numpy()
    rast_o = oracle.rasterize(clip, ref_vao['tris'].numpy(), (H, W))
    rast = rp.rasterize()[0]
    np.testing.assert_array_equal(_u32(rast), rast_o.view(np.uint32))
    owners = set(ref_vao['stencils'][glue.float_to_triidx(torch.from_numpy(rast_o[..., 3])).long()].unique().tolist())
    assert owners == {0, 1, 2, 3}, owners  # every object and the background are visible
    # G-buffers of the layer against the restated collect, one material per object
    r = torch.from_numpy(rast_o)
    tris, st = ref_vao['tris'], ref_vao['stencils']

    def albedo(pi, it):
        if pi == 1:
            return _procedural_cpu(it.interpolate(ref_vao['world_pos']))
        rgb = it.interpolate(ref_vao['color'])[..., :3]
        return rgb * tint if pi == 2 else rgb

    def world_normal(pi, it):
        M = cpu_objs[pi]['M']
        return glue.normalized(it.interpolate(glue.normalized(glue.transform_vector3x3(ref_vao['normals'], M))))

    def alpha(pi, it):
        return torch.ones_like(it.vi_data[..., :1])

    want_alb = glue.gbuffer_collect_layer(r, tris, st, 3, albedo, [1.0, 0.0, 1.0])
    want_nrm = glue.gbuffer_collect_layer(r, tris, st, 3, world_normal, [0.0, 0.0, 0.0])
    want_alpha = glue.gbuffer_collect_layer(r, tris, st, 3, alpha, [0.0])
    torch.testing.assert_close(rp.albedo_layered()[0].cpu(), want_alb, rtol=1e-5, atol=1e-5)
    torch.testing.assert_close(rp.world_space_normal_layered()[0].cpu(), want_nrm, rtol=1e-5, atol=1e-5)
    assert torch.equal(rp.alpha_layered()[0].cpu(), want_alpha)
    torch.testing.assert_close(rp.albedo().cpu(), glue.compose_layers([want_alb], [want_alpha]),
                               rtol=1e-5, atol=1e-5)
