"""GPU parity of interpolate / antialias / topology and of all backward passes.

Forward outputs use the oracle's float operation sequence (compared bit-exact
where the sequence is identical, else within 1e-6); gradients are float
atomics on the GPU (order-dependent sums) and are compared with the oracle's
sequential float32 gradients at 1e-4 relative to the gradient scale.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import glue
from diffrp_amd import nvdr
from tests.scenes import noisy_clip, readme_clip

pytestmark = pytest.mark.gpu


def _rast(pos, tri, res, grad_db=False):
    ctx = nvdr.RasterizeCudaContext()
    return nvdr.rasterize(ctx, pos, tri, res, grad_db=grad_db)


def _close_rel(a, b, rtol, what):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(np.abs(b).max(), 1e-30)
    err = np.abs(a - b).max() / scale
    assert err <= rtol, f'{what}: max |diff| / max |ref| = {err:.3e} > {rtol}'


@pytest.mark.parametrize('C', [3, 4, 8, 5])
def test_interpolate_forward(cuda, C):
    pos, tri, _ = noisy_clip(5, 0.01, 256)
    g = torch.Generator().manual_seed(1)
    attr = torch.rand(pos.shape[0], C, generator=g)
    print(f'case: interpolate C={C}', flush=True)
    rast, _ = _rast(pos.cuda(), tri.cuda(), (256, 256))
    out, out_da = nvdr.interpolate(attr.cuda(), rast, tri.cuda())
    assert out_da.shape[-1] == 0
    r = rast.cpu().numpy()
    o = oracle.interpolate(attr.numpy(), r, tri.numpy())
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))
    # against diffrp's own torch formula (addcmul) on CPU, 1e-5 relative
    rt = torch.from_numpy(r)
    mask = rt[..., 3] > 0
    vi = rt[mask]
    ti = tri[glue.float_to_triidx(vi[..., 3]) - 1].long()
    ref = glue.polyfill_interpolate(attr, vi, ti)
    got = out.cpu()[mask]
    assert torch.allclose(got, ref, rtol=1e-5, atol=1e-6)


def test_interpolate_batched_attr_and_da(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 128, views=3)
    g = torch.Generator().manual_seed(2)
    attr = torch.rand(3, pos.shape[1], 6, generator=g)
    rast, db = _rast(pos.cuda(), tri.cuda(), (128, 128), grad_db=True)
    out, da = nvdr.interpolate(attr.cuda(), rast, tri.cuda(), rast_db=db, diff_attrs=[0, 2, 5])
    o, oda = oracle.interpolate(attr.numpy(), rast.cpu().numpy(), tri.numpy(), rast_db=db.cpu().numpy(),
                                diff_attrs=[0, 2, 5])
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))
    np.testing.assert_array_equal(da.cpu().numpy().view(np.uint32), oda.view(np.uint32))


def test_topology_matches_oracle(cuda):
    _, tri, _ = noisy_clip(5, 0.0, 64)
    th = nvdr.antialias_construct_topology_hash(tri.cuda())
    np.testing.assert_array_equal(th.opp.cpu().numpy(), oracle.topology(tri.numpy()))
    # soup: every edge is a boundary edge
    soup = torch.arange(300, dtype=torch.int32).reshape(100, 3)
    assert (nvdr.antialias_construct_topology_hash(soup.cuda()).opp.cpu() == -1).all()


@pytest.mark.parametrize('views', [None, 2])
def test_antialias_forward(cuda, views):
    pos, tri, _ = noisy_clip(5, 0.01, 192, views=views)
    n = 1 if views is None else views
    g = torch.Generator().manual_seed(3)
    color = torch.rand(n, 192, 192, 3, generator=g)
    print(f'case: antialias views={views}', flush=True)
    rast, _ = _rast(pos.cuda(), tri.cuda(), (192, 192))
    out = nvdr.antialias(color.cuda(), rast, pos.cuda(), tri.cuda())
    o = oracle.antialias(color.numpy(), rast.cpu().numpy(), pos.numpy(), tri.numpy())
    np.testing.assert_array_equal(out.cpu().numpy().view(np.uint32), o.view(np.uint32))
    assert (out.cpu() != color).any()


def test_rasterize_backward(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 128)
    g = torch.Generator().manual_seed(4)
    pos_d = pos.cuda().requires_grad_(True)
    rast, _ = _rast(pos_d, tri.cuda(), (128, 128))
    gr = torch.randn(rast.shape, generator=g)
    gr[..., 3] = 0
    rast.backward(gr.cuda())
    o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), rast.detach().cpu().numpy(), gr.numpy())
    _close_rel(pos_d.grad.cpu().numpy(), o, 1e-4, 'grad pos (rasterize)')


def test_rasterize_backward_db(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 96)
    g = torch.Generator().manual_seed(5)
    pos_d = pos.cuda().requires_grad_(True)
    rast, db = _rast(pos_d, tri.cuda(), (96, 96), grad_db=True)
    gr = torch.randn(rast.shape, generator=g)
    gr[..., 3] = 0
    gd = torch.randn(db.shape, generator=g)
    torch.autograd.backward([rast, db], [gr.cuda(), gd.cuda()])
    o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), rast.detach().cpu().numpy(), gr.numpy(), grad_db=gd.numpy())
    _close_rel(pos_d.grad.cpu().numpy(), o, 1e-4, 'grad pos (rasterize + db)')


def test_interpolate_backward(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 128)
    g = torch.Generator().manual_seed(6)
    attr = torch.rand(pos.shape[0], 8, generator=g)
    rast, db = _rast(pos.cuda(), tri.cuda(), (128, 128), grad_db=True)
    rast = rast.detach().requires_grad_(True)
    db = db.detach().requires_grad_(True)
    a = attr.cuda().requires_grad_(True)
    out, da = nvdr.interpolate(a, rast, tri.cuda(), rast_db=db, diff_attrs=[1, 7])
    go = torch.randn(out.shape, generator=g)
    gda = torch.randn(da.shape, generator=g)
    torch.autograd.backward([out, da], [go.cuda(), gda.cuda()])
    oa, orr, odb = oracle.interpolate_bwd(attr.numpy(), rast.detach().cpu().numpy(), tri.numpy(), go.numpy(),
                                          rast_db=db.detach().cpu().numpy(), diff_attrs=[1, 7], grad_da=gda.numpy())
    _close_rel(a.grad.cpu().numpy(), oa, 1e-4, 'grad attr')
    _close_rel(rast.grad.cpu().numpy(), orr, 1e-5, 'grad rast')
    _close_rel(db.grad.cpu().numpy(), odb, 1e-5, 'grad db')


def test_antialias_backward(cuda):
    pos, tri, _ = noisy_clip(5, 0.01, 160)
    g = torch.Generator().manual_seed(7)
    color = torch.rand(1, 160, 160, 4, generator=g)
    rast, _ = _rast(pos.cuda(), tri.cuda(), (160, 160))
    c = color.cuda().requires_grad_(True)
    p = pos.cuda().requires_grad_(True)
    out = nvdr.antialias(c, rast, p, tri.cuda(), pos_gradient_boost=2.0)
    go = torch.randn(out.shape, generator=g)
    out.backward(go.cuda())
    oc, op = oracle.antialias_bwd(color.numpy(), rast.cpu().numpy(), pos.numpy(), tri.numpy(), go.numpy(),
                                  pos_gradient_boost=2.0)
    _close_rel(c.grad.cpu().numpy(), oc, 1e-5, 'grad color')
    _close_rel(p.grad.cpu().numpy(), op, 1e-4, 'grad pos (antialias)')
    assert np.abs(op).max() > 0


def test_readme_pipeline_colors_match_reference(cuda):
    """GPU raster + diffrp's shading glue reproduce the reference's 512^2 README output."""
    import os
    fx = np.load(os.path.join(os.path.dirname(__file__), 'golden', 'readme_normals_512.npz'))['rgba_bottom_first']
    pos, tri, verts, cam = readme_clip(512)
    sv, st, sn = glue.face_soup_flat(verts, tri)
    clip = torch.nn.functional.pad(sv, (0, 1), value=1.0) @ (cam.P('cpu') @ cam.V('cpu')).T
    rast, _ = _rast(clip.cuda(), st.cuda(), (512, 512))
    img = glue.false_color_camera_space_normal(rast.cpu(), sv, st, sn, torch.eye(4), cam.V('cpu'))
    ours = glue.to_uint8(img)[::-1]
    np.testing.assert_array_equal(ours, fx)


def test_incidence_table(cuda):
    _, tri, _ = noisy_clip(4, 0.0, 64)
    V = int(tri.max()) + 3          # two vertices without corners
    off, inc = nvdr.incidence(tri.cuda(), V)
    off, inc = off.cpu().numpy(), inc.cpu().numpy()
    flat = tri.numpy().reshape(-1)
    counts = np.bincount(flat, minlength=V)
    np.testing.assert_array_equal(np.diff(off), counts)
    for v in range(0, V, 97):
        np.testing.assert_array_equal(inc[off[v]:off[v + 1]], np.nonzero(flat == v)[0])


@pytest.mark.parametrize('C', [3, 8, 16])
def test_interpolate_backward_triangle_centric(cuda, C):
    """Backward with the rasteriser's pixel boxes (no atomics) against the oracle."""
    pos, tri, _ = noisy_clip(4, 0.01, 96, views=5)
    g = torch.Generator().manual_seed(8)
    attr = torch.rand(pos.shape[1], C, generator=g)
    ctx = nvdr.RasterizeCudaContext()
    rast, _ = nvdr.rasterize(ctx, pos.cuda(), tri.cuda(), (96, 96), grad_db=False)
    assert hasattr(rast, '_drp_boxes')
    a = attr.cuda().requires_grad_(True)
    out, _ = nvdr.interpolate(a, rast, tri.cuda())
    go = torch.randn(out.shape, generator=g)
    out.backward(go.cuda())
    ga1 = a.grad.clone()
    oa, orr = oracle.interpolate_bwd(attr.numpy(), rast.detach().cpu().numpy(), tri.numpy(), go.numpy())
    _close_rel(ga1.cpu().numpy(), oa, 1e-5, 'grad attr (triangle-centric)')
    # deterministic: a second backward gives identical bits
    a.grad = None
    out, _ = nvdr.interpolate(a, rast, tri.cuda())
    out.backward(go.cuda())
    assert torch.equal(a.grad, ga1)


def test_interpolate_backward_batched_attr_triangle_centric(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 80, views=3)
    g = torch.Generator().manual_seed(9)
    attr = torch.rand(3, pos.shape[1], 4, generator=g)
    rast, _ = _rast(pos.cuda(), tri.cuda(), (80, 80))
    a = attr.cuda().requires_grad_(True)
    out, _ = nvdr.interpolate(a, rast, tri.cuda())
    go = torch.randn(out.shape, generator=g)
    out.backward(go.cuda())
    oa, _ = oracle.interpolate_bwd(attr.numpy(), rast.cpu().numpy(), tri.numpy(), go.numpy())
    _close_rel(a.grad.cpu().numpy(), oa, 1e-5, 'grad attr batched')


def test_rasterize_backward_instanced_triangle_centric(cuda):
    pos, tri, _ = noisy_clip(4, 0.01, 112, views=4)
    g = torch.Generator().manual_seed(10)
    p = pos.cuda().requires_grad_(True)
    rast, _ = _rast(p, tri.cuda(), (112, 112))
    gr = torch.randn(rast.shape, generator=g)
    gr[..., 3] = 0
    rast.backward(gr.cuda())
    o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), rast.detach().cpu().numpy(), gr.numpy())
    _close_rel(p.grad.cpu().numpy(), o, 1e-4, 'grad pos instanced')
    g1 = p.grad.clone()
    p.grad = None
    rast, _ = _rast(p, tri.cuda(), (112, 112))
    rast.backward(gr.cuda())
    assert torch.equal(p.grad, g1)


def test_rasterize_backward_range_mode_uses_atomics(cuda):
    pos, tri, _ = noisy_clip(3, 0.01, 64)
    g = torch.Generator().manual_seed(11)
    p = pos.cuda().requires_grad_(True)
    ranges = torch.tensor([[0, 600], [400, 880]], dtype=torch.int32)
    rast, _ = nvdr.rasterize(nvdr.RasterizeCudaContext(), p, tri.cuda(), (64, 64), ranges=ranges, grad_db=False)
    gr = torch.randn(rast.shape, generator=g)
    gr[..., 3] = 0
    rast.backward(gr.cuda())
    o = oracle.rasterize_bwd(pos.numpy(), tri.numpy(), rast.detach().cpu().numpy(), gr.numpy())
    _close_rel(p.grad.cpu().numpy(), o, 1e-4, 'grad pos range mode')


def test_antialias_backward_with_and_without_forward_state(cuda):
    """The backward that reads the forward's pair decisions equals the one that searches again."""
    from diffrp_amd import _lib
    from diffrp_amd._lib import ptr, stream_of
    pos, tri, _ = noisy_clip(5, 0.01, 128, views=2)
    g = torch.Generator().manual_seed(12)
    n, H, W, C = 2, 128, 128, 4
    color = torch.rand(n, H, W, C, generator=g).cuda()
    go = torch.randn(n, H, W, C, generator=g).cuda()
    p, t = pos.cuda().contiguous(), tri.cuda().int().contiguous()
    rast, _ = _rast(p, t, (H, W))
    opp = nvdr.antialias_construct_topology_hash(t).opp
    L = _lib.lib()
    V, F = p.shape[1], t.shape[0]
    ws_bytes = L.drp_antialias_workspace_size(n, H, W, V, F, 1)
    out = torch.empty_like(color)
    res = []
    for reuse in (0, 1):
        ws = torch.empty(ws_bytes, dtype=torch.uint8, device='cuda')
        s = stream_of(color)
        _lib.check(L.drp_antialias_fwd(s, ptr(color), n, H, W, C, ptr(rast), ptr(p), V, 1, ptr(t), F, ptr(opp),
                                       ptr(out), ptr(ws), ws_bytes), 'antialias')
        gc, gp = torch.empty_like(color), torch.zeros_like(p)
        _lib.check(L.drp_antialias_bwd(s, ptr(color), n, H, W, C, ptr(rast), ptr(p), V, 1, ptr(t), F, ptr(opp),
                                       ptr(go), ptr(gc), ptr(gp), 1.0, reuse, ptr(ws), ws_bytes), 'antialias bwd')
        torch.cuda.synchronize()
        res.append((gc.cpu(), gp.cpu()))
    assert torch.equal(res[0][0], res[1][0])
    assert res[0][1].abs().max() > 0
    _close_rel(res[1][1].numpy(), res[0][1].numpy(), 1e-6, 'grad pos with forward state')


def test_backward_with_triangle_ids_beyond_24_bits(cuda):
    """2^24 culled triangles (over 1024 extra vertices behind the eye) in front of
    the mesh: the real triangles' ids exceed 2^24 (the id codec's integer-bits
    range) and the deterministic backward sorts 64-bit keys.  Forward and
    gradients equal the unpadded mesh's bit for bit."""
    pos, tri, _ = noisy_clip(4, 0.01, 96, views=2)
    n_pad, n_dummy = 1 << 24, 1024
    V = pos.shape[1]
    dummy = torch.tensor([0.0, 0.0, 0.0, -1.0]).expand(2, n_dummy, 4)
    pos_p = torch.cat([pos, dummy], 1).contiguous()
    d = V + torch.arange(n_pad, dtype=torch.int32) % n_dummy
    tri_p = torch.cat([torch.stack([d, d, d], 1), tri.int()]).contiguous()
    g = torch.Generator().manual_seed(13)
    attr = torch.rand(V, 4, generator=g)
    go = torch.randn(2, 96, 96, 4, generator=g).cuda()

    def run(p_all, t_all):
        p = p_all.cuda().requires_grad_(True)
        a = torch.cat([attr, torch.zeros(p_all.shape[1] - V, 4)]).cuda().requires_grad_(True)
        rast, _ = _rast(p, t_all.cuda(), (96, 96))
        out, _ = nvdr.interpolate(a, rast, t_all.cuda())
        (out * go).sum().backward()
        return rast.detach(), p.grad[:, :V].cpu(), a.grad[:V].cpu()

    r0, gp0, ga0 = run(pos, tri)
    r1, gp1, ga1 = run(pos_p, tri_p)
    id0 = oracle.float_to_triidx(r0[..., 3].cpu().numpy())
    id1 = oracle.float_to_triidx(r1[..., 3].cpu().numpy())
    assert np.array_equal(id1, np.where(id0 > 0, id0 + n_pad, 0))
    assert (id1 > (1 << 24)).any()
    assert torch.equal(r0[..., :3], r1[..., :3])
    assert torch.equal(gp0, gp1) and gp0.abs().max() > 0
    assert torch.equal(ga0, ga1) and ga0.abs().max() > 0
