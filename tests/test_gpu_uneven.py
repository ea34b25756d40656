"""The packed uneven batch (UnevenBatch): rasterize, interpolate and antialias
and their backward passes over every item in one launch sequence per operator.

Parity: on a small draw every item's rast, attributes and antialiased colours
equal the float32 oracle's bit for bit (ids converted from batch to
item-local), and the position and attribute gradients of an L2 loss agree with
the oracle's float32 backward chain per item to 1e-4 of the item's gradient
scale; the whole step captured in a HIP graph replays to the eager result; at
the full 32-mesh draw the batch equals per-item nvdr calls (forward bit for
bit, gradients to 1e-5 of scale)."""
import numpy as np
import pytest
import torch

import oracle
from diffrp_amd import batch, nvdr
from diffrp_amd.loss import l2_loss
from tests.test_gpu_batch import SMALL_RES, SMALL_S, _scene

pytestmark = pytest.mark.gpu


def _local_ids(rast_item, f0):
    """Batch ids of one item's slice -> item-local ids (same codec)."""
    r = rast_item.clone()
    ids = oracle.float_to_triidx(r[..., 3].cpu().numpy())
    loc = np.where(ids > 0, ids - f0, 0).astype(np.int32)
    r[..., 3] = torch.from_numpy(oracle.triidx_to_float(loc)).to(r.device)
    return r


def _build(pos, tri, res, C, seed):
    g = torch.Generator().manual_seed(seed)
    ctx = nvdr.RasterizeCudaContext()
    ub = batch.UnevenBatch(ctx, [t.cuda() for t in tri], [p.shape[0] for p in pos], res)
    P = torch.cat(pos).cuda().contiguous().requires_grad_(True)
    A = torch.rand(P.shape[0], C, generator=g).cuda().requires_grad_(True)
    T = torch.rand(ub.n_pix, C, generator=g).cuda()
    return ub, P, A, T


def _step(ub, P, A, T):
    rast, boxes = ub.rasterize(P)
    col = ub.interpolate(A, rast, boxes)
    out = ub.antialias(col, rast, P)
    loss = l2_loss(out, T)
    loss.backward()
    return rast, col, out, loss


def test_uneven_batch_matches_oracle_per_item(cuda):
    print('case: 6 meshes s=2..5, resolutions 48x80..256x256, C=4', flush=True)
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 11)
    ub, P, A, T = _build(pos, tri, SMALL_RES, 4, 12)
    rast, col, out, _ = _step(ub, P, A, T)
    torch.cuda.synchronize()
    rs, cs, os_ = ub.split(rast.detach()), ub.split(col.detach()), ub.split(out.detach())
    gp, ga = P.grad.cpu().numpy(), A.grad.cpu().numpy()
    Ts = ub.split(T)
    for i, (p, t, res) in enumerate(zip(pos, tri, SMALL_RES)):
        vb, V, fb, F, pb, H, W = ub.plan.item(i)
        p_np, t_np = p.numpy(), t.numpy()
        r_o = oracle.rasterize(p_np, t_np, res)
        np.testing.assert_array_equal(_local_ids(rs[i], fb).cpu().numpy().view(np.uint32), r_o.view(np.uint32))
        a_i = A.detach()[vb:vb + V].cpu().numpy()
        c_o = oracle.interpolate(a_i, r_o, t_np)
        np.testing.assert_array_equal(cs[i].cpu().numpy().view(np.uint32), c_o.view(np.uint32))
        opp = oracle.topology(t_np)
        y_o = oracle.antialias(c_o, r_o, p_np, t_np, opp)
        np.testing.assert_array_equal(os_[i].cpu().numpy().view(np.uint32), y_o.view(np.uint32))
        gy = 2.0 * (y_o - Ts[i].cpu().numpy())
        gcol, gpa = oracle.antialias_bwd(c_o, r_o, p_np, t_np, gy, opp)
        gat, grast = oracle.interpolate_bwd(a_i, r_o, t_np, gcol)
        gpo = oracle.rasterize_bwd(p_np, t_np, r_o, grast) + gpa
        for got, want, what in ((gp[vb:vb + V], gpo, 'pos'), (ga[vb:vb + V], gat, 'attr')):
            scale = np.abs(want).max()
            assert scale > 0, (i, what)
            err = np.abs(got - want).max() / scale
            assert err <= 1e-4, f'item {i} {what}: {err:.2e}'


def test_uneven_batch_step_replays_from_a_graph(cuda):
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 13)
    ub, P, A, T = _build(pos, tri, SMALL_RES, 8, 14)

    def step():
        P.grad = None
        A.grad = None
        return _step(ub, P, A, T)[3].detach()
    l_ref = step()
    torch.cuda.synchronize()
    gp_ref, ga_ref = P.grad.clone(), A.grad.clone()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(side)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        l_g = step()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(l_g, l_ref)
        assert torch.equal(A.grad, ga_ref)                       # fixed summation order
        scale = float(gp_ref.abs().max())
        assert float((P.grad - gp_ref).abs().max()) <= 1e-5 * scale  # antialias: per-tile float atomics
        torch.rand(1 << 20, device='cuda').sum().item()          # unrelated eager work between replays


def _c4():
    rng = np.random.default_rng(3)
    subdivs = rng.integers(3, 9, 32)
    sizes = rng.choice([256, 512, 1024, 2048], 32)
    res = [(int(s), int(s)) for s in sizes]
    pos, tri = _scene(subdivs, res, 0.02, 3)
    return pos, tri, res


def test_uneven_batch_full_size_equals_per_item_calls(cuda):
    """The 32-mesh draw of seed 3 (subdivision 3-8, 256^2-2048^2, C=4), tuned."""
    pos, tri, res = _c4()
    print(f'case: 32 meshes, {sum(t.shape[0] for t in tri)} faces, {sum(h * w for h, w in res)} pixels', flush=True)
    ctx = nvdr.RasterizeCudaContext()
    batch.tune_batch(ctx, [p.cuda() for p in pos], [t.cuda() for t in tri], res)
    g = torch.Generator().manual_seed(15)
    ub = batch.UnevenBatch(ctx, [t.cuda() for t in tri], [p.shape[0] for p in pos], res)
    P = torch.cat(pos).cuda().contiguous().requires_grad_(True)
    A = torch.rand(P.shape[0], 4, generator=g).cuda().requires_grad_(True)
    T = torch.rand(ub.n_pix, 4, generator=g).cuda()
    rast, col, out, loss = _step(ub, P, A, T)
    gp_b, ga_b = P.grad.clone(), A.grad.clone()
    rs, os_, Ts = ub.split(rast.detach()), ub.split(out.detach()), ub.split(T)
    # per item through the drop-in operators
    loss_items = 0.0
    gp_i, ga_i = torch.zeros_like(P), torch.zeros_like(A)
    for i in range(len(res)):
        vb, V, fb, F, pb, H, W = ub.plan.item(i)
        p = P.detach()[vb:vb + V].clone().requires_grad_(True)
        a = A.detach()[vb:vb + V].clone().requires_grad_(True)
        t = tri[i].cuda()
        r, _ = nvdr.rasterize(ctx, p, t, res[i], grad_db=False)
        assert torch.equal(_local_ids(rs[i], fb).view(torch.int32), r.view(torch.int32)), f'item {i} rast'
        y = nvdr.antialias(nvdr.interpolate(a, r, t)[0], r, p, t)
        assert torch.equal(os_[i].view(torch.int32), y.view(torch.int32)), f'item {i} antialias'
        li = l2_loss(y, Ts[i])
        li.backward()
        loss_items += float(li)
        gp_i[vb:vb + V] = p.grad
        ga_i[vb:vb + V] = a.grad
    assert abs(float(loss) - loss_items) <= 1e-5 * loss_items
    for i in range(len(res)):
        vb, V = ub.plan.item(i)[:2]
        for got, want, what in ((gp_b, gp_i, 'pos'), (ga_b, ga_i, 'attr')):
            w = want[vb:vb + V]
            scale = float(w.abs().max())
            err = float((got[vb:vb + V] - w).abs().max())
            assert err <= 1e-5 * max(scale, 1e-30), f'item {i} {what}: {err:.3e} of {scale:.3e}'


@pytest.mark.parametrize('C', [4, 8])
def test_uneven_fused_backward_equals_separate_passes(cuda, C):
    """The batch's attribute and position backward from one pass over the
    pixels (drp_batch_interpolate_rasterize_bwd) equals the two batch passes,
    bit for bit, with antialias and its position gradient on top."""
    print(f'case: 6 meshes s=2..5, C={C}', flush=True)
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 17)
    grads = []
    for fused in (True, False):
        saved = batch._FUSED_C
        batch._FUSED_C = saved if fused else ()
        try:
            ub, P, A, T = _build(pos, tri, SMALL_RES, C, 18)
            _step(ub, P, A, T)
            torch.cuda.synchronize()
            grads.append((P.grad.clone(), A.grad.clone()))
        finally:
            batch._FUSED_C = saved
    (gp_f, ga_f), (gp_s, ga_s) = grads
    assert torch.equal(ga_f, ga_s)
    # the antialias position gradient is summed with float atomics; the rest is bit-exact
    assert float((gp_f - gp_s).abs().max()) <= 1e-5 * float(gp_s.abs().max())
    assert float(gp_f.abs().max()) > 0


def test_uneven_fused_backward_without_antialias_is_bit_exact(cuda):
    pos, tri = _scene(SMALL_S, SMALL_RES, 0.02, 19)
    grads = []
    for fused in (True, False):
        saved = batch._FUSED_C
        batch._FUSED_C = saved if fused else ()
        try:
            ub, P, A, T = _build(pos, tri, SMALL_RES, 8, 20)
            rast, boxes = ub.rasterize(P)
            col = ub.interpolate(A, rast, boxes)
            l2_loss(col, T).backward()
            torch.cuda.synchronize()
            grads.append((P.grad.clone(), A.grad.clone()))
        finally:
            batch._FUSED_C = saved
    (gp_f, ga_f), (gp_s, ga_s) = grads
    assert torch.equal(ga_f, ga_s) and torch.equal(gp_f, gp_s)
    assert float(gp_f.abs().max()) > 0
