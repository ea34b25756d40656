"""Pins of the oracle's backward passes by float64 central differences (CPU).

Each operator's analytic gradient (float64 oracle) is compared with central
differences of its own forward at frozen visibility, and the float32 oracle
gradients with the float64 ones.  A wrong sign, index or missing term in a
backward formula fails these checks.
"""
import numpy as np
import pytest

import oracle
from tests.fd import central_diff, fd_check, ids_f64, pipeline_f64, pipeline_grad_f64
from tests.scenes import noisy_clip, vp_from_direction

R = np.random.default_rng


@pytest.fixture(scope='module', autouse=True)
def _built(oracle_lib):
    return oracle_lib


def _scene(res=48, views=2, subdiv=3):
    pos, tri, verts = noisy_clip(subdiv, 0.01, res, views=views)
    return pos.double().numpy(), tri.numpy(), verts


def test_rasterize_backward_fd_with_db():
    pos, tri, _ = _scene()
    H = W = 48
    r0 = oracle.rasterize(pos, tri, (H, W), dtype=np.float64)
    rng = R(0)
    wr = rng.standard_normal(r0.shape)
    wr[..., 3] = 0
    wd = rng.standard_normal(r0.shape)
    x = pos.copy()

    def f():
        r, db = oracle.rasterize_fixed(x, tri, r0, with_db=True, dtype=np.float64)
        return float((r * wr).sum() + (db * wd).sum())

    g = oracle.rasterize_bwd(pos, tri, r0, wr, grad_db=wd, dtype=np.float64)
    nz = np.flatnonzero(np.abs(g) > 1e-6 * np.abs(g).max())
    coords = R(1).choice(nz, 60, replace=False)
    n, worst = fd_check(f, x, g, coords, h=1e-6, rtol=1e-5, floor=1e-3 * np.abs(g).max())
    assert n >= 50


def test_interpolate_backward_fd():
    pos, tri, _ = _scene(views=1)
    H = W = 48
    r0, db0 = oracle.rasterize(pos, tri, (H, W), with_db=True, dtype=np.float64)
    rng = R(2)
    attr = rng.random((pos.shape[1], 5))
    wo = rng.standard_normal((1, H, W, 5))
    wda = rng.standard_normal((1, H, W, 4))
    diff = [1, 4]
    ga, gr, gdb = oracle.interpolate_bwd(attr, r0, tri, wo, rast_db=db0, diff_attrs=diff, grad_da=wda,
                                         dtype=np.float64)
    a, r, d = attr.copy(), r0.copy(), db0.copy()

    def f():
        o, da = oracle.interpolate(a, r, tri, rast_db=d, diff_attrs=diff, dtype=np.float64)
        return float((o * wo).sum() + (da * wda).sum())

    # attr (the output is linear in it)
    fd_check(f, a, ga, R(3).choice(a.size, 40, replace=False), h=1e-3, rtol=1e-8, floor=1e-6)
    # rast u, v channels and db at covered pixels
    cov = np.flatnonzero((r0[..., 3] > 0).reshape(-1))
    for ch in (0, 1):
        idx = cov[R(4 + ch).choice(len(cov), 20, replace=False)] * 4 + ch
        fd_check(f, r, gr, idx, h=1e-6, rtol=1e-6, floor=1e-6)
    idx = cov[R(6).choice(len(cov), 20, replace=False)] * 4 + R(7).integers(0, 4, 20)
    fd_check(f, d, gdb, idx, h=1e-6, rtol=1e-6, floor=1e-6)


def test_antialias_backward_fd():
    pos, tri, _ = _scene(res=64, views=2, subdiv=3)
    H = W = 64
    r0 = oracle.rasterize(pos, tri, (H, W), dtype=np.float64)
    rng = R(8)
    color = rng.random((2, H, W, 3))
    wo = rng.standard_normal(color.shape)
    opp = oracle.topology(tri)
    gc, gp = oracle.antialias_bwd(color, r0, pos, tri, wo, opp, pos_gradient_boost=1.0, dtype=np.float64)
    c, x = color.copy(), pos.copy()

    def f():
        return float((oracle.antialias(c, r0, x, tri, opp, dtype=np.float64) * wo).sum())

    fd_check(f, c, gc, R(9).choice(c.size, 40, replace=False), h=1e-3, rtol=1e-8, floor=1e-6)
    nz = np.flatnonzero(np.abs(gp) > 1e-6 * np.abs(gp).max())
    assert len(nz) > 20
    n, _ = fd_check(f, x, gp, R(10).choice(nz, min(60, len(nz)), replace=False), h=1e-7, rtol=1e-4,
                    floor=1e-3 * np.abs(gp).max())
    assert n >= 20


def test_antialias_pos_gradient_boost_scales_position_gradient_only():
    pos, tri, _ = _scene(res=40, views=1)
    r0 = oracle.rasterize(pos, tri, (40, 40), dtype=np.float64)
    color = R(11).random((1, 40, 40, 3))
    wo = R(12).standard_normal(color.shape)
    gc1, gp1 = oracle.antialias_bwd(color, r0, pos, tri, wo, pos_gradient_boost=1.0, dtype=np.float64)
    gc3, gp3 = oracle.antialias_bwd(color, r0, pos, tri, wo, pos_gradient_boost=3.0, dtype=np.float64)
    np.testing.assert_array_equal(gc1, gc3)
    np.testing.assert_allclose(gp3, 3.0 * gp1, rtol=1e-12)


def test_float32_gradients_match_float64():
    pos64, tri, _ = _scene(res=64, views=2)
    pos32 = pos64.astype(np.float32)
    H = W = 64
    r32 = oracle.rasterize(pos32, tri, (H, W))
    r64 = oracle.rasterize_fixed(pos64, tri, ids_f64(r32), dtype=np.float64)
    rng = R(13)
    attr = rng.random((pos64.shape[1], 4))
    go = rng.standard_normal((2, H, W, 4))
    ga32, gr32 = oracle.interpolate_bwd(attr.astype(np.float32), r32, tri, go.astype(np.float32))
    ga64, gr64 = oracle.interpolate_bwd(attr, r64, tri, go, dtype=np.float64)
    np.testing.assert_allclose(ga32, ga64, rtol=0, atol=1e-4 * np.abs(ga64).max())
    gp32 = oracle.rasterize_bwd(pos32, tri, r32, gr32)
    gp64 = oracle.rasterize_bwd(pos64, tri, r64, gr64, dtype=np.float64)
    np.testing.assert_allclose(gp32, gp64, rtol=0, atol=1e-4 * np.abs(gp64).max())
    col = rng.random((2, H, W, 4))
    gc32, gq32 = oracle.antialias_bwd(col.astype(np.float32), r32, pos32, tri, go.astype(np.float32))
    gc64, gq64 = oracle.antialias_bwd(col, r64, pos64, tri, go, dtype=np.float64)
    np.testing.assert_allclose(gc32, gc64, rtol=0, atol=1e-5 * np.abs(gc64).max())
    np.testing.assert_allclose(gq32, gq64, rtol=0, atol=1e-4 * np.abs(gq64).max())


def test_full_pipeline_fd_world_space():
    """World vertices -> clip (per-view VP) -> rasterize -> interpolate -> antialias -> L2."""
    from diffrp_amd.geometry import noisy_icosphere
    verts_t, tri_t = noisy_icosphere(3, 0.8, 0.01, seed=0)
    verts = verts_t.double().numpy()
    tri = tri_t.numpy()
    H = W = 64
    VP = np.stack([vp_from_direction(d, H).double().numpy() for d in ([0, 0, 1], [0.6, 0.8, 0])])
    rng = R(14)
    attr = rng.random((len(verts), 3))
    target = rng.random((2, H, W, 3))
    from tests.fd import clip_f64
    ids = ids_f64(oracle.rasterize(clip_f64(verts, VP).astype(np.float32), tri, (H, W)))
    opp = oracle.topology(tri)
    gv, ga = pipeline_grad_f64(verts, attr, tri, VP, ids, target, opp)
    v, a = verts.copy(), attr.copy()
    f = lambda: pipeline_f64(v, a, tri, VP, ids, target, opp)  # noqa: E731
    nz = np.flatnonzero(np.abs(gv) > 1e-4 * np.abs(gv).max())
    n, _ = fd_check(f, v, gv, R(15).choice(nz, 32, replace=False), h=1e-4, rtol=1e-4,
                    floor=1e-3 * np.abs(gv).max())
    assert n >= 24
    fd_check(f, a, ga, R(16).choice(a.size, 16, replace=False), h=1e-4, rtol=1e-6, floor=1e-6)


def test_window_fd_equals_full_image_fd():
    """The windowed loss used for large-image finite differences moves exactly as
    the full-image loss does when one vertex moves (visibility frozen)."""
    from diffrp_amd.geometry import noisy_icosphere
    from tests.fd import central_diff, clip_f64, vertex_window, window_loss_fn
    verts_t, tri_t = noisy_icosphere(3, 0.8, 0.01, seed=1)
    verts, tri = verts_t.double().numpy(), tri_t.numpy()
    H, W = 96, 80
    VP = vp_from_direction([0.2, 0.9, 0.4], H).double().numpy()[None]
    VP[0, 0] *= H / W  # keep the aspect ratio of a W x H image
    rng = R(31)
    attr = rng.random((len(verts), 4))
    target = rng.random((1, H, W, 4))
    ids = ids_f64(oracle.rasterize(clip_f64(verts, VP).astype(np.float32), tri, (H, W)))
    opp = oracle.topology(tri)
    full = lambda: pipeline_f64(verts, attr, tri, VP, ids, target, opp)  # noqa: E731
    visible = np.unique(tri[np.unique(ids[0, ..., 3]).astype(int)[1:] - 1])
    checked = 0
    for k in R(32).choice(visible, 12, replace=False):
        win = vertex_window(ids[0, ..., 3], tri, k)
        assert win is not None and (win[1] - win[0]) * (win[3] - win[2]) < H * W // 4
        local = window_loss_fn(verts, attr, tri, VP, ids, target, opp, win)
        for x, i in ((verts, 3 * k + int(rng.integers(3))), (attr, 4 * k + int(rng.integers(4)))):
            flat = x.reshape(-1)
            d_full = central_diff(full, flat, i, 1e-5)
            d_win = central_diff(local, flat, i, 1e-5)
            assert abs(d_full - d_win) <= 1e-6 * max(1.0, abs(d_full)), (k, i, d_full, d_win)
            checked += 1
    assert checked == 24
