"""Pins of the C oracle's rasteriser (CPU).

Every check is independent of the oracle's own code path:
  * the reference's documented README render (coverage and colours, exact);
  * brute force coverage in exact integer arithmetic on tiny random scenes;
  * watertightness of shared edges with pixel centres exactly on edges;
  * the depth test as an explicit minimum over per-triangle fragments;
  * barycentrics / depth against screen-space float64 closed forms;
  * the tri-id codec's closed form (continuity at 2^24).
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import glue
from tests.scenes import noisy_clip, random_soup, readme_clip


@pytest.fixture(scope='module', autouse=True)
def _built(oracle_lib):
    return oracle_lib


def test_readme_render_coverage_matches_reference_png():
    fx = np.load('tests/golden/readme_normals_512.npz')['rgba_bottom_first']
    pos, tri, _, _ = readme_clip(512)
    r = oracle.rasterize(pos.numpy(), tri.numpy(), (512, 512))
    np.testing.assert_array_equal(r[0, ..., 3] > 0, fx[..., 3] > 0)
    assert int((fx[..., 3] > 0).sum()) == 190260


def test_readme_render_colors_match_reference_png():
    """Raster + diffrp's normal shading glue + to_pil reproduce the documented image bit for bit."""
    fx = np.load('tests/golden/readme_normals_512.npz')['rgba_bottom_first']
    _, tri, verts, cam = readme_clip(512)
    sv, st, sn = glue.face_soup_flat(verts, tri)
    clip = torch.nn.functional.pad(sv, (0, 1), value=1.0) @ (cam.P('cpu') @ cam.V('cpu')).T
    r = oracle.rasterize(clip.numpy(), st.numpy(), (512, 512))
    img = glue.false_color_camera_space_normal(torch.from_numpy(r), sv, st, sn, torch.eye(4), cam.V('cpu'))
    np.testing.assert_array_equal(glue.to_uint8(img)[::-1], fx)


def test_snapping_precision_is_pinned_by_the_reference_render():
    """8 sub-pixel bits would differ from the documented render (the check above is not vacuous)."""
    fx = np.load('tests/golden/readme_normals_512.npz')['rgba_bottom_first'][..., 3] > 0
    pos, tri, _, _ = readme_clip(512)
    p = pos.double().numpy()
    P = (p[:, :2] / p[:, 3:] + 1) * 256
    def cover(bits):
        Ps = np.rint(P * 2 ** bits) / 2 ** bits
        m = np.zeros((512, 512), bool)
        cy, cx = np.mgrid[0:512, 0:512] + 0.5
        for t in tri.numpy():
            a, b, c = Ps[t]
            if (b[0] - a[0]) * (c[1] - a[1]) - (c[0] - a[0]) * (b[1] - a[1]) <= 0:
                continue
            x0, x1 = int(np.floor(min(a[0], b[0], c[0]))), int(np.ceil(max(a[0], b[0], c[0])))
            y0, y1 = int(np.floor(min(a[1], b[1], c[1]))), int(np.ceil(max(a[1], b[1], c[1])))
            X, Y = cx[y0:y1 + 1, x0:x1 + 1], cy[y0:y1 + 1, x0:x1 + 1]
            ins = np.ones_like(X, bool)
            for p0, q0 in ((a, b), (b, c), (c, a)):
                ins &= (q0[0] - p0[0]) * (Y - p0[1]) - (q0[1] - p0[1]) * (X - p0[0]) >= 0
            m[y0:y1 + 1, x0:x1 + 1] |= ins
        return m
    assert (cover(4) ^ fx).sum() == 0
    assert (cover(8) ^ fx).sum() > 0


# ----------------------------------------------------------------------------
def _snap_f32(pos, H, W):
    """Independent float32 snapping: rint(((x/w)+1) * (W/2) * 16) with numpy float32 ops."""
    p = pos.astype(np.float32)
    nx = p[:, 0] / p[:, 3]
    ny = p[:, 1] / p[:, 3]
    fx = (nx + np.float32(1)) * np.float32(W * 0.5)
    fy = (ny + np.float32(1)) * np.float32(H * 0.5)
    return np.rint(fx * np.float32(16)).astype(np.int64), np.rint(fy * np.float32(16)).astype(np.int64)


def _brute_cover(X, Y, H, W):
    """Top-left rule from its definition: a pixel centre is inside when it is strictly
    inside every edge, or on an edge that is a top edge (horizontal, interior below)
    or a left edge (interior to the right), for y-up coordinates."""
    area = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0])
    cov = np.zeros((H, W), bool)
    if area == 0:
        return cov
    order = [0, 1, 2] if area > 0 else [0, 2, 1]
    xs, ys = [int(X[i]) for i in order], [int(Y[i]) for i in order]
    for r in range(H):
        for c in range(W):
            px, py = 16 * c + 8, 16 * r + 8
            ok = True
            for k in range(3):
                x0, y0, x1, y1 = xs[k], ys[k], xs[(k + 1) % 3], ys[(k + 1) % 3]
                # interior on the left of the directed edge (CCW, y up)
                side = (x1 - x0) * (py - y0) - (y1 - y0) * (px - x0)
                if side > 0:
                    continue
                if side < 0:
                    ok = False
                    break
                top = y1 == y0 and x1 < x0          # horizontal, interior below
                left = y1 < y0                      # going down: interior to its right
                if not (top or left):
                    ok = False
                    break
            cov[r, c] = ok
    return cov


@pytest.mark.parametrize('seed', range(6))
def test_coverage_brute_force_tiny(seed):
    rng = np.random.default_rng(seed)
    H, W = 12, 10
    n = 6
    pos = np.zeros((3 * n, 4))
    # vertices on the pixel-centre lattice half of the time: exact edge ties
    for i in range(3 * n):
        w = rng.uniform(0.5, 2.0) if i % 2 else 1.0
        if rng.random() < 0.5:
            X = (rng.integers(-2, W + 2) + 0.5) / W * 2 - 1
            Y = (rng.integers(-2, H + 2) + 0.5) / H * 2 - 1
        else:
            X, Y = rng.uniform(-1.2, 1.2, 2)
        pos[i] = [X * w, Y * w, rng.uniform(-0.5, 0.5) * w, w]
    pos = pos.astype(np.float32)
    tri = np.arange(3 * n, dtype=np.int32).reshape(n, 3)
    ranges = np.array([[t, 1] for t in range(n)], np.int32)
    r = oracle.rasterize(pos, tri, (H, W), ranges=ranges)
    X, Y = _snap_f32(pos, H, W)
    for t in range(n):
        want = _brute_cover(X[3 * t:3 * t + 3], Y[3 * t:3 * t + 3], H, W)
        got = r[t, ..., 3] > 0
        np.testing.assert_array_equal(got, want, err_msg=f'triangle {t}')


def test_watertight_shared_edges_exact_ties():
    """A jittered grid mesh whose vertices sit on pixel centres: every pixel centre
    inside the grid is covered by exactly one triangle (no cracks, no double hits)."""
    rng = np.random.default_rng(5)
    H = W = 48
    n = 6
    xs = np.linspace(4, W - 4, n + 1)
    g = np.stack(np.meshgrid(xs, xs), -1)
    g[1:-1, 1:-1] += rng.integers(-2, 3, size=g[1:-1, 1:-1].shape)
    g = g + 0.5                                             # pixel centres
    verts = g.reshape(-1, 2)
    ndc = verts / np.array([W, H]) * 2 - 1
    pos = np.concatenate([ndc, np.zeros((len(ndc), 1)), np.ones((len(ndc), 1))], 1).astype(np.float32)
    tris = []
    for i in range(n):
        for j in range(n):
            a, b, c, d = i * (n + 1) + j, i * (n + 1) + j + 1, (i + 1) * (n + 1) + j, (i + 1) * (n + 1) + j + 1
            if (i + j) % 2:
                tris += [(a, b, d), (a, d, c)]
            else:
                tris += [(a, b, c), (b, d, c)]
    tri = np.array(tris, np.int32)
    ranges = np.array([[t, 1] for t in range(len(tri))], np.int32)
    r = oracle.rasterize(pos, tri, (H, W), ranges=ranges)
    count = (r[..., 3] > 0).sum(0)
    full = oracle.rasterize(pos, tri, (H, W))[0, ..., 3] > 0
    assert count.max() == 1
    assert (count[full] == 1).all()
    # interior of the grid's bounding square is fully covered
    assert count[10:38, 10:38].min() == 1


def test_depth_test_is_minimum_over_fragments():
    pos, tri = random_soup(120, seed=11)
    H, W = 40, 36
    n = len(tri)
    per = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W), ranges=np.array([[t, 1] for t in range(n)], np.int32))
    full = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W))[0]
    z = np.where(per[..., 3] > 0, per[..., 2], np.inf)                      # [n, H, W]
    best = np.argmin(z, 0)                                                  # first (smallest id) on ties
    hit = np.isfinite(z.min(0))
    ids = oracle.float_to_triidx(full[..., 3])
    np.testing.assert_array_equal(ids > 0, hit)
    np.testing.assert_array_equal(ids[hit] - 1, best[hit])
    np.testing.assert_array_equal(full[..., 2][hit], z.min(0)[hit])


def test_triangle_order_permutation_invariance():
    pos, tri = random_soup(300, seed=12)
    H, W = 50, 50
    a = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W))[0]
    perm = np.random.default_rng(0).permutation(len(tri))
    b = oracle.rasterize(pos.numpy(), tri.numpy()[perm], (H, W))[0]
    ia, ib = oracle.float_to_triidx(a[..., 3]), oracle.float_to_triidx(b[..., 3])
    mapped = np.where(ib > 0, perm[np.maximum(ib - 1, 0)] + 1, 0)
    np.testing.assert_array_equal(ia, mapped)
    np.testing.assert_array_equal(a[..., :3], b[..., :3])


def test_barycentrics_and_depth_against_screen_space_closed_form():
    """u, v, z/w vs the float64 screen-space form: lambda_i from 2D edge functions,
    beta_i = (lambda_i / w_i) / sum(lambda_j / w_j), z/w = sum lambda_i z_i / w_i."""
    pos, tri, _ = noisy_clip(4, 0.01, 200)
    H = W = 200
    r = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W))[0]
    ids = oracle.float_to_triidx(r[..., 3])
    rr, cc = np.nonzero(ids > 0)
    p = pos.double().numpy()
    t = tri.numpy()[ids[rr, cc] - 1]
    px = (2 * cc + 1 - W) / W
    py = (2 * rr + 1 - H) / H
    sx, sy, sz, sw = (p[t][..., k] for k in range(4))
    qx, qy = sx / sw, sy / sw
    def area(ax, ay, bx, by, cx, cy):
        return (bx - ax) * (cy - ay) - (cx - ax) * (by - ay)
    l0 = area(px, py, qx[:, 1], qy[:, 1], qx[:, 2], qy[:, 2])
    l1 = area(qx[:, 0], qy[:, 0], px, py, qx[:, 2], qy[:, 2])
    l2 = area(qx[:, 0], qy[:, 0], qx[:, 1], qy[:, 1], px, py)
    L = l0 + l1 + l2
    lam = np.stack([l0, l1, l2], -1) / L[:, None]
    b = lam / sw
    b = b / b.sum(-1, keepdims=True)
    np.testing.assert_allclose(r[rr, cc, 0], b[:, 0], atol=1e-4)
    np.testing.assert_allclose(r[rr, cc, 1], b[:, 1], atol=1e-4)
    assert np.quantile(np.abs(r[rr, cc, 0] - b[:, 0]), 0.999) < 2e-5
    zw = (lam * sz / sw).sum(-1)
    np.testing.assert_allclose(r[rr, cc, 2], zw, atol=2e-6)
    # covered centres lie inside (up to the 1/16-pixel snap): barycentrics >= -eps
    assert b.min() > -0.2


def test_homogeneous_path_matches_direct_path():
    """Triangles just inside the guard band (direct, fixed point) and the same
    triangles pushed to the homogeneous path by one vertex behind the camera
    give the same fragments where both are valid (here: interior pixels)."""
    H = W = 32
    pos = np.array([[-0.9, -0.9, 0.1, 1.0], [0.9, -0.8, 0.2, 1.0], [0.0, 0.9, 0.3, 1.0]], np.float32)
    tri = np.array([[0, 1, 2]], np.int32)
    a = oracle.rasterize(pos, tri, (H, W))[0]
    # same plane, but a fourth vertex behind the eye makes the triangle (0, 1, 3)
    # cover part of the same screen area through the homogeneous test
    pos2 = np.concatenate([pos, [[0.0, 0.9 * -0.5, 0.3 * -0.5, -0.5]]]).astype(np.float32)
    b = oracle.rasterize(pos2, np.array([[0, 1, 3]], np.int32), (H, W))[0]
    assert (b[..., 3] > 0).sum() > 0
    assert np.isfinite(b).all()


def test_codec_closed_form():
    # valid domain: 0x4a800000 + n stays a finite positive float (n < 0x35000000)
    n = np.array([0, 1, 2, 1000, 2 ** 24 - 1, 2 ** 24, 2 ** 24 + 1, 2 ** 25, 2 ** 29, 0x35000000 - 1], np.int64)
    f = oracle.triidx_to_float(n.astype(np.int32))
    back = oracle.float_to_triidx(f)
    np.testing.assert_array_equal(back, n)
    assert f.view(np.uint32)[5] == 0x4b800000 == 0x4a800000 + 0x01000000   # continuous at 2^24
    assert (np.diff(f.astype(np.float64)) > 0).all()                          # monotone
    small = np.arange(0, 2 ** 24 + 1, 4099, dtype=np.int32)
    np.testing.assert_array_equal(oracle.triidx_to_float(small), small.astype(np.float32))
    # the reference's torch formula (interpolator.py:24-29) restated in glue agrees
    t = torch.from_numpy(n.astype(np.int32))
    np.testing.assert_array_equal(glue.triidx_to_float(t).numpy().view(np.uint32), f.view(np.uint32))
    np.testing.assert_array_equal(glue.float_to_triidx(torch.from_numpy(f)).numpy(), n)


def test_peeled_layers_enumerate_all_fragments_in_depth_order():
    pos, tri = random_soup(60, seed=13)
    H, W = 24, 24
    n = len(tri)
    per = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W), ranges=np.array([[t, 1] for t in range(n)], np.int32))
    frags = (per[..., 3] > 0).sum(0)
    prev = None
    zs = []
    for k in range(int(frags.max()) + 1):
        lay = oracle.rasterize(pos.numpy(), tri.numpy(), (H, W), prev=prev)
        zs.append(np.where(lay[0, ..., 3] > 0, lay[0, ..., 2], np.inf))
        prev = lay
    zs = np.stack(zs)
    np.testing.assert_array_equal(np.isfinite(zs).sum(0), frags)              # every fragment once
    assert np.isinf(zs[-1]).all()                                             # then empty
    z_sorted = np.sort(np.where(per[..., 3] > 0, per[..., 2], np.inf), 0)[:zs.shape[0]]
    np.testing.assert_array_equal(zs, z_sorted)


def test_clip_space_inputs_outside_frustum_are_finite_and_deterministic():
    g = torch.Generator().manual_seed(7)
    pos = torch.randn(600, 4, generator=g) * torch.tensor([3.0, 3.0, 1.0, 1.0])
    pos[:, 3] = pos[:, 3] * 1.5 + 0.3
    pos[:, 2] = pos[:, 2].clamp(-0.9, 0.9) * pos[:, 3].abs()
    pos[::7, 0] *= 1e5
    tri = torch.arange(600, dtype=torch.int32).reshape(200, 3)
    a = oracle.rasterize(pos.numpy(), tri.numpy(), (32, 32))
    b = oracle.rasterize(pos.numpy(), tri.numpy(), (32, 32))
    assert np.isfinite(a).all()
    np.testing.assert_array_equal(a, b)
    assert ((a[..., 2] >= -1) & (a[..., 2] <= 1)).all()
