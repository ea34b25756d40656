"""Independent pins of the oracle's antialias forward, rast_db, out_da and
topology table (CPU).

None of these checks re-evaluates the oracle's own formula:
  * antialias forward: closed-form blends of the analytic antialiasing method
    (Laine et al. 2020, "Modular Primitives for High-Performance
    Differentiable Rendering", sec. 3.3; diffrp's call:
    diffrp/rendering/surface_deferred.py:775) for straight silhouette edges
    placed at known sub-pixel offsets, the identity away from id
    discontinuities, the non-blending of an edge shared by two triangles of
    one orientation, and the nearer-triangle-first rule;
  * rast_db and out_da: central differences of u, v and of the interpolated
    attribute at pixel centres moved by +-h pixels (moving the sample point by
    2h/W in NDC x is the same as moving every clip vertex by -2h/W * w in x),
    visibility frozen, on a non-square image so a swapped axis or scale fails;
  * topology: the invariants of a closed two-manifold (every edge has exactly
    one neighbour, and the relation is symmetric over the same undirected
    edge), edges shared by three triangles and boundary edges giving -1, and a
    dictionary count of undirected edges.
"""
import numpy as np
import pytest

import oracle
from tests.scenes import random_soup

R = np.random.default_rng


@pytest.fixture(scope='module', autouse=True)
def _built(oracle_lib):
    return oracle_lib


# ---------------------------------------------------------------------------
# antialias forward
# ---------------------------------------------------------------------------
W_AA, H_AA = 16, 12


def _ndc(p, n):
    """NDC coordinate of pixel coordinate p (pixel c has its centre at c + 0.5)."""
    return 2.0 * p / n - 1.0


def _edge_triangle(axis, side, edge_pix, z=0.0):
    """One clip-space triangle (w = 1) with a straight edge on the line
    {axis coordinate = edge_pix} (pixel units) crossing the whole image, and its
    other two edges outside the image: it covers the pixels below (side = -1)
    or above (side = +1) that line."""
    if axis == 'x':
        e = _ndc(edge_pix, W_AA)
        verts = [(e, -3.0), (e, 3.0), (3.0 * side, 0.0)]
    else:
        e = _ndc(edge_pix, H_AA)
        verts = [(-3.0, e), (3.0, e), (0.0, 3.0 * side)]
    return [(x, y, z, 1.0) for x, y in verts]


def _blend_expected(color, a_idx, b_idx, d):
    """Laine et al. box-filter blend for a pair of pixel centres one pixel apart:
    pixel a shows the triangle whose edge crosses the segment at distance d from
    a's centre, pixel b shows the other surface.  d < 1/2: a receives
    (1/2 - d) of (b - a); else b receives (d - 1/2) of (a - b)."""
    out = color.copy()
    ca, cb = color[a_idx], color[b_idx]
    if d < 0.5:
        out[a_idx] = ca + (0.5 - d) * (cb - ca)
    else:
        out[b_idx] = cb + (d - 0.5) * (ca - cb)
    return out


def _colour_by_id(rast, palette):
    ids = oracle.float_to_triidx(rast[..., 3])
    return palette[ids].astype(np.float32)


@pytest.mark.parametrize('delta', [0.1, 0.4, 0.6, 0.9])
@pytest.mark.parametrize('axis', ['x', 'y'])
@pytest.mark.parametrize('side', [-1, 1])
def test_antialias_straight_silhouette_closed_form(delta, axis, side):
    k = 7 if axis == 'x' else 5                  # pair of pixels k, k + 1 across the edge
    edge = k + 0.5 + delta
    pos = np.asarray(_edge_triangle(axis, side, edge), np.float32)
    tri = np.asarray([[0, 1, 2]], np.int32)
    rast = oracle.rasterize(pos, tri, (H_AA, W_AA))
    ids = oracle.float_to_triidx(rast[0, ..., 3])
    # coverage: the triangle's side of the edge, whole columns (rows)
    lin = np.arange(W_AA if axis == 'x' else H_AA) + 0.5
    inside = lin < edge if side < 0 else lin > edge
    want_ids = np.broadcast_to(inside[None, :] if axis == 'x' else inside[:, None], (H_AA, W_AA)).astype(np.int32)
    np.testing.assert_array_equal(ids, want_ids)
    palette = R(40).random((2, 3))
    color = _colour_by_id(rast, palette)
    out = oracle.antialias(color, rast, pos, tri)
    # distance from the triangle's pixel centre to the edge
    tri_pix, bg_pix = (k, k + 1) if side < 0 else (k + 1, k)
    d = abs(edge - (tri_pix + 0.5))
    want = color[0].astype(np.float64)
    for line in range(H_AA if axis == 'x' else W_AA):
        a = (line, tri_pix) if axis == 'x' else (tri_pix, line)
        b = (line, bg_pix) if axis == 'x' else (bg_pix, line)
        want = _blend_expected(want, a, b, d)
    np.testing.assert_allclose(out[0], want, rtol=0, atol=2e-6)
    # and float64 agrees with the closed form to rounding
    out64 = oracle.antialias(color.astype(np.float64), oracle.rasterize(pos, tri, (H_AA, W_AA), dtype=np.float64),
                             pos.astype(np.float64), tri, dtype=np.float64)
    np.testing.assert_allclose(out64[0], want, rtol=0, atol=1e-6)


def test_antialias_is_identity_away_from_id_discontinuities():
    pos, tri = random_soup(40, seed=41, spread=1.0)
    pos, tri = pos.numpy(), tri.numpy()
    H, W = 40, 56
    rast = oracle.rasterize(pos, tri, (H, W))
    color = R(42).random((1, H, W, 4)).astype(np.float32)
    out = oracle.antialias(color, rast, pos, tri)
    ids = rast[0, ..., 3]
    edge = np.zeros((H, W), bool)
    dx = ids[:, 1:] != ids[:, :-1]
    dy = ids[1:, :] != ids[:-1, :]
    edge[:, 1:] |= dx
    edge[:, :-1] |= dx
    edge[1:, :] |= dy
    edge[:-1, :] |= dy
    np.testing.assert_array_equal(out[0][~edge], color[0][~edge])
    assert (out[0][edge] != color[0][edge]).any(1).sum() > 50  # the soup's silhouettes do blend


def _quad(flip_second):
    """Two triangles sharing the diagonal of a square larger than the image."""
    pos = np.asarray([(-3, -3, 0, 1), (3, -3, 0, 1), (3, 3, 0, 1), (-3, 3, 0, 1)], np.float32)
    t2 = [0, 3, 2] if flip_second else [0, 2, 3]
    return pos, np.asarray([[0, 1, 2], t2], np.int32)


def test_antialias_shared_edge_of_one_orientation_never_blends():
    pos, tri = _quad(flip_second=False)
    H, W = 24, 20
    rast = oracle.rasterize(pos, tri, (H, W))
    assert set(np.unique(rast[..., 3])) == {1.0, 2.0}
    color = _colour_by_id(rast, R(43).random((3, 3)))
    np.testing.assert_array_equal(oracle.antialias(color, rast, pos, tri), color)
    # control: the same geometry with the second triangle wound the other way is
    # a fold, its shared edge a silhouette, and the diagonal does blend
    pos_f, tri_f = _quad(flip_second=True)
    rast_f = oracle.rasterize(pos_f, tri_f, (H, W))
    np.testing.assert_array_equal(rast_f[..., 3], rast[..., 3])
    assert (oracle.antialias(color, rast_f, pos_f, tri_f) != color).any(-1).sum() >= H


@pytest.mark.parametrize('near', ['left', 'right'])
def test_antialias_tries_the_nearer_triangle_first(near):
    # left triangle: pixels left of x = k + 0.5 + d1; right triangle: pixels right of
    # x = k + 0.5 + d2 (d2 < d1: they overlap between the two edges)
    k, d1, d2 = 7, 0.7, 0.2
    z_left, z_right = (-0.5, 0.5) if near == 'left' else (0.5, -0.5)
    pos = np.asarray(_edge_triangle('x', -1, k + 0.5 + d1, z_left) + _edge_triangle('x', 1, k + 0.5 + d2, z_right),
                     np.float32)
    tri = np.asarray([[0, 1, 2], [3, 4, 5]], np.int32)
    rast = oracle.rasterize(pos, tri, (H_AA, W_AA))
    ids = oracle.float_to_triidx(rast[0, ..., 3])
    np.testing.assert_array_equal(ids[:, k], 1)
    np.testing.assert_array_equal(ids[:, k + 1], 2)
    palette = R(44).random((3, 3))
    color = _colour_by_id(rast, palette)
    out = oracle.antialias(color, rast, pos, tri)
    want = color[0].astype(np.float64)
    for row in range(H_AA):
        if near == 'left':   # the left triangle's edge, seen from pixel k: distance d1
            want = _blend_expected(want, (row, k), (row, k + 1), d1)
        else:                # the right triangle's edge, seen from pixel k + 1: distance 1 - d2
            want = _blend_expected(want, (row, k + 1), (row, k), 1.0 - d2)
    np.testing.assert_allclose(out[0], want, rtol=0, atol=2e-6)


# ---------------------------------------------------------------------------
# rast_db and out_da against screen-space central differences
# ---------------------------------------------------------------------------
H_DB, W_DB = 32, 48


def _shifted(pos, axis, h):
    """Clip positions whose barycentrics at pixel centre p equal those of `pos` at
    p moved by h pixels along `axis` (x: 0, y: 1)."""
    n = W_DB if axis == 0 else H_DB
    p = pos.copy()
    p[..., axis] -= (2.0 * h / n) * p[..., 3]
    return p


def _db_scene():
    pos32, tri = random_soup(24, seed=45, spread=1.0, w_range=(0.6, 2.5))
    pos = pos32.double().numpy()
    tri = tri.numpy()
    rast, db = oracle.rasterize(pos, tri, (H_DB, W_DB), with_db=True, dtype=np.float64)
    cov = rast[0, ..., 3] > 0
    assert cov.sum() > 400
    return pos, tri, rast, db, cov


def test_rast_db_equals_screen_space_central_differences():
    pos, tri, rast, db, cov = _db_scene()
    h = 1e-4
    fd = np.zeros_like(db)
    for axis in (0, 1):
        up = oracle.rasterize_fixed(_shifted(pos, axis, h), tri, rast, dtype=np.float64)
        dn = oracle.rasterize_fixed(_shifted(pos, axis, -h), tri, rast, dtype=np.float64)
        d = (up - dn) / (2 * h)
        fd[..., 0 + axis] = d[..., 0]    # du/dX, du/dY
        fd[..., 2 + axis] = d[..., 1]    # dv/dX, dv/dY
    scale = np.abs(fd[0][cov]).max()
    np.testing.assert_allclose(db[0][cov], fd[0][cov], rtol=0, atol=1e-6 * scale)
    np.testing.assert_array_equal(db[0][~cov], 0)
    # the float32 oracle (the GPU's bit-exact checker) agrees to float32 rounding
    r32, db32 = oracle.rasterize(pos.astype(np.float32), tri, (H_DB, W_DB), with_db=True)
    same = (r32[0, ..., 3] == rast[0, ..., 3]) & cov
    np.testing.assert_allclose(db32[0][same], fd[0][same], rtol=0, atol=1e-4 * scale)


def test_out_da_equals_screen_space_central_differences():
    pos, tri, rast, db, cov = _db_scene()
    attr = R(46).random((pos.shape[0], 5))
    diff = [0, 3, 4]
    _, da = oracle.interpolate(attr, rast, tri, rast_db=db, diff_attrs=diff, dtype=np.float64)
    h = 1e-4
    fd = np.zeros_like(da)
    for axis in (0, 1):
        up = oracle.interpolate(attr, oracle.rasterize_fixed(_shifted(pos, axis, h), tri, rast, dtype=np.float64),
                                tri, dtype=np.float64)
        dn = oracle.interpolate(attr, oracle.rasterize_fixed(_shifted(pos, axis, -h), tri, rast, dtype=np.float64),
                                tri, dtype=np.float64)
        d = (up - dn) / (2 * h)
        for j, ch in enumerate(diff):
            fd[..., 2 * j + axis] = d[..., ch]
    scale = np.abs(fd[0][cov]).max()
    np.testing.assert_allclose(da[0][cov], fd[0][cov], rtol=0, atol=1e-6 * scale)
    np.testing.assert_array_equal(da[0][~cov], 0)


# ---------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------
def _edge_dict(tri):
    edges = {}
    for t, f in enumerate(tri):
        for k in range(3):
            a, b = int(f[k]), int(f[(k + 1) % 3])
            edges.setdefault((min(a, b), max(a, b)), []).append((t, k))
    return edges


def _check_against_edge_count(tri, opp):
    for lst in _edge_dict(tri).values():
        for t, k in lst:
            if len(lst) == 2:
                other = lst[1][0] if lst[0] == (t, k) else lst[0][0]
                assert opp[t, k] == other
            else:
                assert opp[t, k] == -1


def test_topology_closed_mesh_invariants():
    from diffrp_amd.geometry import icosphere
    _, f = icosphere(3, 1.0)
    tri = np.ascontiguousarray(f, np.int32)
    opp = oracle.topology(tri)
    assert (opp >= 0).all()                      # closed: every edge has a neighbour
    for t in range(len(tri)):
        for k in range(3):
            n = opp[t, k]
            assert n != t
            e = {tri[t, k], tri[t, (k + 1) % 3]}
            js = [j for j in range(3) if {tri[n, j], tri[n, (j + 1) % 3]} == e]
            assert len(js) == 1 and opp[n, js[0]] == t  # symmetric, over the same undirected edge
    _check_against_edge_count(tri, opp)


def test_topology_boundary_and_non_manifold_edges():
    from diffrp_amd.geometry import icosphere
    v, f = icosphere(2, 1.0)
    tri = np.ascontiguousarray(f, np.int32)
    V = len(v)
    # a hole: drop triangle 0; its three neighbours lose one neighbour each
    holed = tri[1:]
    opp_h = oracle.topology(holed)
    assert (opp_h == -1).sum() == 3
    _check_against_edge_count(holed, opp_h)
    # a fin on edge (v0, v1) of triangle 5: three triangles share it, all get -1
    a, b = tri[5, 0], tri[5, 1]
    fin = np.concatenate([tri, np.asarray([[a, b, V]], np.int32)])
    opp_f = oracle.topology(fin)
    on_edge = [(t, k) for t in range(len(fin)) for k in range(3)
               if {fin[t, k], fin[t, (k + 1) % 3]} == {a, b}]
    assert len(on_edge) == 3
    for t, k in on_edge:
        assert opp_f[t, k] == -1
    assert (opp_f == -1).sum() == 3 + 2          # plus the fin's two open edges
    _check_against_edge_count(fin, opp_f)
    # a doubled triangle (same three vertices twice): every edge has three triangles
    dup = np.concatenate([tri, tri[7:8]])
    opp_d = oracle.topology(dup)
    assert (opp_d[7] == -1).all() and (opp_d[-1] == -1).all()
    _check_against_edge_count(dup, opp_d)
