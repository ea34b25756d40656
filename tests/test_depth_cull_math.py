"""
The fine raster's depth cull (``depth_culled`` in the raster kernels), modelled
in numpy with the kernel's float32 arithmetic: eval_pixel's fmas and rounded
products, a reciprocal that may be one ulp off, and the bound's own float
operations in the kernel's order.  The cull skips a triangle when its lower
bound L exceeds the tile's farthest current depth, so it is exact only if L
never exceeds the depth that eval_pixel and the float division give at any
pixel centre of the box.  Checked by brute force over every pixel centre of
the box on random, sliver, near-degenerate, steep-perspective and huge
triangles (the float64 emulation of an fma rounds twice; that differs from the
hardware's single rounding in about one case in 2^29 and stays far inside the
bound's slack), and the bound must be tight enough to cull.
"""
import numpy as np

F32 = np.float32
F64 = np.float64
U = F32(2.0 ** -24)


def _fma(a, b, c):
    return (np.asarray(a, F64) * np.asarray(b, F64) + np.asarray(c, F64)).astype(F32)


def _mul(a, b):
    return (np.asarray(a, F64) * np.asarray(b, F64)).astype(F32)


def _div(a, b):
    return (np.asarray(a, F64) / np.asarray(b, F64)).astype(F32)


def _rcp(x, rng):
    r = _div(F32(1), x)
    k = int(rng.integers(-1, 2))  # the hardware reciprocal is within one ulp
    return r if k == 0 else np.nextafter(r, F32(np.inf) if k > 0 else F32(-np.inf))


def _eval(v, px, py):
    """eval_pixel: (zn, wn) at pixel centres px, py (arrays broadcast)."""
    x, y, z, w = v
    X = [_fma(-w[k], px, x[k]) for k in range(3)]
    Y = [_fma(-w[k], py, y[k]) for k in range(3)]
    e = [None] * 3
    for k in range(3):
        i, j = (k + 1) % 3, (k + 2) % 3
        e[k] = _fma(X[i], Y[j], -_mul(X[j], Y[i]))
    zn = _fma(e[0], z[0], _fma(e[1], z[1], _mul(e[2], z[2])))
    wn = _fma(e[0], w[0], _fma(e[1], w[1], _mul(e[2], w[2])))
    return zn, wn


def _bound(v, px0, px1, py0, py1, rng):
    """The kernel's lower bound L on every computed depth in the box (None: no cull)."""
    x, y, z, w = v
    MX = [max(abs(_fma(-w[k], px0, x[k])), abs(_fma(-w[k], px1, x[k]))) for k in range(3)]
    MY = [max(abs(_fma(-w[k], py0, y[k])), abs(_fma(-w[k], py1, y[k]))) for k in range(3)]
    ez = ew = F32(0)
    for k in range(3):
        i, j = (k + 1) % 3, (k + 2) % 3
        T = _fma(MX[i], MY[j], _mul(MX[j], MY[i]))
        ez = _fma(T, abs(z[k]), ez)
        ew = _fma(T, w[k], ew)
    ez, ew = _mul(ez, F32(8) * U), _mul(ew, F32(8) * U)
    qmin, qa, w0 = F32(np.inf), F32(0), F32(0)
    for c in range(4):
        zn, wn = _eval(v, px1 if c & 1 else px0, py1 if c & 2 else py0)
        if c == 0:
            w0 = F32(abs(wn) - ew)
        q = _mul(zn, _rcp(wn, rng))
        qmin = min(qmin, q) if not np.isnan(q) else F32(np.nan)
        qa = max(qa, abs(q)) if not np.isnan(q) else F32(np.nan)
    if not (w0 > _mul(F32(3), ew)) or not (qa < np.inf) or np.isnan(qmin):
        return None
    d = F32(w0 - ew)
    zb = _mul(F32(2), F32(_mul(qa, F32(1) + F32(5) * U) + _div(ez, d)))
    E = _div(F32(ez + _mul(zb, ew)), d)
    m = F32(F32(_mul(F32(2), E) + _mul(F32(8) * U, qa)) + _mul(F32(2) * U, F32(zb + E)))
    return F32(qmin - m)


def _pix(c, W):
    return _div(F32(2 * c + 1 - W), F32(W))


def _check(v, c0, c1, r0, r1, H, W, rng):
    """(bound or None, smallest computed depth over the box's pixel centres)."""
    cols = _pix(np.arange(c0, c1 + 1), W)
    rows = _pix(np.arange(r0, r1 + 1), H)
    PX, PY = np.meshgrid(cols, rows)
    zn, wn = _eval(v, PX, PY)
    with np.errstate(all='ignore'):
        zw = _div(zn, wn)
    L = _bound(v, cols[0], cols[-1], rows[0], rows[-1], rng)
    return L, zw


def _clip(sx, sy, zd, w):
    w = np.asarray(w, F32)
    return (_mul(sx, w), _mul(sy, w), _mul(zd, w), w)


def _cases(rng, n):
    for _ in range(n):
        H = W = int(rng.choice([64, 256, 512, 2048]))
        kind = rng.integers(0, 6)
        cx, cy = rng.uniform(-1, 1, 2)
        px_size = 2.0 / W
        if kind == 0:    # ordinary, 0.05 .. 40 pixels
            s = px_size * 10 ** rng.uniform(-1.3, 1.6)
            sx, sy = cx + s * rng.normal(size=3), cy + s * rng.normal(size=3)
        elif kind == 1:  # sliver: one vertex near the opposite edge's line
            s = px_size * 10 ** rng.uniform(0, 1.5)
            a, b = rng.normal(size=2) * s, rng.normal(size=2) * s
            t = rng.uniform(0, 1)
            off = rng.normal(size=2) * px_size * 10 ** rng.uniform(-6, -1)
            p2 = a + t * (b - a) + off
            sx, sy = cx + np.array([a[0], b[0], p2[0]]), cy + np.array([a[1], b[1], p2[1]])
        elif kind == 2:  # long thin (crumpled-mesh) sliver
            d = rng.normal(size=2)
            d /= np.linalg.norm(d)
            L = px_size * 10 ** rng.uniform(1, 2.5)
            wdt = px_size * 10 ** rng.uniform(-3, 0)
            nrm = np.array([-d[1], d[0]])
            pts = [np.zeros(2), d * L, d * L * rng.uniform(0.2, 0.8) + nrm * wdt]
            sx, sy = cx + np.array([p[0] for p in pts]), cy + np.array([p[1] for p in pts])
        elif kind == 3:  # sub-pixel
            s = px_size * 10 ** rng.uniform(-3, -0.5)
            sx, sy = cx + s * rng.normal(size=3), cy + s * rng.normal(size=3)
        elif kind == 4:  # huge, vertices far off screen
            s = 10 ** rng.uniform(0, 3)
            sx, sy = cx + s * rng.normal(size=3), cy + s * rng.normal(size=3)
        else:            # near-degenerate: three almost collinear vertices
            d = rng.normal(size=2)
            s = px_size * 10 ** rng.uniform(0, 2)
            ts = rng.uniform(-1, 1, 3)
            off = rng.normal(size=3) * px_size * 10 ** rng.uniform(-7, -3)
            sx, sy = cx + ts * d[0] * s - off * d[1], cy + ts * d[1] * s + off * d[0]
        zd = np.clip(rng.uniform(-1, 1) + rng.normal(size=3) * 10 ** rng.uniform(-4, 0), -1.5, 1.5)
        w = 10 ** rng.uniform(-1, 2, 3) if rng.uniform() < 0.5 else np.full(3, 10 ** rng.uniform(-1, 2)) * \
            (1 + rng.normal(size=3) * 1e-3)
        v = _clip(np.asarray(sx, F32), np.asarray(sy, F32), np.asarray(zd, F32), w)
        # the kernel's box: pixel centres inside the triangle's screen box, within
        # one 32 x 8 tile of the image; half the cases instead take any box
        # (up to 32 x 32) near the triangle (the bound holds for every box)
        fx, fy = (np.asarray(sx) + 1) * 0.5 * W - 0.5, (np.asarray(sy) + 1) * 0.5 * H - 0.5
        c0, c1 = int(np.ceil(fx.min())), int(np.floor(fx.max()))
        r0, r1 = int(np.ceil(fy.min())), int(np.floor(fy.max()))
        if rng.uniform() < 0.5 and c0 <= c1 and r0 <= r1 and c1 >= 0 and r1 >= 0 and c0 < W and r0 < H:
            tx = int(rng.integers(max(c0, 0), min(c1, W - 1) + 1)) // 32 * 32
            ty = int(rng.integers(max(r0, 0), min(r1, H - 1) + 1)) // 8 * 8
            yield (v, max(c0, tx), min(c1, tx + 31, W - 1), max(r0, ty), min(r1, ty + 7, H - 1), H, W, True)
            continue
        ccx = int(np.clip((np.mean(sx) + 1) * 0.5 * W + rng.normal() * 8, 0, W - 1))
        ccy = int(np.clip((np.mean(sy) + 1) * 0.5 * H + rng.normal() * 8, 0, H - 1))
        bw, bh = int(rng.integers(1, 33)), int(rng.integers(1, 33))
        c0, r0 = max(0, ccx - bw // 2), max(0, ccy - bh // 2)
        yield v, c0, min(W - 1, c0 + bw - 1), r0, min(H - 1, r0 + bh - 1), H, W, False


def test_depth_cull_bound_is_below_every_computed_depth():
    rng = np.random.default_rng(7)
    used = 0
    for v, c0, c1, r0, r1, H, W, _ in _cases(rng, 6000):
        L, zw = _check(v, c0, c1, r0, r1, H, W, rng)
        if L is None:
            continue
        assert np.isfinite(zw).all(), 'a cullable triangle gave a non-finite depth'
        zmin = float(zw.min())
        assert float(L) <= zmin, (float(L), zmin, v, (c0, c1, r0, r1, H, W))
        used += 1
    assert used > 3000  # the bound applies to most triangles


def test_depth_cull_separates_crumpled_mesh_surfaces():
    """Slivers of a crumpled sphere (10-50 px long, 0.3-3 px wide, w ~ 3,
    near 0.1, far 10): the bound applies and its margin (below ~1e-4 at the
    median, 7e-4 at most) is far below the 0.03 that separates the sphere's
    front and back surfaces in depth and within the 5e-4 of its noise."""
    rng = np.random.default_rng(1)
    margins, n = [], 0
    for W in (256, 2048):
        for wid in (0.3, 1.0, 3.0):
            for _ in range(100):
                px = 2.0 / W
                d = rng.normal(size=2)
                d /= np.linalg.norm(d)
                Lp = px * rng.uniform(10, 50)
                pts = [np.zeros(2), d * Lp, d * Lp * rng.uniform(0.2, 0.8) + np.array([-d[1], d[0]]) * px * wid]
                c = rng.uniform(-0.5, 0.5, 2)
                sx = np.array([c[0] + q[0] for q in pts])
                sy = np.array([c[1] + q[1] for q in pts])
                v = _clip(sx.astype(F32), sy.astype(F32), (0.97 + rng.normal(size=3) * 1e-3).astype(F32),
                          3.0 + rng.normal(size=3) * 0.05)
                fx, fy = (sx + 1) * 0.5 * W - 0.5, (sy + 1) * 0.5 * W - 0.5
                c0, c1 = int(np.ceil(fx.min())), int(np.floor(fx.max()))
                r0, r1 = int(np.ceil(fy.min())), int(np.floor(fy.max()))
                if c0 > c1 or r0 > r1:
                    continue
                n += 1
                L, zw = _check(v, c0, min(c1, c0 + 31), r0, min(r1, r0 + 7), W, W, rng)
                assert L is not None
                assert float(L) <= float(zw.min())
                margins.append(float(zw.min()) - float(L))
    assert n > 500
    assert np.median(margins) < 2e-4 and max(margins) < 1e-3


def test_depth_cull_refuses_slivers_and_crossings():
    rng = np.random.default_rng(8)
    # a degenerate triangle (collinear vertices) is never culled
    sx = np.array([0.1, 0.2, 0.3], F32)
    v = _clip(sx, sx * F32(0.5), np.array([0.1, 0.2, 0.3], F32), np.array([2, 2, 2], F32))
    assert _bound(v, _pix(10, 64), _pix(20, 64), _pix(10, 64), _pix(20, 64), rng) is None
    # a depth plane seen edge-on across the box still bounds correctly
    v = _clip(np.array([-0.5, 0.5, 0.0], F32), np.array([-0.5, -0.5, 0.5], F32),
              np.array([-0.9, 0.9, 0.0], F32), np.array([1.0, 50.0, 3.0], F32))
    L, zw = _check(v, 10, 40, 10, 40, 64, 64, rng)
    assert L is None or float(L) <= float(np.nanmin(zw))
