"""
The fine raster's 32-bit row spans (records whose vertices lie within 2^14
sub-pixels of the tile's first pixel centre), modelled in numpy with the
kernel's arithmetic: int32 edge constants, a float32 crossing estimate from a
reciprocal that may be one ulp off, clamped to the span window, then one exact
int32 step either way.  The span must equal brute-force coverage (every edge
function >= 0 with the top-left rule) on random, sliver and coordinate-limit
triangles, and no intermediate may leave int32.
"""
import math

import numpy as np

F32 = np.float32
LIM = 1 << 14


def _i32(v):
    assert -(1 << 31) <= v < (1 << 31), v
    return v


def _rcp(x, rng):
    r = F32(1.0) / F32(x)
    k = int(rng.integers(-1, 2))  # the hardware reciprocal is within one ulp
    return r if k == 0 else np.nextafter(r, F32(np.inf) if k > 0 else F32(-np.inf))


def _edges(X, Y):
    out = []
    for e in range(3):
        j = 0 if e == 2 else e + 1
        dx, dy = X[j] - X[e], Y[j] - Y[e]
        out.append((dx, dy, (dy < 0) or (dy == 0 and dx < 0)))
    return out


def _span32(X, Y, edges, y, xl, xr, rng):
    for e, (dx, dy, tl) in enumerate(edges):
        a = _i32(-dy * 16)
        cc = _i32(_i32(dx * (16 * y - Y[e])) + _i32(dy * X[e]) - (0 if tl else 1))
        if a == 0:
            if cc < 0:
                xr = xl - 1
            continue
        q = F32(F32(-F32(cc)) * _rcp(F32(F32(-dy) * F32(16)), rng))
        q = min(max(q, F32(xl - 1)), F32(xr + 1))
        # one step d towards the uncovered side, the same for both directions
        rising = a > 0
        x0 = int(math.ceil(q)) if rising else int(math.floor(q))
        d = -1 if rising else 1
        if _i32(cc + _i32(a * (x0 + d))) >= 0:
            x0 += d
        elif _i32(cc + _i32(a * x0)) < 0:
            x0 -= d
        if rising:
            xl = max(xl, x0)
        else:
            xr = min(xr, x0)
    return xl, xr


def test_rowspan32_equals_brute_force():
    rng = np.random.default_rng(1)
    rows = 0
    for trial in range(6000):
        scale = [40, 400, 4000, LIM - 1][trial % 4]
        X = [int(v) for v in rng.integers(-scale, scale + 1, 3)]
        Y = [int(v) for v in rng.integers(-scale, scale + 1, 3)]
        if trial % 7 == 0:
            Y[1] = Y[0]  # horizontal edge
        if trial % 11 == 0:
            X[2] = X[1]  # vertical edge
        if trial % 13 == 0:  # sliver
            X[1] = X[0] + int(rng.integers(-2, 3))
            Y[2] = Y[0] + int(rng.integers(-2, 3))
        X = [max(-LIM + 1, min(LIM - 1, v)) for v in X]
        Y = [max(-LIM + 1, min(LIM - 1, v)) for v in Y]
        area = (X[1] - X[0]) * (Y[2] - Y[0]) - (X[2] - X[0]) * (Y[1] - Y[0])
        if area == 0:
            continue
        if area < 0:
            X[1], X[2], Y[1], Y[2] = X[2], X[1], Y[2], Y[1]
        edges = _edges(X, Y)
        bx0 = int(rng.integers(0, 32))
        bw = min(int(rng.integers(1, 33)), 32 - bx0)
        for y in range(0, 32, 3):
            xl, xr = _span32(X, Y, edges, y, bx0, bx0 + bw - 1, rng)
            want = [x for x in range(bx0, bx0 + bw)
                    if all(dx * (16 * y - Y[e]) - dy * (16 * x - X[e]) - (0 if tl else 1) >= 0
                           for e, (dx, dy, tl) in enumerate(edges))]
            assert list(range(xl, xr + 1)) == want, (X, Y, y, bx0, bw)
            rows += 1
    assert rows > 50000
