"""
CPU oracle of the differentiable rasterisation path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` import this package, always as the checker.  The product
(``diffrp_amd``) never imports it and has no CPU fallback.

Two layers:
  * ``drp_oracle.c`` (built into ``build/libdrp_oracle.so``): the rasteriser,
    interpolation, antialias and their backward passes as scalar loops, in
    float32 (bit-exact checker of the HIP kernels) and float64 (finite
    differences).  Wrapped below with numpy arrays.
  * ``glue.py``: diffrp's own PyTorch glue on the path (interpolation formula,
    tri-id codec, G-buffer scatter, layer composition, normal shading),
    restated on torch-CPU.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, 'build', 'libdrp_oracle.so')
_lib = None


def build(force: bool = False) -> str:
    """Compile the C oracle with its Makefile (gcc, no fast-math, no contraction)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(['make', '-s', '-C', _HERE] + (['-B'] if force else []), check=True)
    return _LIB_PATH


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        for name in ('ora_rasterize', 'ora_rasterize_fixed', 'ora_rasterize_bwd', 'ora_interpolate',
                     'ora_interpolate_bwd', 'ora_antialias', 'ora_antialias_bwd'):
            for sfx in ('_f32', '_f64'):
                getattr(_lib, name + sfx).restype = ctypes.c_int
        _lib.ora_triidx_to_float.restype = ctypes.c_float
        _lib.ora_triidx_to_float.argtypes = [ctypes.c_int32]
        _lib.ora_float_to_triidx.restype = ctypes.c_int32
        _lib.ora_float_to_triidx.argtypes = [ctypes.c_float]
        _lib.ora_set_threads.argtypes = [ctypes.c_int]
    return _lib


def set_threads(n: int):
    """Threads of the oracle's loops (n <= 0: every processor); results do not depend on it."""
    lib().ora_set_threads(int(n))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _real(dtype):
    dt = np.dtype(dtype)
    if dt == np.float32:
        return np.float32, '_f32'
    if dt == np.float64:
        return np.float64, '_f64'
    raise TypeError(f'oracle works in float32 or float64, got {dt}')


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


def _pos_shape(pos):
    pos = np.asarray(pos)
    if pos.ndim == 2:
        return 0, pos.shape[0]
    return 1, pos.shape[1]


def _check(rc, name):
    if rc != 0:
        raise RuntimeError(f'oracle {name} failed with code {rc}')


def rasterize(pos, tri, resolution, ranges=None, prev=None, with_db=False, dtype=np.float32):
    """rast [n,H,W,4] (and rast_db) for clip-space ``pos`` [V,4] or [n,V,4].

    ``ranges`` [n,2] (start, count) selects range mode for 2-D ``pos``.
    ``prev`` is the previous depth layer for peeling.  In float64 the id
    channel holds (triangle + 1) as a plain number instead of the codec.
    """
    real, sfx = _real(dtype)
    pos = _c(pos, real)
    tri = _c(tri, np.int32)
    H, W = resolution
    batched, V = _pos_shape(pos)
    if batched:
        n = pos.shape[0]
        rng = None
    elif ranges is not None:
        rng = _c(ranges, np.int32).reshape(-1, 2)
        n = rng.shape[0]
    else:
        n, rng = 1, None
    prev = _c(prev, real)
    rast = np.zeros((n, H, W, 4), real)
    db = np.zeros((n, H, W, 4), real) if with_db else None
    rc = getattr(lib(), 'ora_rasterize' + sfx)(
        _ptr(pos), ctypes.c_int(n), ctypes.c_int(V), ctypes.c_int(batched), _ptr(tri), ctypes.c_int(len(tri)),
        _ptr(rng), ctypes.c_int(H), ctypes.c_int(W), _ptr(prev), _ptr(rast), _ptr(db))
    _check(rc, 'rasterize')
    return (rast, db) if with_db else rast


def rasterize_fixed(pos, tri, rast_in, with_db=False, dtype=np.float64):
    """(u, v, z/w) re-evaluated at the pixels of ``rast_in``'s id buffer (visibility frozen)."""
    real, sfx = _real(dtype)
    pos = _c(pos, real)
    tri = _c(tri, np.int32)
    rast_in = _c(rast_in, real)
    n, H, W, _ = rast_in.shape
    batched, V = _pos_shape(pos)
    out = np.zeros_like(rast_in)
    db = np.zeros_like(rast_in) if with_db else None
    rc = getattr(lib(), 'ora_rasterize_fixed' + sfx)(
        _ptr(pos), ctypes.c_int(n), ctypes.c_int(V), ctypes.c_int(batched), _ptr(tri), ctypes.c_int(len(tri)),
        ctypes.c_int(H), ctypes.c_int(W), _ptr(rast_in), _ptr(out), _ptr(db))
    _check(rc, 'rasterize_fixed')
    return (out, db) if with_db else out


def rasterize_bwd(pos, tri, rast, grad_rast, grad_db=None, dtype=np.float32):
    real, sfx = _real(dtype)
    pos = _c(pos, real)
    tri = _c(tri, np.int32)
    rast = _c(rast, real)
    grad_rast = _c(grad_rast, real)
    grad_db = _c(grad_db, real)
    n, H, W, _ = rast.shape
    batched, V = _pos_shape(pos)
    gpos = np.zeros_like(pos)
    rc = getattr(lib(), 'ora_rasterize_bwd' + sfx)(
        _ptr(pos), ctypes.c_int(n), ctypes.c_int(V), ctypes.c_int(batched), _ptr(tri), ctypes.c_int(len(tri)),
        ctypes.c_int(H), ctypes.c_int(W), _ptr(rast), _ptr(grad_rast), _ptr(grad_db), _ptr(gpos))
    _check(rc, 'rasterize_bwd')
    return gpos


def interpolate(attr, rast, tri, rast_db=None, diff_attrs=None, dtype=np.float32):
    """out [n,H,W,C] (and out_da [n,H,W,2*len(diff_attrs)] when ``diff_attrs`` is given)."""
    real, sfx = _real(dtype)
    attr = _c(attr, real)
    rast = _c(rast, real)
    tri = _c(tri, np.int32)
    n, H, W, _ = rast.shape
    batched = 1 if attr.ndim == 3 else 0
    V, C = attr.shape[-2], attr.shape[-1]
    out = np.zeros((n, H, W, C), real)
    didx = None if diff_attrs is None else np.ascontiguousarray(diff_attrs, np.int32)
    nd = 0 if didx is None else len(didx)
    out_da = np.zeros((n, H, W, 2 * nd), real) if didx is not None else None
    rc = getattr(lib(), 'ora_interpolate' + sfx)(
        _ptr(attr), ctypes.c_int(n), ctypes.c_int(V), ctypes.c_int(batched), ctypes.c_int(C), _ptr(rast),
        ctypes.c_int(H), ctypes.c_int(W), _ptr(tri), ctypes.c_int(len(tri)), _ptr(_c(rast_db, real)), _ptr(didx),
        ctypes.c_int(nd), _ptr(out), _ptr(out_da))
    _check(rc, 'interpolate')
    return (out, out_da) if didx is not None else out


def interpolate_bwd(attr, rast, tri, grad_out, rast_db=None, diff_attrs=None, grad_da=None, dtype=np.float32):
    """(grad_attr, grad_rast[, grad_db])."""
    real, sfx = _real(dtype)
    attr = _c(attr, real)
    rast = _c(rast, real)
    tri = _c(tri, np.int32)
    grad_out = _c(grad_out, real)
    n, H, W, _ = rast.shape
    batched = 1 if attr.ndim == 3 else 0
    V, C = attr.shape[-2], attr.shape[-1]
    didx = None if diff_attrs is None else np.ascontiguousarray(diff_attrs, np.int32)
    nd = 0 if didx is None else len(didx)
    gattr = np.zeros_like(attr)
    grast = np.zeros((n, H, W, 4), real)
    gdb = np.zeros((n, H, W, 4), real) if didx is not None else None
    rc = getattr(lib(), 'ora_interpolate_bwd' + sfx)(
        _ptr(attr), ctypes.c_int(n), ctypes.c_int(V), ctypes.c_int(batched), ctypes.c_int(C), _ptr(rast),
        ctypes.c_int(H), ctypes.c_int(W), _ptr(tri), ctypes.c_int(len(tri)), _ptr(_c(rast_db, real)), _ptr(didx),
        ctypes.c_int(nd), _ptr(grad_out), _ptr(_c(grad_da, real)), _ptr(gattr), _ptr(grast), _ptr(gdb))
    _check(rc, 'interpolate_bwd')
    return (gattr, grast, gdb) if didx is not None else (gattr, grast)


def topology(tri):
    """opp [F,3]: triangle across edge (v_k, v_k+1), -1 for boundary / non-manifold edges."""
    tri = _c(tri, np.int32)
    opp = np.zeros((len(tri), 3), np.int32)
    _check(lib().ora_topology(_ptr(tri), ctypes.c_int(len(tri)), _ptr(opp)), 'topology')
    return opp


def antialias(color, rast, pos, tri, opp=None, dtype=np.float32):
    real, sfx = _real(dtype)
    color = _c(color, real)
    rast = _c(rast, real)
    pos = _c(pos, real)
    tri = _c(tri, np.int32)
    opp = topology(tri) if opp is None else _c(opp, np.int32)
    n, H, W, C = color.shape
    batched, V = _pos_shape(pos)
    out = np.zeros_like(color)
    rc = getattr(lib(), 'ora_antialias' + sfx)(
        _ptr(color), ctypes.c_int(n), ctypes.c_int(H), ctypes.c_int(W), ctypes.c_int(C), _ptr(rast), _ptr(pos),
        ctypes.c_int(V), ctypes.c_int(batched), _ptr(tri), ctypes.c_int(len(tri)), _ptr(opp), _ptr(out))
    _check(rc, 'antialias')
    return out


def antialias_bwd(color, rast, pos, tri, grad_out, opp=None, pos_gradient_boost=1.0, dtype=np.float32):
    """(grad_color, grad_pos)."""
    real, sfx = _real(dtype)
    color = _c(color, real)
    rast = _c(rast, real)
    pos = _c(pos, real)
    tri = _c(tri, np.int32)
    grad_out = _c(grad_out, real)
    opp = topology(tri) if opp is None else _c(opp, np.int32)
    n, H, W, C = color.shape
    batched, V = _pos_shape(pos)
    gcol = np.zeros_like(color)
    gpos = np.zeros_like(pos)
    ctype = ctypes.c_float if real == np.float32 else ctypes.c_double
    rc = getattr(lib(), 'ora_antialias_bwd' + sfx)(
        _ptr(color), ctypes.c_int(n), ctypes.c_int(H), ctypes.c_int(W), ctypes.c_int(C), _ptr(rast), _ptr(pos),
        ctypes.c_int(V), ctypes.c_int(batched), _ptr(tri), ctypes.c_int(len(tri)), _ptr(opp), _ptr(grad_out),
        _ptr(gcol), _ptr(gpos), ctype(pos_gradient_boost))
    _check(rc, 'antialias_bwd')
    return gcol, gpos


def triidx_to_float(n):
    n = np.ascontiguousarray(n, np.int32)
    out = np.zeros(n.shape, np.float32)
    lib().ora_encode_ids(_ptr(n), _ptr(out), ctypes.c_int64(n.size))
    return out


def float_to_triidx(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(x.shape, np.int32)
    lib().ora_decode_ids(_ptr(x), _ptr(out), ctypes.c_int64(x.size))
    return out
