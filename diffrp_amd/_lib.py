"""
ctypes binding of ``libdrp_hip.so`` (the C-ABI declared in ``include/drp.h``).

The library is built in-tree by ``diffrp_amd.build_ext`` (``__graft_entry__.build()``).
There is no fallback: if the library is missing or a tensor is not on a ROCm
device, the call raises.
"""
import ctypes
import os
import threading

import torch

from .build_ext import LIB_PATH

_lib = None
_lock = threading.Lock()

c_void_p = ctypes.c_void_p
c_int = ctypes.c_int
c_ll = ctypes.c_longlong
c_size = ctypes.c_size_t
c_float = ctypes.c_float

_SIGS = {
    'drp_last_error': (ctypes.c_char_p, []),
    'drp_abi_version': (c_int, []),
    'drp_rasterize_workspace_size': (c_size, [c_int, c_int, c_int, c_ll, c_int, c_int, c_int]),
    'drp_rasterize_fwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_size, c_int, c_int, c_int, c_ll, c_int,
                                  c_void_p]),
    'drp_rasterize_fwd_noalloc': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int,
                                          c_int, c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_ll, c_int,
                                          c_void_p]),
    'drp_rasterize_bwd_workspace_size': (c_size, [c_int, c_int, c_int, c_int]),
    'drp_rasterize_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_void_p, c_int, c_int,
                                  c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  c_size]),
    'drp_batch_plan_size': (c_size, [c_void_p, c_int, c_int, c_int]),
    'drp_batch_plan': (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_size]),
    'drp_batch_plan_info': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
    'drp_rasterize_batch_workspace_size': (c_size, [c_void_p]),
    'drp_rasterize_batch_fwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_size, c_void_p, c_int]),
    'drp_batch_interpolate_fwd': (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p]),
    'drp_batch_interpolate_bwd_workspace_size': (c_size, [c_void_p, c_int]),
    'drp_batch_interpolate_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p,
                                          c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                          c_size]),
    'drp_batch_interpolate_rasterize_bwd_workspace_size': (c_size, [c_void_p, c_int]),
    'drp_batch_interpolate_rasterize_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                                                    c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                                    c_void_p, c_void_p, c_void_p, c_size]),
    'drp_batch_rasterize_bwd_workspace_size': (c_size, [c_void_p]),
    'drp_batch_rasterize_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        c_size]),
    'drp_batch_antialias_workspace_size': (c_size, [c_void_p]),
    'drp_batch_antialias_fwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_size]),
    'drp_batch_antialias_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                        c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int, c_void_p, c_size]),
    'drp_sample2d_fwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_ll, c_int, c_int, c_void_p]),
    'drp_sample2d_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_ll, c_int, c_int, c_void_p,
                                 c_void_p, c_void_p]),
    'drp_sample3d_fwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_ll, c_int, c_int,
                                 c_void_p]),
    'drp_sample3d_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_ll, c_int, c_int,
                                 c_void_p, c_void_p, c_void_p]),
    'drp_stencil_partition_workspace_size': (c_size, [c_int, c_int, c_int]),
    'drp_stencil_partition': (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_int, c_int, c_void_p, c_void_p,
                                      c_void_p, c_void_p, c_size]),
    'drp_count_covered': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p]),
    'drp_incidence_workspace_size': (c_size, [c_int, c_int]),
    'drp_incidence_build': (c_int, [c_void_p, c_void_p, c_int, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_interpolate_fwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_int,
                                    c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p]),
    'drp_interpolate_bwd_workspace_size': (c_size, [c_int, c_int, c_int, c_int, c_int, c_int, c_int]),
    'drp_interpolate_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_int, c_int,
                                    c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                    c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_interpolate_rasterize_bwd_workspace_size': (c_size, [c_int, c_int, c_int, c_int, c_int, c_int]),
    'drp_interpolate_rasterize_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_int,
                                              c_void_p, c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_edge_grad_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                  c_int, c_void_p]),
    'drp_topology_workspace_size': (c_size, [c_int]),
    'drp_topology_build': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_size]),
    'drp_antialias_workspace_size': (c_size, [c_int, c_int, c_int, c_int, c_int, c_int]),
    'drp_antialias_fwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                  c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_antialias_prepare': (c_int, [c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_int, c_void_p, c_int,
                                      c_void_p, c_void_p, c_size]),
    'drp_rasterize_fwd_aa_prepare': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_void_p, c_int, c_int, c_int,
                                             c_void_p, c_void_p, c_void_p, c_size, c_int, c_int, c_int, c_void_p,
                                             c_void_p, c_void_p, c_size]),
    'drp_antialias_fwd_prepared': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int,
                                           c_int, c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_antialias_bwd': (c_int, [c_void_p, c_void_p, c_int, c_int, c_int, c_int, c_void_p, c_void_p, c_int, c_int,
                                  c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p, c_float, c_int, c_void_p,
                                  c_size]),
    'drp_clip_transform_fwd': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    'drp_clip_transform_bwd': (c_int, [c_void_p, c_void_p, c_int, c_void_p, c_int, c_void_p]),
    'drp_l2_loss_workspace_size': (c_size, []),
    'drp_l2_loss_fwd': (c_int, [c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p, c_size]),
    'drp_l2_loss_bwd': (c_int, [c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p, c_void_p]),
    'drp_l2_loss_fwd_grad': (c_int, [c_void_p, c_void_p, c_void_p, c_ll, c_void_p, c_void_p, c_void_p, c_size]),
    'drp_l2_loss_scale': (c_int, [c_void_p, c_void_p, c_ll, c_void_p]),
}

EXPORTED = tuple(_SIGS)


def lib() -> ctypes.CDLL:
    """Load the HIP library (once).  Raises if it has not been built."""
    global _lib
    if _lib is None:
        with _lock:
            if _lib is None:
                if not os.path.exists(LIB_PATH):
                    raise RuntimeError(
                        f'{LIB_PATH} is missing: build it with `python -m diffrp_amd.build_ext` '
                        '(or __graft_entry__.build()); there is no CPU fallback')
                L = ctypes.CDLL(LIB_PATH)
                for name, (res, args) in _SIGS.items():
                    f = getattr(L, name)
                    f.restype = res
                    f.argtypes = args
                _lib = L
    return _lib


def check(rc: int, what: str):
    if rc != 0:
        msg = lib().drp_last_error()
        raise RuntimeError(f'{what} failed ({rc}): {msg.decode() if msg else "unknown error"}')


def ptr(t):
    return None if t is None else c_void_p(t.data_ptr())


def stream_of(t: torch.Tensor):
    return c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def require_device(*tensors, names=None):
    for i, t in enumerate(tensors):
        if t is None:
            continue
        if not t.is_cuda:
            n = names[i] if names else f'argument {i}'
            raise RuntimeError(f'{n} must be a ROCm device tensor (got {t.device}); the rasteriser has no CPU path')
