"""
The C-ABI library without a GPU: it loads, exports every entry point that
``include/drp.h`` declares with the parameter count the ctypes binding uses,
answers the host-only size queries, and rejects bad arguments with an error
status and message before touching a device (the reference's operators raise on
such input as well).  No kernel is launched here.
"""
import ctypes
import os
import re

import pytest
import torch

from diffrp_amd import _lib, build_ext

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'drp.h')


def _declarations():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    decls = {}
    for m in re.finditer(r'\b(?:int|size_t|const char \*)\s*(drp_\w+)\s*\(([^;{]*)\)\s*;', text):
        args = m.group(2).strip()
        decls[m.group(1)] = 0 if args in ('', 'void') else len(args.split(','))
    return decls


@pytest.fixture(scope='module')
def L():
    if not os.path.exists(build_ext.LIB_PATH):
        build_ext.build()
    return _lib.lib()


def test_header_and_binding_agree():
    decls = _declarations()
    assert len(decls) >= 17
    assert set(decls) == set(_lib.EXPORTED)
    for name, n in decls.items():
        assert len(_lib._SIGS[name][1]) == n, name


def test_library_exports_every_declared_symbol(L):
    for name in _declarations():
        assert hasattr(L, name), name
    m = re.search(r'#define DRP_ABI_VERSION (\d+)', open(HEADER).read())
    assert L.drp_abi_version() == int(m.group(1))


def test_workspace_queries(L):
    ws = L.drp_rasterize_workspace_size(64, 512, 512, 64 * 81920, 16, 8, 512)
    tiles = 64 * 32 * 64
    assert ws >= tiles * 512 * 4 + 64 * 81920 * 12
    assert ws % 256 == 0
    assert L.drp_rasterize_workspace_size(0, 512, 512, 0, 16, 8, 512) == 0
    # partials plus one record slot per pixel (triangle + sums)
    px = 64 * 512 * 512
    assert L.drp_rasterize_bwd_workspace_size(64, 512, 512, 81920) >= 64 * 81920 * 48 + px * (4 + 48)
    assert L.drp_interpolate_bwd_workspace_size(64, 512, 512, 81920, 8, 0, 8) >= 8 * 81920 * 96 + px * (4 + 96)
    assert L.drp_interpolate_bwd_workspace_size(64, 512, 512, 81920, 8, 0, 8) < L.drp_interpolate_bwd_workspace_size(
        64, 512, 512, 81920, 8, 0, 1)
    assert L.drp_rasterize_bwd_workspace_size(64, 0, 512, 81920) == 0
    assert L.drp_topology_workspace_size(81920) > 0
    assert L.drp_incidence_workspace_size(40962, 81920) > 0
    assert L.drp_antialias_workspace_size(1, 64, 64, 3, 100, 50) > 0


def _fwd(L, **kw):
    a = dict(stream=None, pos=None, n=1, V=0, pb=0, tri=None, F=0, ranges=None, H=16, W=16, prev=None,
             rast=None, rast_db=None, ws=None, ws_bytes=0, tw=8, th=8, cap=64, lt=0, mr=0, boxes=None)
    a.update(kw)
    return L.drp_rasterize_fwd(*a.values())


def test_argument_errors_are_reported(L):
    dummy = ctypes.c_void_p(4096)
    assert _fwd(L) < 0
    assert b'rast' in L.drp_last_error()
    assert _fwd(L, rast=dummy, tw=12, th=8) < 0
    assert b'tile' in L.drp_last_error()
    assert _fwd(L, rast=dummy, H=-1) < 0
    assert _fwd(L, rast=dummy, V=3, F=1) < 0          # pos / tri missing
    assert b'null' in L.drp_last_error()
    assert _fwd(L, rast=dummy, ws=None, ws_bytes=0) < 0
    assert b'workspace' in L.drp_last_error()
    assert _fwd(L, rast=ctypes.c_void_p(4100)) < 0    # misaligned output
    assert b'aligned' in L.drp_last_error()
    assert _fwd(L, n=0) == 0                          # empty batch: nothing to do
    assert L.drp_edge_grad_bwd(None, dummy, dummy, 16, 16, 2, dummy, dummy, dummy, 1, dummy) < 0
    assert b'channels' in L.drp_last_error()
    assert L.drp_interpolate_bwd(None, dummy, 1, 3, 0, 0, dummy, 16, 16, dummy, 1, None, None, 0, dummy, None,
                                 dummy, dummy, None, None, 1, None, None, None, 0) < 0
    assert b'channel' in L.drp_last_error()


def test_operators_reject_host_tensors(L):
    from diffrp_amd import nvdr
    ctx = nvdr.RasterizeCudaContext(device='cuda:0')
    pos = torch.tensor([[[-1., -1., 0., 1.], [1., -1., 0., 1.], [0., 1., 0., 1.]]])
    tri = torch.tensor([[0, 1, 2]], dtype=torch.int32)
    with pytest.raises(RuntimeError, match='CPU path|ROCm device'):
        nvdr.rasterize(ctx, pos, tri, (8, 8))
    with pytest.raises(RuntimeError, match='CPU path|ROCm device'):
        nvdr.interpolate(torch.zeros(3, 2), torch.zeros(1, 8, 8, 4), tri)
    with pytest.raises(RuntimeError):
        nvdr.RasterizeGLContext()
