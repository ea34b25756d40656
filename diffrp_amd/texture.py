"""
Texture sampling on the HIP kernels: diffrp's ``sample2d`` / ``sample3d``
(``diffrp/utils/shader_ops.py:198-310``) and glTF samplers
(``diffrp/materials/gltf_material.py:9-22``).

Same signatures and conventions as diffrp: textures are channels-last
([H, W, C] / [D, H, W, C]), coordinates (..., 2) / (..., 3) with (0, 0) the
bottom-left corner; wrap 'border', 'reflection', 'cyclic' (2-D) or 'latlong';
mode 'bilinear', 'nearest' or (2-D) 'bicubic' -- grid_sample's modes.
Differentiable w.r.t. the texture and the coordinates (nearest has no
coordinate gradient, as in grid_sample).
"""
from dataclasses import dataclass

import torch

from . import _lib
from ._lib import check, ptr, require_device, stream_of

_WRAP = {'border': 0, 'reflection': 1, 'cyclic': 2, 'latlong': 3}
_MODE = {'bilinear': 0, 'nearest': 1, 'bicubic': 2}


def _codes(wrap, mode, dims):
    if wrap not in _WRAP or (dims == 3 and wrap == 'cyclic'):
        raise ValueError(f'unsupported wrap {wrap!r} for sample{dims}d')
    if mode not in _MODE or (dims == 3 and mode == 'bicubic'):
        raise ValueError(f'unsupported mode {mode!r} for sample{dims}d')
    return _WRAP[wrap], _MODE[mode]


class _SampleFunc(torch.autograd.Function):
    @staticmethod
    def forward(fctx, tex, coords, wrap, mode):
        dims = coords.shape[-1]
        tex_c = tex.float().contiguous()
        tc = coords.float().contiguous()
        n = tc.numel() // dims
        C = tex_c.shape[-1]
        out = torch.empty(n, C, dtype=torch.float32, device=tex.device)
        L = _lib.lib()
        if dims == 2:
            H, W = tex_c.shape[0], tex_c.shape[1]
            rc = L.drp_sample2d_fwd(stream_of(tex_c), ptr(tex_c), H, W, C, ptr(tc), n, wrap, mode, ptr(out))
        else:
            D, H, W = tex_c.shape[0], tex_c.shape[1], tex_c.shape[2]
            rc = L.drp_sample3d_fwd(stream_of(tex_c), ptr(tex_c), D, H, W, C, ptr(tc), n, wrap, mode, ptr(out))
        check(rc, f'sample{dims}d')
        fctx.save_for_backward(tex_c, tc)
        fctx.args = (wrap, mode, dims)
        return out

    @staticmethod
    def backward(fctx, gout):
        tex_c, tc = fctx.saved_tensors
        wrap, mode, dims = fctx.args
        need_tex, need_tc = fctx.needs_input_grad[0], fctx.needs_input_grad[1]
        if not (need_tex or need_tc):
            return None, None, None, None
        g = gout.float().contiguous()
        gtex = torch.zeros_like(tex_c) if need_tex else None
        gtc = torch.empty_like(tc) if need_tc else None
        n = tc.numel() // dims
        C = tex_c.shape[-1]
        L = _lib.lib()
        if dims == 2:
            rc = L.drp_sample2d_bwd(stream_of(tex_c), ptr(tex_c), tex_c.shape[0], tex_c.shape[1], C, ptr(tc), n, wrap,
                                    mode, ptr(g), ptr(gtex), ptr(gtc))
        else:
            rc = L.drp_sample3d_bwd(stream_of(tex_c), ptr(tex_c), tex_c.shape[0], tex_c.shape[1], tex_c.shape[2], C,
                                    ptr(tc), n, wrap, mode, ptr(g), ptr(gtex), ptr(gtc))
        check(rc, f'sample{dims}d backward')
        return gtex, gtc, None, None


def _sample(tex, coords, wrap, mode, dims):
    if tex.dim() != dims + 1:
        raise ValueError(f'texture must be {"[H, W, C]" if dims == 2 else "[D, H, W, C]"} (got {tuple(tex.shape)})')
    if coords.shape[-1] != dims:
        raise ValueError(f'coordinates must end in {dims} (got {tuple(coords.shape)})')
    require_device(tex, coords, names=('texture', 'texcoords'))
    w, m = _codes(wrap, mode, dims)
    out = _SampleFunc.apply(tex, coords, w, m)
    return out.reshape(*coords.shape[:-1], tex.shape[-1])


def sample2d(texture2d: torch.Tensor, texcoords: torch.Tensor, wrap: str = 'border', mode: str = 'bilinear'):
    """Sample an [H, W, C] texture at (..., 2) coordinates -> (..., C)."""
    return _sample(texture2d, texcoords, wrap, mode, 2)


def sample3d(texture3d: torch.Tensor, texcoords: torch.Tensor, wrap: str = 'border', mode: str = 'bilinear'):
    """Sample a [D, H, W, C] texture at (..., 3) coordinates -> (..., C)."""
    return _sample(texture3d, texcoords, wrap, mode, 3)


@dataclass
class GLTFSampler:
    """glTF texture sampler (``gltf_material.py:9-22``): 'repeat' wraps uv into
    [0, 1) first, 'clamp' samples with border padding, 'mirror' with
    reflection; 'point' / 'linear' interpolation."""
    image: torch.Tensor
    wrap_mode: str = 'repeat'
    interpolation: str = 'linear'

    def sample(self, uv: torch.Tensor) -> torch.Tensor:
        if self.wrap_mode == 'repeat':
            uv = uv.remainder(1.0)
        return sample2d(self.image, uv, wrap='border' if self.wrap_mode == 'clamp' else 'reflection',
                        mode='bilinear' if self.interpolation == 'linear' else 'nearest')
