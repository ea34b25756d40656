"""
MI355X-native differentiable rasterisation path of diffrp.

``diffrp_amd.nvdr`` is the drop-in for the ``nvdiffrast.torch`` operators diffrp
calls (rasterize / DepthPeeler / interpolate / antialias); the HIP kernels
behind it live in ``csrc/`` and are reached through the C-ABI of
``lib/libdrp_hip.so`` (``include/drp.h``).
"""
__version__ = '0.1.0'
