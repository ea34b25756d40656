"""Write tests/golden/procedural_albedo_512.npz from the reference's documented output.

Source: docs/source/assets/procedural-albedo.png of diffrp 0.2.7: the
procedural material of docs/source/writing_a_material.md:22-31 on the README
scene (icosphere r=0.8, default camera), albedo_srgb through to_pil, 512x512
(its coverage mask is bit-identical to the direct 512^2 README render).
Stored: the RGBA image flipped to bottom-row-first, uint8.  Run in the build
container, where the reference is mounted.
"""
import os
import sys

import numpy as np
from PIL import Image

SRC = '/root/reference/docs/source/assets/procedural-albedo.png'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'procedural_albedo_512.npz')

if __name__ == '__main__':
    img = np.array(Image.open(SRC))
    assert img.shape == (512, 512, 4) and img.dtype == np.uint8
    np.savez_compressed(OUT, rgba_bottom_first=img[::-1].copy())
    print(OUT, os.path.getsize(OUT), 'bytes; covered px', int((img[..., 3] > 0).sum()))
    sys.exit(0)
