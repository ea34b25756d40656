"""Write tests/golden/readme_normals_512.npz from the reference's documented output.

Source: docs/source/assets/procedural-normals.png of diffrp 0.2.7, the direct
512x512 output of the README example (README.md:53-63: icosphere r=0.8, default
camera, DefaultMaterial, flat normals, false_color_camera_space_normal, to_pil).
Stored: the RGBA image flipped to bottom-row-first (the rasteriser's
orientation), uint8.  Run in the build container, where the reference is mounted.
"""
import os
import sys

import numpy as np
from PIL import Image

SRC = '/root/reference/docs/source/assets/procedural-normals.png'
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'readme_normals_512.npz')

if __name__ == '__main__':
    img = np.array(Image.open(SRC))
    assert img.shape == (512, 512, 4) and img.dtype == np.uint8
    np.savez_compressed(OUT, rgba_bottom_first=img[::-1].copy())
    print(OUT, os.path.getsize(OUT), 'bytes; covered px', int((img[..., 3] > 0).sum()))
    sys.exit(0)
