"""
Build of the HIP library ``diffrp_amd/lib/libdrp_hip.so`` for gfx950.

hipcc compiles each ``csrc/*.hip`` file to an object and links one shared
library with a plain C ABI (``include/drp.h``).  Flags:
  -ffp-contract=off      the float operation sequence is the one written
                         (bit-exact ids / depth against a scalar CPU restatement)
  -packed-fp32-ops off   no v_pk_*_f32 instructions are generated
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'lib')
LIB_PATH = os.path.join(OUT_DIR, 'libdrp_hip.so')
ARCH = os.environ.get('DRP_OFFLOAD_ARCH', 'gfx950')

SOURCES = ['drp_raster.hip', 'drp_tribwd.hip', 'drp_interp.hip', 'drp_aa.hip', 'drp_mesh.hip', 'drp_edge.hip',
           'drp_gbuffer.hip', 'drp_texture.hip', 'drp_loss.hip', 'drp_clip.hip', 'drp_fork.hip', 'drp_capi.hip']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
CFLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    '-ffp-contract=off', '-fno-fast-math',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-munsafe-fp-atomics',
    '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
]


def _compile(src: str) -> str:
    obj = os.path.join(OUT_DIR, os.path.splitext(src)[0] + '.o')
    cmd = [HIPCC] + CFLAGS + ['-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    return obj


def build(verbose: bool = False) -> str:
    """Compile every HIP source and link ``libdrp_hip.so`` (always rebuilds)."""
    os.makedirs(OUT_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB_PATH + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', f'--offload-arch={ARCH}', '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{" ".join(cmd)}\n{r.stdout}\n{r.stderr}')
    os.replace(tmp, LIB_PATH)
    for o in objs:
        os.remove(o)
    if verbose:
        print('built', LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
