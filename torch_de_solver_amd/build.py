"""Build of the native library ``libtdx.so`` (gfx950) from ``csrc/``.

The library is built in-tree with hipcc; nothing is compiled at import time.
Every HIP source is compiled with floating-point contraction off (the
stencil and residual arithmetic must keep the operation order written in the
kernels) and with packed FP32 instructions disabled.  The library links
hipRTC, which compiles the 'mat' kernels generated per program (same
options).
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
LIB = os.path.join(HERE, 'libtdx.so')
ARCH = os.environ.get('TDX_OFFLOAD_ARCH', 'gfx950')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ROCM_LIB = '/opt/rocm/lib'
FLAGS = ['-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
         '-ffp-contract=off',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-Wno-unused-result', '-I', INCLUDE, '-I', CSRC]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src: str, build_dir: str) -> str:
    obj = os.path.join(build_dir, os.path.basename(src) + '.o')
    cmd = [HIPCC, *FLAGS, '-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-6000:]}")
    return obj


def build(verbose: bool = True) -> str:
    build_dir = os.path.join(HERE, '_build')
    os.makedirs(build_dir, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir), srcs))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', f'--offload-arch={ARCH}', '-o', tmp, *objs,
           '-L', ROCM_LIB, '-lhiprtc', f'-Wl,-rpath,{ROCM_LIB}']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == '__main__':
    build()
    sys.exit(0)
