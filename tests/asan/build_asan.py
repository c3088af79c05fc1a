"""Host-only AddressSanitizer build of the native library (CPU test
infrastructure for tests/test_host_asan.py).

Every source of torch_de_solver_amd/csrc is compiled as for the product
(gfx950 device code embedded, so the library loads and registers its
kernels), with AddressSanitizer on the host code only, into
tests/asan/_build/libtdx_asan.so.  Plan creation, table validation, work
lists, the 'mat' kernel generator and the division check all run on the
host, so a process without a GPU exercises them; device uploads then fail
with an error code, which exercises the failure paths too.  The library
expects the ASan runtime to be preloaded (ASAN_RUNTIME below).
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
CSRC = os.path.join(ROOT, 'torch_de_solver_amd', 'csrc')
INCLUDE = os.path.join(ROOT, 'include')
BUILD = os.path.join(HERE, '_build')
LIB = os.path.join(BUILD, 'libtdx_asan.so')
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ROCM_LIB = '/opt/rocm/lib'
sys.path.insert(0, ROOT)
from torch_de_solver_amd import build as product_build  # noqa: E402

FLAGS = ['--offload-arch=gfx950', '-O1', '-g', '-std=c++17', '-fPIC', '-fno-omit-frame-pointer',
         '-Xarch_host', '-fsanitize=address', '-ffp-contract=off', '-Wno-unused-result',
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I', INCLUDE, '-I', CSRC, '-I', product_build.GEN]


def asan_runtime() -> str:
    r = subprocess.run(['/opt/rocm/llvm/bin/clang', '-print-file-name=libclang_rt.asan-x86_64.so'],
                       capture_output=True, text=True, check=True)
    return r.stdout.strip()


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


STAMP = os.path.join(BUILD, 'flags.txt')


def up_to_date() -> bool:
    if not os.path.exists(LIB) or not os.path.exists(STAMP):
        return False
    if open(STAMP).read() != ' '.join(FLAGS):
        return False
    t = os.path.getmtime(LIB)
    deps = sources() + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.h')]
    deps.append(os.path.join(INCLUDE, 'tdx.h'))
    return all(os.path.getmtime(f) <= t for f in deps)


def build() -> str:
    # the generated include of the embedded device headers (as the product build)
    product_build.write_embed()
    if up_to_date():
        return LIB
    os.makedirs(BUILD, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(BUILD, os.path.basename(src) + '.o')
        r = subprocess.run([HIPCC, *FLAGS, '-c', src, '-o', obj], capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
        return obj

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=4) as ex:
        objs = list(ex.map(compile_one, sources()))
    tmp = LIB + '.tmp'
    r = subprocess.run([HIPCC, '--offload-arch=gfx950', '-shared', '-shared-libasan',
                        '-Xarch_host', '-fsanitize=address', '-o', tmp, *objs, '-L', ROCM_LIB,
                        '-lhiprtc', f'-Wl,-rpath,{ROCM_LIB}'], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    os.replace(tmp, LIB)
    with open(STAMP, 'w') as f:
        f.write(' '.join(FLAGS))
    return LIB


if __name__ == '__main__':
    print(build())
    sys.exit(0)
