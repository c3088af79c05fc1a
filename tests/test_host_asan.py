"""Host code of the native library under AddressSanitizer (CPU only): plan
creation, table validation (with fuzzed tables), work lists, the 'mat'
kernel generator and the division check, from a host-only instrumented
build (tests/asan/build_asan.py) driven by tests/asan/drive_plans.py in a
subprocess with the ASan runtime preloaded.  Any invalid read or write
aborts the subprocess with a report."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, 'asan'))

import build_asan  # noqa: E402


def test_host_code_under_address_sanitizer():
    try:
        lib = build_asan.build()
        rt = build_asan.asan_runtime()
    except (RuntimeError, FileNotFoundError, subprocess.CalledProcessError) as exc:
        pytest.skip(f"no host ASan build here: {exc}")
    env = dict(os.environ, LD_PRELOAD=rt, TDX_LIB_PATH=lib,
               ASAN_OPTIONS='detect_leaks=0:abort_on_error=1:halt_on_error=1',
               HIP_VISIBLE_DEVICES='', CUDA_VISIBLE_DEVICES='')
    r = subprocess.run([sys.executable, os.path.join(HERE, 'asan', 'drive_plans.py')], env=env,
                       capture_output=True, text=True, timeout=900)
    assert r.returncode == 0 and 'drive ok' in r.stdout, (r.stdout[-2000:], r.stderr[-6000:])
