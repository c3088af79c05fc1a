"""Drives the host code of the native library for tests/test_host_asan.py
(run in a subprocess with the AddressSanitizer runtime preloaded and
TDX_LIB_PATH pointing at tests/asan/_build/libtdx_asan.so; no GPU needed).

For every test problem of every mode it lowers the tables, creates the plan
(validation, work lists and layout run; the device upload then fails with an
error code since there is no device, which exercises the cleanup path),
fuzzes single table entries (each must give an error code or pass
validation, never an out-of-bounds access), runs the 'mat' kernel generator
and the division check.  Prints "drive ok" at the end."""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, 'tests')]

import torch  # noqa: E402

import problems  # noqa: E402
from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd import native  # noqa: E402
from torch_de_solver_amd.ad_engine import ADTables  # noqa: E402
from torch_de_solver_amd.compiler import compile_ad, compile_mat, compile_nn  # noqa: E402
from torch_de_solver_amd.mat_engine import MatTables  # noqa: E402
from torch_de_solver_amd.nn_engine import MLPArch, NNTables  # noqa: E402

OK_CODES = {0, -1, -2, -3, -4}


def create(fn, desc, destroy):
    h = ctypes.c_void_p()
    rc = fn(ctypes.byref(desc), ctypes.byref(h))
    assert rc in OK_CODES, rc
    if rc == 0:
        destroy(h)
    return rc


def fuzz(arr, make_desc, fn, destroy, rng, n):
    """Single-entry mutations of an int32 table: every one must come back
    as an error code or a valid plan."""
    codes = []
    for _ in range(n):
        i = int(rng.integers(arr.size))
        old = int(arr[i])
        arr[i] = int(rng.choice([-1, 0, 1, old + 1, old - 1, old * 2 + 7, 1 << 20, -(1 << 20)]))
        codes.append(create(fn, make_desc(), destroy))
        arr[i] = old
    return codes


def main():
    lib = native.lib()
    assert 'asan' in native.LIB_PATH, native.LIB_PATH
    rng = np.random.default_rng(0)
    n_plans = 0
    # 'NN'
    for name, (builder, layers) in problems.NN_CASES.items():
        for dtype in ('float32', 'float64'):
            d, eq, b, kw = builder(api, dtype=dtype)
            prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                              h=kw.get('h', 1e-3), lambda_operator=kw['lambda_operator'],
                              lambda_bound=kw['lambda_bound'])
            arch = MLPArch(d_in=layers[0], width=layers[1], n_out=layers[-1],
                           n_layers=len(layers) - 1)
            npd = np.float32 if dtype == 'float32' else np.float64
            for probs in ([prob], [prob, prob, prob]):
                t = NNTables(probs, arch, npd)
                for tile_rows, chunk in ((0, 0), (-1, 0), (32, 3)):
                    create(lib.tdx_nn_plan_create, t.desc(arch, tile_rows, chunk),
                           lib.tdx_nn_plan_destroy)
                    n_plans += 1
            t = NNTables([prob], arch, npd)
            fuzz(t.ints, lambda: t.desc(arch), lib.tdx_nn_plan_create, lib.tdx_nn_plan_destroy,
                 rng, 150)
            fuzz(t.groups, lambda: t.desc(arch), lib.tdx_nn_plan_create,
                 lib.tdx_nn_plan_destroy, rng, 40)
    # 'mat'
    for name, builder in problems.MAT_CASES.items():
        for dtype in ('float32', 'float64'):
            d, eq, b, kw = builder(api, dtype=dtype)
            prob = compile_mat(d.build('mat'), eq.equation_lst, b.build(d.variable_dict),
                               n_eq=len(eq.equation_lst),
                               derivative_points=kw.get('derivative_points', 2),
                               lambda_operator=kw['lambda_operator'],
                               lambda_bound=kw['lambda_bound'])
            t = MatTables(prob, np.float32 if dtype == 'float32' else np.float64)
            for ex in (0, 1, 2):
                create(lib.tdx_mat_plan_create, t.desc(ex), lib.tdx_mat_plan_destroy)
                n_plans += 1
            fuzz(t.ints, lambda: t.desc(1), lib.tdx_mat_plan_create, lib.tdx_mat_plan_destroy,
                 rng, 150)
    # the generator of the 'mat' interior kernels (host code only)
    for name, (builder, nx, nt, dp, _) in problems.SPECIALIZED_CASES.items():
        d, eq, b, dp, n_eq = builder(api, nx, nt, dp)
        prob = compile_mat(d.build('mat'), eq.equation_lst, b.build(d.variable_dict), n_eq=n_eq,
                           derivative_points=dp, lambda_operator=1, lambda_bound=10)
        src = MatTables(prob, np.float32).specialized_source(1, compile=False)
        assert 'tdx_mat_jit_step' in src, name
    # the generator of the fused 'NN' kernel's residual programs (host code
    # only), on valid tables and on single-entry mutations of them
    n_fused = 0
    for name, (builder, layers) in problems.NN_CASES.items():
        d, eq, b, kw = builder(api, dtype='float32')
        prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                          h=kw.get('h', 1e-3), lambda_operator=kw['lambda_operator'],
                          lambda_bound=kw['lambda_bound'])
        arch = MLPArch(d_in=layers[0], width=layers[1], n_out=layers[-1],
                       n_layers=len(layers) - 1)
        t = NNTables([prob, prob], arch, np.float32)
        src = t.fused_source(arch, compile=False)
        assert src == '' or 'tdx_fused_jit' in src, name
        n_fused += 1 if src else 0
        for _ in range(60):
            i = int(rng.integers(t.ints.size))
            old = int(t.ints[i])
            t.ints[i] = int(rng.choice([-1, 0, 1, old + 1, old - 1, 1 << 20]))
            desc = t.desc(arch)
            rc = lib.tdx_nn_fused_source(ctypes.byref(desc), None, 0, 0)
            assert rc >= 0 or rc in (-1, -2, -3, -4, -5), rc
            t.ints[i] = old
    assert n_fused >= 1
    # 'autograd'
    for name, (builder, layers) in problems.AD_CASES.items():
        d, eq, b, kw = builder(api, dtype='float64')
        prob = compile_ad(d.build('autograd'), eq.equation_lst, b.build(d.variable_dict),
                          n_out=layers[-1], lambda_operator=kw['lambda_operator'],
                          lambda_bound=kw['lambda_bound'])
        arch = MLPArch(d_in=layers[0], width=layers[1], n_out=layers[-1], n_layers=len(layers) - 1)
        t = ADTables(prob, np.float64)
        create(lib.tdx_ad_plan_create, t.desc(prob, arch), lib.tdx_ad_plan_destroy)
        n_plans += 1
        fuzz(t.ints, lambda: t.desc(prob, arch), lib.tdx_ad_plan_create,
             lib.tdx_ad_plan_destroy, rng, 100)
    # the division check
    for h in (1e-3, 1 / 127, 0.01):
        assert lib.tdx_quotient_check(ctypes.c_float(h), 0) >= 0
    assert lib.tdx_nn_plan_create(None, None) == -1
    print(f"drive ok: {n_plans} plans", flush=True)


if __name__ == '__main__':
    main()
