"""The no-allocation entry point tdx_nn_train_step_noalloc against
tdx_nn_train_step with a caller workspace: the same kernels on the same
state give the same bits, on two calls in a row, the second on a smaller
problem (the scratch array keeps the first call's contents, which the
second call must not depend on); plans larger than the scratch array are
refused with an error before anything runs."""
import sys

import numpy as np
import pytest
import torch

import problems
from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd import native  # noqa: E402
from torch_de_solver_amd import workloads as W  # noqa: E402
from torch_de_solver_amd.compiler import compile_nn  # noqa: E402
from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch  # noqa: E402

DEV = torch.device('cuda')


def _engine(nx, nt, layers):
    d, eq, b, kw = problems.burgers_nn(api, nx=nx, nt=nt)
    net = problems.randomize_biases(problems.make_mlp(layers))
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                      lambda_operator=1, lambda_bound=10)
    eng = NNEngine([prob], mlp_arch(net), DEV)
    flat = torch.cat([p.detach().reshape(-1) for p in net.parameters()]).to(DEV)
    return eng, flat


def _lib():
    return native.lib()


def _step(eng, state, noalloc):
    p, m, v, step, loss = state
    lib = _lib()
    args = (eng._plan, native.ptr(p), native.ptr(m), native.ptr(v), native.ptr(step),
            1e-3, 0.9, 0.999, 1e-8, native.ptr(loss))
    if noalloc:
        rc = lib.tdx_nn_train_step_noalloc(*args, native.stream_ptr())
    else:
        rc = lib.tdx_nn_train_step(*args, native.ptr(eng.workspace), native.stream_ptr())
    native.check(rc)


def _state(flat):
    return (flat.clone(), torch.zeros_like(flat), torch.zeros_like(flat),
            torch.zeros(1, dtype=torch.int64, device=DEV),
            torch.zeros(2, dtype=flat.dtype, device=DEV))


def test_noalloc_step_matches_workspace_step():
    lib = _lib()
    cap = int(lib.tdx_noalloc_scratch_bytes())
    # first call: the larger problem, then a smaller one
    for nx, nt, layers in ((40, 30, [2, 32, 32, 32, 1]), (12, 7, [2, 12, 12, 1])):
        eng, flat = _engine(nx, nt, layers)
        need = int(lib.tdx_nn_workspace_bytes(eng._plan))
        print(f"[noalloc] burgers {nx}x{nt} net {layers}: workspace {need} of {cap} bytes",
              flush=True)
        assert need <= cap
        a, b = _state(flat), _state(flat)
        for _ in range(2):
            _step(eng, a, noalloc=False)
            _step(eng, b, noalloc=True)
        torch.cuda.synchronize()
        for x, y in zip(a, b):
            assert torch.equal(x, y)
        assert int(b[3]) == 2 and torch.isfinite(b[4]).all()


def test_noalloc_refuses_plans_beyond_the_scratch_size():
    lib = _lib()
    cap = int(lib.tdx_noalloc_scratch_bytes())
    insts = W.sweep_instances(128)
    probs, nets = [], []
    for inst in insts:
        d, eq, b, net, ckw = W.sweep_problem(inst)
        probs.append(compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                                h=ckw['h'], lambda_operator=1, lambda_bound=10))
        nets.append(net)
    eng = NNEngine(probs, mlp_arch(nets[0]), DEV)
    need = int(lib.tdx_nn_workspace_bytes(eng._plan))
    print(f"[noalloc] 128-instance sweep: workspace {need} of {cap} bytes", flush=True)
    assert need > cap
    flat = torch.zeros(eng.n_inst * eng.n_params, dtype=torch.float32, device=DEV)
    st = _state(flat)
    rc = lib.tdx_nn_train_step_noalloc(eng._plan, native.ptr(st[0]), native.ptr(st[1]),
                                       native.ptr(st[2]), native.ptr(st[3]), 1e-3, 0.9, 0.999,
                                       1e-8, native.ptr(st[4]), native.stream_ptr())
    assert rc == -1
    assert b'no-allocation' in lib.tdx_last_error()
    torch.cuda.synchronize()
    assert int(st[3]) == 0
