"""The fused 'autograd' training step (k_ad_step: forward jets, residual
program, adjoint and weight gradients of a tile of whole entries in one
kernel) against the float32 oracle and against the layer kernels.

One train step with lr = 0 and beta1 = 0 leaves the parameters as they
are and writes the gradient itself into the first moment (m = g), and the
loss pair of the parameters into the loss buffer.

Tolerances (float32):
* against the float32 oracle (nested torch autograd on the CPU): loss to
  1e-5 relative, gradient to 1e-5 in norm -- the bound the layer kernels
  meet (tests/test_autograd_gpu.py);
* against the layer kernels: loss to 1e-6 relative, gradient to 1e-5 in
  norm (the same arithmetic per row, summed in other orders);
* two fused steps from the same state: bit for bit.
"""
import os
import sys

import numpy as np
import pytest
import torch

import problems
from conftest import gpu_available
from oracle import tedeous_oracle as O

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

import torch_de_solver_amd as tdx  # noqa: E402
from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd import workloads  # noqa: E402
from torch_de_solver_amd.nn_engine import flatten_module_params  # noqa: E402
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _net(layers, seed=0):
    net = problems.make_mlp(layers, seed=seed, dtype=torch.float64)
    problems.randomize_biases(net, scale=0.3)
    return net


def _solution(d, eq, b, kw, net_dev, fused, jit=None):
    """A Solution whose plan is created with TDX_AD_FUSED = fused and, if
    given, TDX_AD_JIT = jit (generated jet maps and residual programs)."""
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    env = {'TDX_AD_FUSED': '1' if fused else '0'}
    if jit is not None:
        env['TDX_AD_JIT'] = str(int(jit))
    old = {k: os.environ.get(k) for k in env}
    os.environ.update(env)
    try:
        sol = Solution(grid, pre, net_dev, 'autograd', None, kw['lambda_operator'],
                       kw['lambda_bound'])
    finally:
        for k, v in old.items():
            if v is None:
                del os.environ[k]
            else:
                os.environ[k] = v
    return sol


def _step_gradient(sol, net_dev):
    """(loss pair, gradient) of one train step at lr 0, beta1 0."""
    eng = sol.engine
    params = flatten_module_params(net_dev)
    before = params.clone()
    m = torch.zeros_like(params)
    v = torch.zeros_like(params)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss = torch.zeros(2, dtype=params.dtype, device=DEV)
    eng.train_step(params, m, v, step, 0.0, 0.0, 0.999, 1e-8, loss)
    torch.cuda.synchronize()
    assert torch.equal(params, before)
    assert int(step.item()) == 1
    return loss.cpu().numpy().astype(np.float64), m.cpu().numpy().astype(np.float64)


def _case(name):
    builder, layers = problems.AD_CASES[name]
    d, eq, b, kw = builder(api, dtype='float32')
    return d, eq, b, kw, layers


def _oracle(d, eq, b, kw, net64):
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    net = net64.to(torch.float32)
    loss, norm, *_ = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    return float(loss), float(norm), grad.astype(np.float64)


@pytest.mark.parametrize('jit', [0, 1])
@pytest.mark.parametrize('name', list(problems.AD_CASES))
def test_ad_step_against_float32_oracle(name, jit):
    """Interpreted (jit 0) and generated (jit 1) jet maps and residual
    programs."""
    _print(f"[ad step] {name} jit {jit}")
    d, eq, b, kw, layers = _case(name)
    net_dev = _net(layers).to(torch.float32).to(DEV)
    sol = _solution(d, eq, b, kw, net_dev, fused=True, jit=jit)
    assert sol.engine.step_kernel == ('fused-generated' if jit else 'fused'), name
    loss, grad = _step_gradient(sol, net_dev)
    rl, rn, rg = _oracle(d, eq, b, kw, _net(layers))
    rel = abs(loss[0] - rl) / rl
    reln = abs(loss[1] - rn) / rn
    gn = np.linalg.norm(grad - rg) / np.linalg.norm(rg)
    _print(f"  loss rel {rel:.2e}, normalized {reln:.2e}, gradient {gn:.2e}")
    assert rel < 1e-5 and reln < 1e-5, (loss, rl, rn)
    assert gn < 1e-5, gn


@pytest.mark.parametrize('name', list(problems.AD_CASES))
def test_ad_step_against_layer_kernels(name):
    _print(f"[ad step vs layers] {name}")
    d, eq, b, kw, layers = _case(name)
    res = []
    for fused in (True, False):
        net_dev = _net(layers).to(torch.float32).to(DEV)
        sol = _solution(d, eq, b, kw, net_dev, fused=fused, jit=0)
        assert sol.engine.step_kernel == ('fused' if fused else 'layers')
        res.append(_step_gradient(sol, net_dev))
    (lf, gf), (ll, gl) = res
    gn = np.linalg.norm(gf - gl) / np.linalg.norm(gl)
    _print(f"  loss {lf[0]:.9g} / {ll[0]:.9g}, gradient {gn:.2e}")
    np.testing.assert_allclose(lf, ll, rtol=1e-6)
    assert gn < 1e-5, gn


@pytest.mark.parametrize('grid_res', [20, 200])
def test_ad_step_poisson(grid_res):
    """The bench workload's problem and network (2-100x5-1: 7 column blocks,
    four H x H layers, three points of five jets per wavefront), at 21^2 and
    at the full 201^2 points (where the plan compiles the generated jet
    maps and residual programs by default): fused against the layer
    kernels, and two fused steps from the same state bit for bit."""
    _print(f"[ad step] poisson {grid_res}")
    res = []
    for fused in (True, False):
        d, eq, b, net, kw = workloads.poisson_ad(grid_res)
        net_dev = net.to(torch.float32).to(DEV)
        sol = _solution(d, eq, b, kw, net_dev, fused=fused)
        expect = ('fused-generated' if grid_res == 200 else 'fused') if fused else 'layers'
        assert sol.engine.step_kernel == expect
        res.append(_step_gradient(sol, net_dev))
        if fused:
            again = _step_gradient(sol, net_dev)
            assert np.array_equal(res[0][0], again[0]) and np.array_equal(res[0][1], again[1])
    (lf, gf), (ll, gl) = res
    gn = np.linalg.norm(gf - gl) / np.linalg.norm(gl)
    _print(f"  loss {lf[0]:.9g} / {ll[0]:.9g}, gradient {gn:.2e}")
    np.testing.assert_allclose(lf, ll, rtol=1e-6)
    assert gn < 1e-5, gn


def test_ad_step_poisson_small_against_oracle():
    """Poisson at 21^2 points on the 2-100x5-1 network, generated kernel,
    against the float32 oracle."""
    _print("[ad step] poisson 20 vs oracle")
    d, eq, b, net, kw = workloads.poisson_ad(20)
    net64 = net.to(torch.float64)
    net_dev = workloads.poisson_ad(20)[3].to(torch.float32).to(DEV)
    sol = _solution(d, eq, b, kw, net_dev, fused=True, jit=1)
    assert sol.engine.step_kernel == 'fused-generated'
    loss, grad = _step_gradient(sol, net_dev)
    rl, rn, rg = _oracle(d, eq, b, kw, net64)
    rel = abs(loss[0] - rl) / rl
    gn = np.linalg.norm(grad - rg) / np.linalg.norm(rg)
    _print(f"  loss rel {rel:.2e}, gradient {gn:.2e}")
    assert rel < 1e-5, (loss, rl)
    assert gn < 1e-5, gn


def test_ad_step_model_train():
    """Model.train in float32 through the graph-captured fused step against
    torch's Adam over the native autograd Function (layer kernels): the
    same parameters after 21 epochs to 1e-5."""
    _print("[ad step model] wave_ad float32")
    res = []
    for fused in (True, False):
        d, eq, b, kw = problems.wave_ad(api, dtype='float32')
        net = _net([2, 12, 12, 1]).to(torch.float32)
        model = tdx.Model(net, d, eq, b)
        model.verbose = False
        model.compile('autograd', lambda_operator=1, lambda_bound=100)
        if fused:
            assert model.solution_cls.engine.step_kernel in ('fused', 'fused-generated')
        model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 21, callbacks=None if fused else [object()])
        res.append(torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu())
    np.testing.assert_allclose(res[0].numpy(), res[1].numpy(), rtol=1e-5, atol=1e-6)
