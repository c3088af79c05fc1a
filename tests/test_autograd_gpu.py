"""GPU parity of the 'autograd' mode engine (jets through the network) with
the CPU oracle's nested autograd (oracle/tedeous_oracle.py: ADOracle).

Tolerances:
* float64: loss, residuals, boundary values to 1e-9 relative, parameter
  gradient to 1e-8 relative (the same mathematics: Faa di Bruno's formula
  against repeated reverse-mode differentiation, summed in other orders).
* float32 against the float64 oracle: no step size divides the rounding
  here (derivatives are exact), so the loss agrees to 1e-4 relative and the
  gradient to 1e-3 in norm.
"""
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
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _net(layers, dtype, seed=0):
    net = problems.make_mlp(layers, seed=seed, dtype=dtype)
    problems.randomize_biases(net, scale=0.3)
    return net


def _oracle(name, net_cpu, dtype='float64'):
    builder, _ = problems.AD_CASES[name]
    d, eq, b, kw = builder(api, dtype=dtype)
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    loss, norm, op, bv, tv, keys, lens = orc.evaluate(net_cpu, kw['lambda_operator'],
                                                      kw['lambda_bound'])
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net_cpu.parameters()]).numpy()
    return (float(loss), float(norm), op.detach().numpy(), bv.detach().numpy(),
            tv.detach().numpy(), keys, lens, grad)


def _engine(name, net_dev, dtype):
    builder, _ = problems.AD_CASES[name]
    d, eq, b, kw = builder(api, dtype=dtype)
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    sol = Solution(grid, pre, net_dev, 'autograd', None, kw['lambda_operator'],
                   kw['lambda_bound'])
    return sol, kw


@pytest.mark.parametrize('name', list(problems.AD_CASES))
def test_ad_float64_against_oracle(name):
    _print(f"[ad f64] {name}")
    _, layers = problems.AD_CASES[name]
    net = _net(layers, torch.float64)
    ref = _oracle(name, net)
    net_dev = _net(layers, torch.float64).to(DEV)
    sol, _ = _engine(name, net_dev, 'float64')
    loss, norm = sol.evaluate()
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net_dev.parameters()]).cpu().numpy()
    rl, rn, rop, rbv, rtv, rkeys, rlens, rgrad = ref
    np.testing.assert_allclose(float(loss), rl, rtol=1e-9)
    np.testing.assert_allclose(float(norm), rn, rtol=1e-9)
    op = sol.op.cpu().numpy()
    np.testing.assert_allclose(op, rop, rtol=1e-9, atol=1e-9 * np.abs(rop).max())
    np.testing.assert_allclose(sol.bval.cpu().numpy(), rbv, rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(sol.true_bval.cpu().numpy(), rtv)
    assert sol.bval_keys == rkeys and sol.bval_length == rlens
    np.testing.assert_allclose(grad, rgrad, rtol=1e-8, atol=1e-10 * np.abs(rgrad).max())


@pytest.mark.parametrize('name', ['burgers_ad', 'kdv_ad', 'mixed_ad', 'beam_ad'])
def test_ad_float32_against_float64_oracle(name):
    _print(f"[ad f32] {name}")
    _, layers = problems.AD_CASES[name]
    ref = _oracle(name, _net(layers, torch.float64))
    net32 = _net(layers, torch.float64).to(torch.float32).to(DEV)
    sol, _ = _engine(name, net32, 'float32')
    loss, _ = sol.evaluate()
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net32.parameters()]).cpu().numpy()
    assert abs(float(loss) - ref[0]) / ref[0] < 1e-4, (float(loss), ref[0])
    gn = np.linalg.norm(grad - ref[7]) / np.linalg.norm(ref[7])
    assert gn < 1e-3, gn


@pytest.mark.parametrize('name', ['burgers_ad', 'kdv_ad', 'mixed_ad', 'beam_ad'])
def test_ad_float32_against_float32_oracle(name):
    """float32 engine against the float32 oracle (nested torch autograd on
    the CPU, the reference's algorithm at the reference's precision): loss
    to 1e-5 relative and gradient to 1e-5 in norm."""
    _print(f"[ad f32 vs f32] {name}")
    _, layers = problems.AD_CASES[name]
    ref = _oracle(name, _net(layers, torch.float64).to(torch.float32), dtype='float32')
    net32 = _net(layers, torch.float64).to(torch.float32).to(DEV)
    sol, _ = _engine(name, net32, 'float32')
    loss, _ = sol.evaluate()
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net32.parameters()]).cpu().numpy()
    rel = abs(float(loss) - ref[0]) / ref[0]
    gn = np.linalg.norm(grad - ref[7]) / np.linalg.norm(ref[7])
    _print(f"  loss rel {rel:.3e}, gradient rel {gn:.3e}")
    assert rel < 1e-5, (float(loss), ref[0])
    assert gn < 1e-5, gn


@pytest.mark.parametrize('width', [100, 160])
def test_ad_wide_network_against_oracle(width):
    """Width 100 (the reference examples' width; its 16-wide MFMA tiles are
    partly padding) and 160 (inputs wider than 128: 64 k-steps of weight
    fragments held per lane), four hidden layers."""
    _print(f"[ad f64] burgers_ad width {width}")
    layers = [2, width, width, width, width, 1]
    ref = _oracle('burgers_ad', _net(layers, torch.float64))
    net_dev = _net(layers, torch.float64).to(DEV)
    sol, _ = _engine('burgers_ad', net_dev, 'float64')
    loss, _ = sol.evaluate()
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net_dev.parameters()]).cpu().numpy()
    np.testing.assert_allclose(float(loss), ref[0], rtol=1e-9)
    np.testing.assert_allclose(grad, ref[7], rtol=1e-8, atol=1e-10 * np.abs(ref[7]).max())


def test_ad_model_fused_and_torch_optimizer_agree():
    """Model.compile('autograd')/train: the fused graph-captured Adam step
    and torch's Adam over the native autograd Function reach the same
    parameters, and the oracle's Adam loop gives the same loss history."""
    _print("[ad model] wave_ad")
    res = []
    for fused in (True, False):
        d, eq, b, kw = problems.wave_ad(api)
        net = _net([2, 12, 12, 1], torch.float64)
        model = tdx.Model(net, d, eq, b)
        model.verbose = False
        model.compile('autograd', lambda_operator=1, lambda_bound=100)
        model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 21, callbacks=None if fused else [object()])
        res.append(torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu())
    np.testing.assert_allclose(res[0].numpy(), res[1].numpy(), rtol=1e-9, atol=1e-12)
    d, eq, b, kw = problems.wave_ad(api)
    net = _net([2, 12, 12, 1], torch.float64)
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    O.train_adam(list(net.parameters()), lambda: orc.evaluate(net, 1, 100)[0], 1e-3, 21)
    ref = torch.cat([p.detach().reshape(-1) for p in net.parameters()])
    np.testing.assert_allclose(res[0].numpy(), ref.numpy(), rtol=1e-7, atol=1e-10)


def test_ad_model_trains_legendre():
    """float32 training in 'autograd' mode reaches the Legendre polynomial
    P3 = (5x^3 - 3x)/2 (u(-1) = -1, u(1) = 1)."""
    _print("[ad train] legendre_ad float32")
    d, eq, b, kw = problems.legendre_ad(api, dtype='float32')
    net = _net([1, 32, 32, 1], torch.float32)
    model = tdx.Model(net, d, eq, b)
    model.verbose = False
    model.compile('autograd', lambda_operator=1, lambda_bound=10)
    model.train(tdx.Optimizer('Adam', {'lr': 3e-3}), 3001)
    x = torch.linspace(-1, 1, 41, device=DEV).reshape(-1, 1)
    with torch.no_grad():
        u = net(x).reshape(-1).cpu().numpy()
    xs = x.reshape(-1).cpu().numpy()
    p3 = (5 * xs ** 3 - 3 * xs) / 2
    err = np.linalg.norm(u - p3) / np.linalg.norm(p3)
    _print(f"rel L2 to P3: {err:.3e}")
    assert err < 0.05, err


@pytest.mark.parametrize('name', ['burgers_ad', 'system_ad'])
def test_ad_causal_against_oracle(name):
    """The causal loss (tedeous/losses.py:137-182) in 'autograd' mode: the
    residual read as slices of the first coordinate, weights
    exp(-tol * earlier slices) without gradient; float64 to 1e-9 / 1e-8."""
    _print(f"[ad causal f64] {name}")
    builder, layers = problems.AD_CASES[name]
    d, eq, b, kw = builder(api)
    net = _net(layers, torch.float64)
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    lo, no, *_ = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'], tol=0.5)
    lo.sum().backward()
    go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    net_dev = _net(layers, torch.float64).to(DEV)
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    sol = Solution(grid, pre, net_dev, 'autograd', None, kw['lambda_operator'],
                   kw['lambda_bound'], 0.5)
    ln, nn = sol.evaluate()
    ln.backward()
    gn = torch.cat([p.grad.reshape(-1) for p in net_dev.parameters()]).cpu().numpy()
    np.testing.assert_allclose(float(ln), float(lo.sum()), rtol=1e-9)
    np.testing.assert_allclose(float(nn), float(no.sum()), rtol=1e-9)
    np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())


def test_ad_causal_model_fused_and_torch_optimizer_agree():
    _print("[ad causal model] burgers_ad")
    res = []
    for fused in (True, False):
        d, eq, b, kw = problems.burgers_ad(api)
        net = _net([2, 16, 16, 16, 1], torch.float64)
        model = tdx.Model(net, d, eq, b)
        model.verbose = False
        model.compile('autograd', lambda_operator=1, lambda_bound=10, tol=0.5)
        model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 21, callbacks=None if fused else [object()])
        res.append(torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu())
    np.testing.assert_allclose(res[0].numpy(), res[1].numpy(), rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize('case', ['coupled', 'legendre'])
def test_ad_weak_form_against_oracle(case):
    """The weak form in 'autograd' mode on the grid points: a 2-D system of
    two equations (each equation's seeds rescaled by its own run sums) and a
    1-D equation; float64 to 1e-9 / 1e-8."""
    _print(f"[ad weak f64] {case}")
    if case == 'coupled':
        d, eq, b, kw, weak = problems.coupled_weak_ad(api)
        layers = [2, 16, 16, 2]
    else:
        d, eq, b, kw = problems.legendre_ad(api)
        weak = [lambda g: torch.cos(0.5 * np.pi * g[:, 0]) + 0.1]
        layers = [1, 32, 32, 1]
    net = _net(layers, torch.float64)
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    lo, no, *_ = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'], weak_form=weak)
    lo.sum().backward()
    go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    net_dev = _net(layers, torch.float64).to(DEV)
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    sol = Solution(grid, pre, net_dev, 'autograd', weak, kw['lambda_operator'], kw['lambda_bound'])
    ln, nn = sol.evaluate()
    ln.backward()
    gn = torch.cat([p.grad.reshape(-1) for p in net_dev.parameters()]).cpu().numpy()
    np.testing.assert_allclose(float(ln), float(lo.sum()), rtol=1e-9)
    np.testing.assert_allclose(float(nn), float(no.sum()), rtol=1e-9)
    np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())
