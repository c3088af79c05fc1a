"""GPU parity of the 'NN' weak form (tedeous/eval.py:13-52, 195-221,
tedeous/losses.py:184-228) with the oracle's restatement.

Tolerances: float64 loss and normalized loss to 1e-9 relative, parameter
gradient to 1e-8 relative (the trapezoid sums are regrouped: per-point
weights instead of pairwise terms).
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


@pytest.mark.parametrize('name', list(problems.WEAK_CASES))
def test_weak_form_against_oracle(name):
    _print(f"[weak f64] {name}")
    builder, layers = problems.WEAK_CASES[name]
    d, eq, b, kw, weak = builder(api)
    net = problems.randomize_biases(problems.make_mlp(layers, dtype=torch.float64), scale=0.3)
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=kw.get('h', 1e-3))
    lo, no, *_ = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'], weak_form=weak)
    lo.sum().backward()
    go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    for p in net.parameters():
        p.grad = None
    net = net.to(DEV)
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=kw.get('h', 1e-3)).set_strategy('NN')
    sol = Solution(grid, pre, net, 'NN', weak, kw['lambda_operator'], kw['lambda_bound'])
    ln, nn = sol.evaluate()
    ln.backward()
    gn = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    np.testing.assert_allclose(float(ln), float(lo.sum()), rtol=1e-9)
    np.testing.assert_allclose(float(nn), float(no.sum()), rtol=1e-9)
    np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())


def test_weak_form_model_fused_and_torch_optimizer_agree():
    """Model.compile(weak_form=...)/train: the fused graph-captured step and
    torch's Adam over the native autograd Function reach the same
    parameters, and the oracle's Adam loop the same ones."""
    _print("[weak model] lotka_weak_nn")
    res = []
    for fused in (True, False):
        d, eq, b, kw, weak = problems.lotka_weak_nn(api)
        net = problems.randomize_biases(problems.make_mlp([1, 16, 16, 2], dtype=torch.float64))
        model = tdx.Model(net, d, eq, b)
        model.verbose = False
        model.compile('NN', lambda_operator=[1, 2], lambda_bound=10, h=1e-3, weak_form=weak)
        model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 21, callbacks=None if fused else [object()])
        res.append(torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu())
    np.testing.assert_allclose(res[0].numpy(), res[1].numpy(), rtol=1e-9, atol=1e-12)
    d, eq, b, kw, weak = problems.lotka_weak_nn(api)
    net = problems.randomize_biases(problems.make_mlp([1, 16, 16, 2], dtype=torch.float64))
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    O.train_adam(list(net.parameters()),
                 lambda: orc.evaluate(net, [1, 2], 10, weak_form=weak)[0].sum(), 1e-3, 21)
    ref = torch.cat([p.detach().reshape(-1) for p in net.parameters()])
    np.testing.assert_allclose(res[0].numpy(), ref.numpy(), rtol=1e-7, atol=1e-10)


def test_nn_two_output_gradient_strong_form():
    """A network of one input and two outputs (the Lotka-Volterra system,
    ordinary loss): every parameter gradient matches the oracle's, including
    the last hidden layer's weights next to the output layer's in the
    wavefront-private backward's small-term write-out."""
    _print("[nn f64] lotka strong form, 1-16-16-2")
    d, eq, b, kw, _ = problems.lotka_weak_nn(api)
    net = problems.randomize_biases(problems.make_mlp([1, 16, 16, 2], dtype=torch.float64),
                                    scale=0.3)
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    lo = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])[0]
    lo.sum().backward()
    go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    for p in net.parameters():
        p.grad = None
    net = net.to(DEV)
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('NN')
    sol = Solution(grid, pre, net, 'NN', None, kw['lambda_operator'], kw['lambda_bound'])
    ln, _ = sol.evaluate()
    ln.backward()
    gn = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    np.testing.assert_allclose(float(ln), float(lo.sum()), rtol=1e-9)
    np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())
