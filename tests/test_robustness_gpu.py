"""GPU tests of behaviour around the numerics: non-finite inputs, lambdas
changed between steps (as the reference's adaptive-lambda callback does,
tedeous/callbacks/adaptive_lambda.py:114), and a backward that follows an
unrelated forward pass.

NaN inputs: the reference's torch ops propagate NaN into the loss and the
gradient without an error; the kernels must do the same (no fault, no
exception, NaN out).
"""
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
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _nn(lam_bound=10, dtype='float64', layers=(2, 12, 12, 12, 1)):
    d, eq, b, kw = problems.burgers_nn(api, dtype=dtype)
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    net = problems.randomize_biases(problems.make_mlp(list(layers), dtype=tdt)).to(DEV)
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('NN')
    return Solution(grid, pre, net, 'NN', None, 1, lam_bound), net


def test_nan_parameters_give_nan_loss_and_gradient():
    for dtype in ('float64', 'float32'):
        _print(f"[nan] NN {dtype}")
        sol, net = _nn(dtype=dtype)
        with torch.no_grad():
            next(net.parameters()).view(-1)[3] = float('nan')
        loss, norm = sol.evaluate()
        assert torch.isnan(loss).all() and torch.isnan(norm).all()
        loss.backward()
        g = torch.cat([p.grad.reshape(-1) for p in net.parameters()])
        assert torch.isnan(g).any()
        # the fused step too
        from torch_de_solver_amd.nn_engine import flatten_module_params
        flat = flatten_module_params(net)
        m, v = torch.zeros_like(flat), torch.zeros_like(flat)
        step = torch.zeros(1, dtype=torch.int64, device=DEV)
        lb = torch.zeros(2, dtype=flat.dtype, device=DEV)
        sol.engine.train_step(flat, m, v, step, 1e-3, 0.9, 0.999, 1e-8, lb)
        torch.cuda.synchronize()
        assert torch.isnan(lb).all() and int(step) == 1


def test_nan_grid_value_gives_nan_loss_mat():
    _print("[nan] mat")
    d, eq, b, kw = problems.wave_mat(api)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u0 = problems.mat_init([1, 12, 12])
    u0[0, 5, 6] = float('nan')
    u = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, u, 'mat', None, 1, 100)
    loss, norm = sol.evaluate()
    assert torch.isnan(loss).all() and torch.isnan(norm).all()
    loss.backward()
    assert torch.isnan(u.grad).any()


def test_nan_parameters_autograd():
    _print("[nan] autograd")
    builder, layers = problems.AD_CASES['burgers_ad']
    d, eq, b, kw = builder(api, dtype='float64')
    net = problems.make_mlp(layers, dtype=torch.float64).to(DEV)
    with torch.no_grad():
        list(net.parameters())[2].view(-1)[0] = float('nan')
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    sol = Solution(grid, pre, net, 'autograd', None, kw['lambda_operator'], kw['lambda_bound'])
    loss, norm = sol.evaluate()
    assert torch.isnan(loss).all()
    loss.backward()
    assert torch.isnan(torch.cat([p.grad.reshape(-1) for p in net.parameters()])).any()


def test_lambda_change_between_steps_is_honoured():
    """Setting ``lambda_bound`` on the solution (as the reference's
    AdaptiveLambda does) changes the next loss exactly as a solution built
    with that lambda computes it; an equal value in another form (the
    reference's [1, n] tensor) keeps the plan."""
    sol, net = _nn(lam_bound=10)
    l10 = float(sol.evaluate()[0])
    engine = sol.engine
    sol.lambda_bound = torch.tensor([[10.0]], dtype=torch.float64)
    assert float(sol.evaluate()[0]) == l10 and sol.engine is engine
    sol.lambda_bound = 100
    l100 = sol.evaluate()[0]
    assert sol.engine is not engine
    fresh, _ = _nn(lam_bound=100)
    with torch.no_grad():
        for p, q in zip(fresh.model.parameters(), net.parameters()):
            p.copy_(q)
    assert float(l100) == float(fresh.evaluate()[0])
    assert float(l100) != l10
    # gradient through the rebuilt plan
    l100.backward()
    g = torch.cat([p.grad.reshape(-1) for p in net.parameters()])
    lf = fresh.evaluate()[0]
    lf.backward()
    gf = torch.cat([p.grad.reshape(-1) for p in fresh.model.parameters()])
    np.testing.assert_array_equal(g.cpu().numpy(), gf.cpu().numpy())


def test_backward_after_an_unrelated_forward():
    """A second forward pass (another parameter vector) between a loss and
    its backward does not change that loss's gradient."""
    sol, net = _nn()
    loss, _ = sol.evaluate()
    params = list(net.parameters())
    saved = [p.detach().clone() for p in params]
    with torch.no_grad():
        for p in params:
            p.mul_(1.5)
    sol.evaluate()             # overwrites the engine's seeds
    _ = sol.op                 # and a side-output pass
    with torch.no_grad():
        for p, s in zip(params, saved):
            p.copy_(s)
    loss.backward()
    g = torch.cat([p.grad.reshape(-1) for p in params]).clone()
    for p in params:
        p.grad = None
    l2, _ = sol.evaluate()
    l2.backward()
    g2 = torch.cat([p.grad.reshape(-1) for p in params])
    np.testing.assert_array_equal(g.cpu().numpy(), g2.cpu().numpy())


def test_model_change_rebinds_the_solution():
    """Solution._model_change (tedeous/solution.py:109-127): after the swap,
    evaluate() is the new network's loss (the plan kept for the same
    architecture, lowered anew for another), equal to a solution built
    with that network from the start."""
    sol, net = _nn()
    l0 = float(sol.evaluate()[0])
    for layers in ((2, 12, 12, 12, 1), (2, 8, 8, 1)):
        _print(f"[model change] to {layers}")
        other = problems.randomize_biases(problems.make_mlp(list(layers), seed=5,
                                                            dtype=torch.float64)).to(DEV)
        plan = sol.engine
        sol._model_change(other)
        assert (sol.engine is plan) == (layers == (2, 12, 12, 12, 1))
        l1 = float(sol.evaluate()[0])
        ref, _ = _nn(layers=layers)
        ref._model_change(other)
        fresh, _ = _nn(layers=layers)
        fresh.model = other
        assert l1 == float(ref.evaluate()[0])
        assert abs(l1 - float(fresh.evaluate()[0])) == 0.0
        assert l1 != l0


def test_mixed_precision_keeps_the_reference_control_flow():
    """Model.train(..., mixed_precision=True): the reference's closure
    (tedeous/optimizers/closure.py:49-64) scales the loss, lets the
    GradScaler step the optimizer inside the closure and the optimizer's
    step(closure) then steps again with the same gradient.  Against the
    oracle on the same control flow (two Adam updates per epoch from one
    float64 gradient): losses to 1e-9."""
    import torch_de_solver_amd as tdx
    from oracle import tedeous_oracle as O
    d, eq, b, kw = problems.burgers_nn(api, dtype='float64')
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    net_o = problems.randomize_biases(problems.make_mlp([2, 8, 8, 1], dtype=torch.float64))
    opt = torch.optim.Adam(net_o.parameters(), lr=1e-3)
    ref = []
    for _ in range(4):
        opt.zero_grad()
        lo = orc.evaluate(net_o, kw['lambda_operator'], kw['lambda_bound'])[0]
        lo.backward()
        opt.step()
        opt.step()
        ref.append(float(orc.evaluate(net_o, kw['lambda_operator'], kw['lambda_bound'])[0]))
    net = problems.randomize_biases(problems.make_mlp([2, 8, 8, 1], dtype=torch.float64)).to(DEV)
    model = tdx.Model(net, d, eq, b)
    model.verbose = False
    model.compile('NN', lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    got = []

    class Record:
        def set_model(self, m):
            pass

        def on_train_begin(self, logs=None):
            pass

        def on_epoch_begin(self, logs=None):
            pass

        def on_epoch_end(self, logs=None):
            got.append(float(model.solution_cls.evaluate()[0]))

        def on_train_end(self, logs=None):
            pass

    model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 5, mixed_precision=True, callbacks=[Record()])
    _print(f"[mixed precision] {got} vs {ref}")
    np.testing.assert_allclose(got, ref, rtol=1e-9)
