"""Condition types beyond Dirichlet/operator/periodic: the native engines
against the CPU oracle.

* 'NN' periodic conditions with an operator (the reference's
  example_wave_1d_basic_periodic.py: du/dx at x = 0 minus du/dx at x = 1,
  each side with the one-sided scheme of its own point type);
* 'data' conditions, plain values and operator values, in both modes;
* 'mat' Robin conditions: alpha u + sum_beta beta * (the whole operator),
  tedeous/eval.py:357-388 (the reference evaluates Robin in 'mat' mode only:
  in 'NN' mode its prepared operator is a list and the coefficient lookup
  fails);
* the causal loss in 'NN' mode (tedeous/losses.py:137-182): several points
  per slice (2-D) and one point per slice (1-D), and the fused train step
  against torch.optim.Adam over the autograd path.

float64: loss to 1e-10, condition values to 1e-12 of their scale, gradients
to 1e-8.  float32 'mat': residual field and condition values bit-identical
(correctly rounded division), loss to 1e-6, gradients to 1e-5.  float32
'NN': the float32 bound of the other NN tests (stencil amplification 1/h^2
of last-bit differences of the network output).
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

from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _nn_problem(kind, dtype):
    d = api.Domain()
    d.variable('x', [0, 1], 12, dtype=dtype)
    d.variable('t', [0, 1], 8, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    if kind == 'periodic_operator':
        b.periodic([{'x': 0, 't': [0, 1]}, {'x': 1, 't': [0, 1]}])
        b.periodic([{'x': 0, 't': [0, 1]}, {'x': 1, 't': [0, 1]}],
                   operator={'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1, 'var': 0}})
    else:
        b.data({'x': [0, 1], 't': 0.5}, operator=None, value=0.3)
        b.data({'x': [0, 1], 't': 1},
               operator={'2 du/dt': {'coeff': 2.0, 'du/dt': [1], 'pow': 1, 'var': 0},
                         'u^2': {'coeff': -1.0, 'u': [None], 'pow': 2, 'var': 0}},
               value=0.1)
    eq = api.Equation()
    eq.add(problems.wave_eq())
    return d, eq, b


@pytest.mark.parametrize('kind', ['periodic_operator', 'data'])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_nn_conditions_against_oracle(kind, dtype):
    _print(f"[conditions nn] {kind} {dtype}")
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    d, eq, b = _nn_problem(kind, dtype)
    h = 0.01
    net = problems.make_mlp([2, 24, 24, 1], seed=11, dtype=tdt)
    problems.randomize_biases(net, seed=12)
    ref = problems.make_mlp([2, 24, 24, 1], seed=11, dtype=tdt)
    ref.load_state_dict(net.state_dict())
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=h)
    lo, no, op, bv, tv, keys, lens = orc.evaluate(ref, 1, 100)
    lo.backward()
    g_ref = torch.cat([p.grad.reshape(-1) for p in ref.parameters()]).numpy()

    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=h).set_strategy('NN')
    net = net.to(DEV)
    sol = Solution(grid, pre, net, 'NN', None, 1, 100)
    ln, nn_ = sol.evaluate()
    ln.backward()
    g = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    assert sol.bval_keys == keys
    bvn = sol.bval.detach().cpu().numpy()
    bvo = bv.detach().numpy()
    scale = np.abs(bvo).max()
    _print(f"  loss {float(ln):.10g} (oracle {float(lo):.10g}); bval max diff "
           f"{np.abs(bvn - bvo).max():.3e} of {scale:.3e}")
    if dtype == 'float64':
        np.testing.assert_allclose(float(ln), float(lo), rtol=1e-10)
        np.testing.assert_allclose(bvn, bvo, rtol=0, atol=1e-12 * scale)
        np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-10 * np.abs(g_ref).max())
    else:
        eps = np.finfo(np.float32).eps
        np.testing.assert_allclose(bvn, bvo, rtol=0, atol=8 * eps * 2.0 / h)
        assert abs(float(ln) - float(lo)) / float(lo) < 2e-3
        assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 2e-2


def _mat_problem(dtype):
    d = api.Domain()
    d.variable('x', [0, 1], 40, dtype=dtype)
    d.variable('t', [0, 1], 30, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.robin({'x': 1, 't': [0, 1]}, value=0.2,
            operator={'2u': {'coeff': 2.0, 'u': [None], 'pow': 1, 'var': 0},
                      'du/dx': {'coeff': 0.5, 'du/dx': [0], 'pow': 1, 'var': 0},
                      'du/dt': {'coeff': -0.25, 'du/dt': [1], 'pow': 1, 'var': 0}})
    b.robin({'x': 0, 't': [0, 1]}, value=0.0,
            operator={'u': {'coeff': 1.5, 'u': [None], 'pow': 1, 'var': 0}})
    b.data({'x': 0.5, 't': [0, 1]}, operator=None, value=0.1)
    b.data({'x': [0, 1], 't': 1},
           operator={'du/dt': {'coeff': 3.0, 'du/dt': [1], 'pow': 1, 'var': 0}}, value=0.0)
    eq = api.Equation()
    eq.add(problems.wave_eq())
    return d, eq, b


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_mat_robin_and_data_against_oracle(dtype):
    _print(f"[conditions mat] robin + data {dtype}")
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    d, eq, b = _mat_problem(dtype)
    orc = O.MatOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    u0 = problems.mat_init([1, 41, 31], seed=13, dtype=tdt, std=0.3)
    uo = u0.clone().requires_grad_()
    lo, no, op, bv, tv, keys, lens = orc.evaluate(uo, 1, 10)
    lo.backward()
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    un = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, un, 'mat', None, 1, 10)
    ln, nn_ = sol.evaluate()
    ln.backward()
    assert sol.bval_keys == keys == ['dirichlet', 'robin', 'data']
    bvn = sol.bval.detach().cpu().numpy()
    bvo = bv.detach().numpy()
    gn, go = un.grad.cpu().numpy(), uo.grad.numpy()
    _print(f"  loss {float(ln):.10g} (oracle {float(lo):.10g}); bval max diff "
           f"{np.abs(bvn - bvo).max():.3e}; grad max diff {np.abs(gn - go).max():.3e}")
    if dtype == 'float64':
        np.testing.assert_allclose(sol.op.cpu().numpy(), op.detach().numpy(), rtol=1e-12,
                                   atol=1e-12 * float(op.abs().max()))
        np.testing.assert_allclose(bvn, bvo, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(float(ln), float(lo), rtol=1e-10)
        np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())
    else:
        np.testing.assert_array_equal(sol.op.cpu().numpy(), op.detach().numpy())
        np.testing.assert_array_equal(bvn, bvo)
        np.testing.assert_allclose(float(ln), float(lo), rtol=1e-6)
        np.testing.assert_allclose(gn, go, rtol=1e-5, atol=1e-6 * np.abs(go).max())


@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_mat_causal_loss_against_oracle(dtype):
    """'mat' causal loss with one equation: the oracle's restatement (one
    slice: the operator matrix has one column) against the engine, with a
    lambda_operator the causal loss ignores."""
    _print(f"[conditions mat] causal {dtype}")
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    d, eq, b = _mat_problem(dtype)
    orc = O.MatOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    u0 = problems.mat_init([1, 41, 31], seed=5, dtype=tdt, std=0.3)
    uo = u0.clone().requires_grad_()
    lo, no, op, bv, tv, keys, lens = orc.evaluate(uo, 7, 10, tol=0.5)
    lo.backward()
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    un = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, un, 'mat', None, 7, 10, tol=0.5)
    ln, nn_ = sol.evaluate()
    ln.backward()
    gn, go = un.grad.cpu().numpy(), uo.grad.numpy()
    rt = 1e-10 if dtype == 'float64' else 1e-6
    np.testing.assert_allclose(float(ln), float(lo), rtol=rt)
    np.testing.assert_allclose(float(nn_), float(no), rtol=rt)
    if dtype == 'float32':
        np.testing.assert_array_equal(sol.op.cpu().numpy(), op.detach().numpy())
    np.testing.assert_allclose(gn, go, rtol=1e-8 if dtype == 'float64' else 1e-5,
                               atol=(1e-10 if dtype == 'float64' else 1e-6) * np.abs(go).max())


# ---------------------------------------------------------------------------
# causal loss ('NN' mode, tedeous/losses.py:137-182)
# ---------------------------------------------------------------------------

def _causal_problem(name, dtype):
    if name == 'wave':
        d, eq, b = _nn_problem('data', dtype)
        return d, eq, b, [2, 24, 24, 1], 0.5
    d = api.Domain()
    d.variable('x', [-1, 1], 30, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': -1}, value=-1)
    b.dirichlet({'x': 1}, value=1)
    eq = api.Equation()
    eq.add({
        '(1-x^2)u_xx': {'coeff': lambda g: 1 - g ** 2, 'u_xx': [0, 0], 'pow': 1, 'var': 0},
        '-2x u_x': {'coeff': lambda g: -2 * g, 'u_x': [0], 'pow': 1, 'var': 0},
        '12u': {'coeff': 12, 'u': [None], 'pow': 1, 'var': 0},
    })
    return d, eq, b, [1, 20, 20, 1], 2.0


@pytest.mark.parametrize('name', ['wave', 'legendre'])
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_nn_causal_loss_against_oracle(name, dtype):
    _print(f"[causal nn] {name} {dtype}")
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    d, eq, b, layers, tol = _causal_problem(name, dtype)
    h = 0.01
    net = problems.make_mlp(layers, seed=21, dtype=tdt)
    problems.randomize_biases(net, seed=22)
    ref = problems.make_mlp(layers, seed=21, dtype=tdt)
    ref.load_state_dict(net.state_dict())
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=h)
    lo, no, op, bv, tv, keys, lens = orc.evaluate(ref, 1, 100, tol=tol)
    lo.backward()
    g_ref = torch.cat([p.grad.reshape(-1) for p in ref.parameters()]).numpy()
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=h).set_strategy('NN')
    net = net.to(DEV)
    sol = Solution(grid, pre, net, 'NN', None, 1, 100, tol=tol)
    _print(f"  slices {sol.problem.causal_nt} of {sol.problem.n_central} points")
    ln, nn_ = sol.evaluate()
    ln.backward()
    g = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    _print(f"  loss {float(ln):.10g} (oracle {float(lo):.10g}), normalized {float(nn_):.10g} "
           f"({float(no):.10g})")
    if dtype == 'float64':
        np.testing.assert_allclose(float(ln), float(lo), rtol=1e-10)
        np.testing.assert_allclose(float(nn_), float(no), rtol=1e-10)
        np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-10 * np.abs(g_ref).max())
    else:
        assert abs(float(ln) - float(lo)) / float(lo) < 2e-3
        assert np.linalg.norm(g - g_ref) / np.linalg.norm(g_ref) < 2e-2


def test_nn_causal_fused_step_equals_autograd_adam():
    """Three fused native Adam steps against torch.optim.Adam over the
    autograd path of the same Solution (float64)."""
    _print("[causal nn] fused step vs torch Adam")
    d, eq, b, layers, tol = _causal_problem('wave', 'float64')
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=0.01).set_strategy('NN')
    nets = []
    for _ in range(2):
        n = problems.make_mlp(layers, seed=23, dtype=torch.float64)
        problems.randomize_biases(n, seed=24)
        nets.append(n.to(DEV))
    sol_a = Solution(grid, pre, nets[0], 'NN', None, 1, 100, tol=tol)
    opt = torch.optim.Adam(nets[0].parameters(), lr=1e-3)
    for _ in range(3):
        opt.zero_grad()
        sol_a.evaluate()[0].backward()
        opt.step()
    sol_b = Solution(grid, pre, nets[1], 'NN', None, 1, 100, tol=tol)
    from torch_de_solver_amd.nn_engine import flatten_module_params
    flat = flatten_module_params(nets[1])
    m, v = torch.zeros_like(flat), torch.zeros_like(flat)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss = torch.zeros(2, dtype=torch.float64, device=DEV)
    for _ in range(3):
        sol_b.engine.train_step(flat, m, v, step, 1e-3, 0.9, 0.999, 1e-8, loss)
    torch.cuda.synchronize()
    pa = torch.cat([p.detach().reshape(-1) for p in nets[0].parameters()]).cpu().numpy()
    np.testing.assert_allclose(flat.cpu().numpy(), pa, rtol=1e-9, atol=1e-12)


# ---------------------------------------------------------------------------
# a reference example end to end: example_schrodinger_matrix.py ('mat')
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_schrodinger_matrix_example_against_oracle(dtype):
    """Two coupled equations, 5-point stencils, plain and operator periodic
    conditions, Dirichlet conditions per variable, on the example's grid."""
    _print(f"[schrodinger mat] {dtype}")
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    d, eq, b, kw = problems.schrodinger_mat(api, dtype=dtype)
    orc = O.MatOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                      derivative_points=5)
    u0 = problems.mat_init([2, 41, 41], seed=17, dtype=tdt, std=0.3)
    uo = u0.clone().requires_grad_()
    lo, no, op, bv, tv, keys, lens = orc.evaluate(uo, 1, 1)
    lo.backward()
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    un = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, un, 'mat', None, 1, 1, derivative_points=5)
    ln, nn_ = sol.evaluate()
    ln.backward()
    assert sol.bval_keys == keys == ['dirichlet', 'periodic']
    gn, go = un.grad.cpu().numpy(), uo.grad.numpy()
    opn = sol.op.cpu().numpy()
    _print(f"  loss {float(ln.detach()):.10g} (oracle {float(lo.detach()):.10g}); grad max diff "
           f"{np.abs(gn - go).max():.3e} of {np.abs(go).max():.3e}")
    if dtype == 'float64':
        np.testing.assert_allclose(opn, op.detach().numpy(), rtol=1e-12,
                                   atol=1e-12 * float(op.abs().max()))
        np.testing.assert_allclose(sol.bval.cpu().numpy(), bv.detach().numpy(), rtol=1e-12,
                                   atol=1e-13)
        np.testing.assert_allclose(float(ln.detach()), float(lo.detach()), rtol=1e-10)
        np.testing.assert_allclose(gn, go, rtol=1e-8, atol=1e-10 * np.abs(go).max())
    else:
        np.testing.assert_array_equal(opn, op.detach().numpy())
        np.testing.assert_array_equal(sol.bval.cpu().numpy(), bv.detach().numpy())
        np.testing.assert_allclose(float(ln.detach()), float(lo.detach()), rtol=1e-6)
        np.testing.assert_allclose(gn, go, rtol=1e-5, atol=1e-6 * np.abs(go).max())


def test_schrodinger_matrix_generated_kernel():
    """The same program on a 201 x 161 grid through the kernel generated for
    it (dp = 5: 16-row rings) against the general kernel, bit for bit."""
    from torch_de_solver_amd.mat_engine import MatEngine
    _print("[schrodinger mat] generated kernel, 201 x 161")
    d, eq, b, kw = problems.schrodinger_mat(api, nx=200, nt=160)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u0 = problems.mat_init([2, 201, 161], seed=18, std=0.3)
    sol = Solution(grid, pre, u0.to(DEV).requires_grad_(), 'mat', None, 1, 1, derivative_points=5)
    gen = MatEngine(sol.problem, DEV, specialize=-1)
    spc = MatEngine(sol.problem, DEV, specialize=1)
    assert spc.specialized_points > 0
    u = u0.to(DEV).reshape(-1).contiguous()
    one = torch.ones(1, device=DEV)
    out = []
    for eng in (gen, spc):
        loss, op, bval = eng.fields(u)
        g = torch.empty_like(u)
        eng.backward(u, one, g)
        torch.cuda.synchronize()
        out.append((loss.cpu().numpy(), op.cpu().numpy(), bval.cpu().numpy(), g.cpu().numpy()))
    _print(f"  specialized points {spc.specialized_points}")
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])
    np.testing.assert_array_equal(out[1][3], out[0][3])
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=1e-6)


def test_nn_callable_pow_against_oracle():
    """Callable powers in 'NN' mode (tedeous/derivative.py:54-55: the
    callable applied to the term's running product of stencil sums): the
    sin_source problem (sin(3 u^2) with a coefficient function, (u u_x)^2 as
    [u, u_x] with pow [1, v -> v * v], exp(-u_x / 2)) in float64 against
    the oracle -- loss 1e-10, operator residual and gradient 1e-8 -- and the
    plan runs the separate kernels (callables stay out of the fused
    kernels' tables)."""
    _print("[nn callable pow] sin_source float64")
    d, eq, b, kw = problems.sin_source_ad(api)
    h = 0.01
    net = problems.make_mlp([2, 16, 16, 1], seed=3, dtype=torch.float64)
    problems.randomize_biases(net, seed=4)
    ref = problems.make_mlp([2, 16, 16, 1], seed=3, dtype=torch.float64)
    ref.load_state_dict(net.state_dict())
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=h)
    lo, no, op, bv, tv, keys, lens = orc.evaluate(ref, 1, 10)
    lo.backward()
    g_ref = torch.cat([p.grad.reshape(-1) for p in ref.parameters()]).numpy()
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=h).set_strategy('NN')
    net = net.to(DEV)
    sol = Solution(grid, pre, net, 'NN', None, 1, 10)
    assert sol.engine.step_kernel == 'separate'
    ln, _ = sol.evaluate()
    ln.backward()
    g = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    opn = sol.op.detach().cpu().numpy().reshape(-1)
    opo = op.detach().numpy().reshape(-1)
    _print(f"  loss {float(ln):.12g} (oracle {float(lo):.12g})")
    np.testing.assert_allclose(float(ln), float(lo), rtol=1e-10)
    np.testing.assert_allclose(opn, opo, rtol=1e-8, atol=1e-8 * np.abs(opo).max())
    np.testing.assert_allclose(g, g_ref, rtol=1e-8, atol=1e-10 * np.abs(g_ref).max())
