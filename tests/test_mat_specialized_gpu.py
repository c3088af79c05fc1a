"""The 'mat' kernel generated for a plan's own program (hipRTC, interior
points) against the general tile kernel on the same plan.

Both evaluate the interior with the same float operations in the same order
(central stencils, the plan's division mode, the residual, seeds and
transposed stencils of the reference's 'mat' operator), so the residual
field, the boundary values and the gradient must be bit-identical; the loss
sums the partials in another order and agrees to 1e-6 relative.  The cases
cover both derivative-point counts the reference uses (2, 3), first to third
derivatives along either axis and mixed, products, integer and real powers,
grid coefficient fields, two coupled equations, and all three division
modes (fast product, IEEE, corrected product).
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
from torch_de_solver_amd.mat_engine import MatEngine  # noqa: E402
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = problems.SPECIALIZED_CASES


def _problem(name):
    builder, nx, nt, dp_, mean = CASES[name]
    d, eq, b, dp, n_eq = builder(api, nx, nt, dp_)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    shape = [n_eq] + list(grid.shape[1:])
    u0 = mean + problems.mat_init(shape, seed=5, std=0.05)
    sol = Solution(grid, pre, u0.to(DEV).requires_grad_(), 'mat', None, 1, 10,
                   derivative_points=dp)
    return sol.problem, u0


def _run(eng, u):
    loss, op, bval = eng.fields(u)
    one = torch.ones(1, dtype=u.dtype, device=DEV)
    grad = torch.empty_like(u)
    eng.backward(u, one, grad)
    torch.cuda.synchronize()
    return loss.cpu().numpy(), op.cpu().numpy(), bval.cpu().numpy(), grad.cpu().numpy()


def _compare(prob, u0, exact_div, specialize=1, steps=3):
    """exact_div: 0 fast product, 1 checked corrected product, 2 IEEE."""
    gen = MatEngine(prob, DEV, exact_div=exact_div, specialize=-1)
    spc = MatEngine(prob, DEV, exact_div=exact_div, specialize=specialize)
    assert gen.specialized_points == 0
    assert spc.specialized_points > 0
    u = u0.to(DEV).contiguous()
    lg, og, bg, gg = _run(gen, u)
    ls, os_, bs, gs = _run(spc, u)
    _print(f"  specialized points {spc.specialized_points} of {u.numel()}; "
           f"op max diff {np.abs(og - os_).max():.3e}, grad max diff {np.abs(gg - gs).max():.3e}")
    np.testing.assert_array_equal(os_, og)
    np.testing.assert_array_equal(bs, bg)
    np.testing.assert_array_equal(gs, gg)
    np.testing.assert_allclose(ls, lg, rtol=1e-6)
    # fused Adam steps from the same start: the same gradients give the same
    # iterates
    ws = []
    for eng in (gen, spc):
        w = u.clone()
        m, v = torch.zeros_like(w), torch.zeros_like(w)
        step = torch.zeros(1, dtype=torch.int64, device=DEV)
        loss = torch.zeros(2, dtype=w.dtype, device=DEV)
        for _ in range(steps):
            eng.train_step(w, m, v, step, 1e-3, 0.9, 0.999, 1e-8, loss)
        torch.cuda.synchronize()
        ws.append(w.cpu().numpy())
    np.testing.assert_array_equal(ws[1], ws[0])


@pytest.mark.parametrize('name', list(CASES))
def test_specialized_equals_general(name):
    _print(f"[mat specialized] {name} exact division")
    prob, u0 = _problem(name)
    _compare(prob, u0.reshape(-1), exact_div=1)


@pytest.mark.parametrize('name', ['wave_dp2', 'mixed_dp2', 'system_dp2'])
def test_specialized_equals_general_fast_division(name):
    _print(f"[mat specialized] {name} fast division")
    prob, u0 = _problem(name)
    _compare(prob, u0.reshape(-1), exact_div=0)


@pytest.mark.parametrize('name', ['wave_dp2', 'mixed_dp3'])
def test_specialized_equals_general_ieee_division(name):
    """exact_div = 2: correctly rounded quotients by IEEE division in both
    kernels (the mode a plan takes when the corrected product is not exact
    for its step)."""
    _print(f"[mat specialized] {name} IEEE division")
    prob, u0 = _problem(name)
    _compare(prob, u0.reshape(-1), exact_div=2)


def test_specialized_by_default_on_large_grids():
    """specialize = 0 (the default) takes the generated kernel once the
    interior holds 2^20 points."""
    _print("[mat specialized] wave 1301 x 1201, default selection")
    builder, _, _, _, _ = CASES['wave_dp2']
    d, eq, b, dp, n_eq = builder(api, 1300, 1200, 2)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u0 = problems.mat_init([1] + list(grid.shape[1:]), seed=7, std=0.05)
    sol = Solution(grid, pre, u0.to(DEV).requires_grad_(), 'mat', None, 1, 10, derivative_points=2)
    assert sol.engine.specialized_points >= (1 << 20)
    _compare(sol.problem, u0.reshape(-1), exact_div=1, specialize=0, steps=2)


# ---------------------------------------------------------------------------
# fused step into a second grid buffer (Adam inside the generated kernel)
# ---------------------------------------------------------------------------

def _compare_step_to(prob, u0, specialize, steps=3, dtype=torch.float32):
    """tdx_mat_train_step_to alternating two buffers against the in-place
    tdx_mat_train_step: identical grids, moments and step counts."""
    eng = MatEngine(prob, DEV, exact_div=1, specialize=specialize)
    hp = (1e-3, 0.9, 0.999, 1e-8)
    u0 = u0.to(DEV, dtype).contiguous()
    w = u0.clone()
    m, v = torch.zeros_like(w), torch.zeros_like(w)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss = torch.zeros(2, dtype=dtype, device=DEV)
    losses = []
    for _ in range(steps):
        eng.train_step(w, m, v, step, *hp, loss)
        losses.append(loss.cpu().numpy().copy())
    bufs = [u0.clone(), torch.empty_like(u0)]
    m2, v2 = torch.zeros_like(w), torch.zeros_like(w)
    step2 = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss2 = torch.zeros(2, dtype=dtype, device=DEV)
    for k in range(steps):
        eng.train_step_to(bufs[k % 2], bufs[(k + 1) % 2], m2, v2, step2, *hp, loss2)
        if k == 0:
            torch.cuda.synchronize()
            np.testing.assert_array_equal(bufs[0].cpu().numpy(), u0.cpu().numpy())
        np.testing.assert_allclose(loss2.cpu().numpy(), losses[k], rtol=1e-6)
    torch.cuda.synchronize()
    out = bufs[steps % 2]
    _print(f"  specialized points {eng.specialized_points}; u max diff "
           f"{float((out - w).abs().max()):.3e}, m {float((m2 - m).abs().max()):.3e}")
    np.testing.assert_array_equal(out.cpu().numpy(), w.cpu().numpy())
    np.testing.assert_array_equal(m2.cpu().numpy(), m.cpu().numpy())
    np.testing.assert_array_equal(v2.cpu().numpy(), v.cpu().numpy())
    assert int(step2.item()) == int(step.item()) == steps
    return eng


@pytest.mark.parametrize('name', list(CASES))
def test_fused_step_to_equals_in_place_step(name):
    _print(f"[mat step_to] {name} specialized")
    prob, u0 = _problem(name)
    eng = _compare_step_to(prob, u0.reshape(-1), specialize=1)
    assert eng.specialized_points > 0


def test_fused_step_to_general_kernel_only():
    _print("[mat step_to] wave_dp2 without the generated kernel")
    prob, u0 = _problem('wave_dp2')
    eng = _compare_step_to(prob, u0.reshape(-1), specialize=-1)
    assert eng.specialized_points == 0


def test_fused_step_to_double():
    _print("[mat step_to] nonlinear float64")
    d, eq, b, kw = problems.nonlinear_mat(api, n=40, dtype='float64')
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u0 = problems.mat_init([1] + list(grid.shape[1:]), seed=8, dtype=torch.float64, std=0.05)
    sol = Solution(grid, pre, u0.to(DEV).requires_grad_(), 'mat', None, 1, 10)
    _compare_step_to(sol.problem, u0.reshape(-1), specialize=0, dtype=torch.float64)


def test_fused_step_to_boundary_inside_interior():
    """A Dirichlet line through the interior (t = 0.5): its adjoint adds to
    points of the generated kernel's rectangle, so the step keeps the
    gradient pass and updates every point afterwards."""
    _print("[mat step_to] wave with an interior Dirichlet line")
    d, eq, b, dp, n_eq = problems.spec_wave(api, 300, 262, 2)
    b.dirichlet({'x': [0, 1], 't': 0.5}, value=0)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u0 = problems.mat_init([1] + list(grid.shape[1:]), seed=9, std=0.05)
    sol = Solution(grid, pre, u0.to(DEV).requires_grad_(), 'mat', None, 1, 10, derivative_points=2)
    assert sol.problem.bcs and any(len(bc.positions[0]) > 0 for bc in sol.problem.bcs)
    _compare_step_to(sol.problem, u0.reshape(-1), specialize=1)


@pytest.mark.parametrize('epochs', [7, 8])
def test_model_train_mat_with_generated_kernel(epochs):
    """Model.train in 'mat' mode on a grid whose interior takes the
    generated kernel by default: the fused steps alternate two grid buffers
    (two captured graphs) and leave the result in the model's tensor, for an
    even and an odd number of steps; identical to in-place steps of the
    general kernel."""
    import torch_de_solver_amd as tdx
    _print(f"[mat model] wave 1301 x 1201, {epochs - 1} steps")
    d, eq, b, dp, n_eq = problems.spec_wave(api, 1300, 1200, 2)
    grid = d.build('mat')
    u0 = problems.mat_init([1] + list(grid.shape[1:]), seed=31, std=0.05).to(DEV)
    u_ref = u0.clone()
    model = tdx.Model(u0.clone(), d, eq, b)
    model.verbose = False
    model.compile('mat', lambda_operator=1, lambda_bound=10)
    assert model.solution_cls.engine.specialized_points >= (1 << 20)
    model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), epochs)
    assert model.steps_done == epochs - 1
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    sol = Solution(grid, pre, u_ref.clone().requires_grad_(), 'mat', None, 1, 10)
    gen = MatEngine(sol.problem, DEV, exact_div=1, specialize=-1)
    w = u_ref.reshape(-1).contiguous()
    m, v = torch.zeros_like(w), torch.zeros_like(w)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss = torch.zeros(2, dtype=w.dtype, device=DEV)
    for _ in range(epochs - 1):
        gen.train_step(w, m, v, step, 1e-3, 0.9, 0.999, 1e-8, loss)
    torch.cuda.synchronize()
    out = model.net.detach().reshape(-1)
    _print(f"  max diff {float((out - w).abs().max()):.3e}")
    np.testing.assert_array_equal(out.cpu().numpy(), w.cpu().numpy())
