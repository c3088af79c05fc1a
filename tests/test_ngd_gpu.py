"""GPU parity of natural-gradient descent ('NGD', tedeous/optimizers/ngd.py)
with the oracle's restatement (oracle/tedeous_oracle.py: ngd_gram,
ngd_lstsq, ngd_line_search, ngd_step), which tests/test_oracle.py pins by
closed forms.

Tolerances (float64):
* Jacobian entries: 1e-10 relative (atol 1e-12 of the largest): the same
  derivatives, summed in another order;
* Gram matrices: 1e-10 relative (atol 1e-12 of the largest);
* natural gradient f = G^+ grad: 1e-6 relative in norm (the Gram matrix of
  a 67-parameter network on 148 residuals has a condition number near
  1e8, which multiplies the 1e-15 differences of G);
* line search: the same candidate chosen, the candidate losses to 1e-9,
  and the new parameters to 1e-6 relative.
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
from torch_de_solver_amd.ngd import NGD  # noqa: E402
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')
LAYERS = [2, 6, 6, 1]


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem():
    d, eq, b, kw = problems.burgers_nn(api, dtype='float64')
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    return d, eq, b, kw, orc


def _net():
    return problems.randomize_biases(problems.make_mlp(LAYERS, dtype=torch.float64), scale=0.3)


def _solution(d, eq, b, kw, net):
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('NN')
    return Solution(grid, pre, net, 'NN', None, kw['lambda_operator'], kw['lambda_bound'])


def _oracle_jacobian(orc, net, kw):
    """J of the flattened operator residuals and of the flattened padded
    boundary residuals, one autograd call per entry."""
    params = list(net.parameters())
    _, _, op, b, t, keys, lens = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])

    def rows(res):
        out = []
        for r in res.reshape(-1):
            g = torch.autograd.grad(r, params, retain_graph=True, allow_unused=True)
            g = [torch.zeros_like(p) if x is None else x for p, x in zip(params, g)]
            out.append(torch.cat([x.reshape(-1) for x in g]))
        return torch.stack(out)
    return rows(op), rows(b - t), b.shape


def test_ngd_jacobian_and_gram_against_oracle():
    d, eq, b, kw, orc = _problem()
    net = _net()
    J_op, J_b, bshape = _oracle_jacobian(orc, net, kw)
    sol = _solution(d, eq, b, kw, _net().to(DEV))
    eng = sol.engine
    flat = torch.cat([p.detach().reshape(-1) for p in sol.model.parameters()])
    jt = eng.jacobian(flat).cpu()
    _print(f"[ngd] J: {jt.shape[0]} parameters x {jt.shape[1]} residuals")
    # the engine's residual entries in the oracle's flattened order
    n_types = bshape[1]
    op_cols, b_cols = [], []
    for r in eng.tables.resid:
        cols = list(range(r.offset, r.offset + r.count))
        if r.column < 0:
            op_cols += cols
        else:
            b_cols += [(int(p) * n_types + r.column, c) for p, c in zip(r.col_pos, cols)]
    J_op_n = jt[:, op_cols].T
    np.testing.assert_allclose(J_op_n.numpy(), J_op.numpy(), rtol=1e-10,
                               atol=1e-12 * float(J_op.abs().max()))
    J_b_n = torch.zeros_like(J_b)
    for flat_pos, c in b_cols:
        J_b_n[flat_pos] = jt[:, c]
    np.testing.assert_allclose(J_b_n.numpy(), J_b.numpy(), rtol=1e-10,
                               atol=1e-12 * float(J_b.abs().max()))
    # Gram matrices
    opt = NGD(sol.model.parameters(), grid_steps_number=20)
    _, _, op_idx, b_idx, n_b = opt._prepare(sol)
    G = opt.gram(eng.jacobian(flat), op_idx, b_idx, n_b).cpu()
    G_o = 1.0 / len(J_op) * J_op.T @ J_op + 1.0 / len(J_b) * J_b.T @ J_b
    np.testing.assert_allclose(G.numpy(), G_o.numpy(), rtol=1e-10,
                               atol=1e-12 * float(G_o.abs().max()))


def test_ngd_step_against_oracle():
    d, eq, b, kw, orc = _problem()
    net_o = _net()
    ref = O.ngd_step(net_o, orc, kw['lambda_operator'], kw['lambda_bound'], grid_steps_number=20)
    net = _net().to(DEV)
    model = tdx.Model(net, d, eq, b)
    model.verbose = False
    model.compile('NN', lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    model.train(tdx.Optimizer('NGD', {'grid_steps_number': 20}), 2)   # one NGD step
    opt = model.optimizer
    last = opt.last
    G_n = last['G'].cpu().numpy()
    np.testing.assert_allclose(G_n, ref['G'].numpy(), rtol=1e-10,
                               atol=1e-12 * float(ref['G'].abs().max()))
    np.testing.assert_allclose(last['grad'].cpu().numpy(), ref['grad'].numpy(), rtol=1e-9,
                               atol=1e-12 * float(ref['grad'].abs().max()))
    cond = np.linalg.cond(ref['G'].numpy())
    f_n, f_o = last['f'].cpu().numpy(), ref['f'].numpy()
    ferr = np.linalg.norm(f_n - f_o) / np.linalg.norm(f_o)
    _print(f"[ngd] cond(G) = {cond:.3e}; natural gradient rel {ferr:.3e}; "
           f"step {ref['step']} (k = {ref['k']})")
    assert ferr < 1e-6, ferr
    assert int(last['k']) == ref['k']
    np.testing.assert_allclose(last['losses'].cpu().numpy(), ref['losses'].numpy(), rtol=1e-9)
    p_n = torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu().numpy()
    p_o = torch.cat([p.detach().reshape(-1) for p in net_o.parameters()]).numpy()
    np.testing.assert_allclose(p_n, p_o, rtol=1e-6, atol=1e-9)


def test_ngd_training_reduces_the_loss():
    """Model.train with NGD: every step takes the line search's best
    candidate (the zero-length step is not among them, so the loss may not
    rise by more than what the smallest step changes); over 30 steps the
    loss falls by a large factor, far more than 30 Adam steps at lr 1e-3."""
    d, eq, b, kw, orc = _problem()
    net = _net().to(DEV)
    model = tdx.Model(net, d, eq, b)
    model.verbose = False
    model.compile('NN', lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    l0 = float(model.solution_cls.evaluate()[0])
    model.train(tdx.Optimizer('NGD', {'grid_steps_number': 20}), 31)
    l1 = float(model.solution_cls.evaluate()[0])
    net2 = _net().to(DEV)
    m2 = tdx.Model(net2, d, eq, b)
    m2.verbose = False
    m2.compile('NN', lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    m2.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 31)
    l2 = float(m2.solution_cls.evaluate()[0])
    _print(f"[ngd] loss {l0:.4e} -> NGD {l1:.4e}, Adam {l2:.4e} (30 steps)")
    assert l1 < 0.1 * l0 and l1 < l2
