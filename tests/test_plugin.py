"""The reference's plugin interface on the CPU: ``Operator_bcond_preproc(...)
.set_strategy(mode)`` preparation objects (operator_prepare, bnd_prepare)
and ``Derivative(model, dp).set_strategy(mode).take_derivative`` against the
oracle's restatement of the reference (float64, to 1e-12 relative: the same
operations on the same grids, the operator sums in term order).

* 'NN': residual on the 'central' points; operator conditions prepared per
  point type (compared as sets of values: the reference concatenates the
  types in an order that is not fixed);
* 'mat': residual fields of 1-D and 2-D problems (dp 2 and 3) and the
  condition positions;
* 'autograd': residuals up to fourth order and with two outputs.
"""
import numpy as np
import pytest
import torch

import problems
from oracle import tedeous_oracle as O
from torch_de_solver_amd import data as api
from torch_de_solver_amd.derivative import Derivative
from torch_de_solver_amd.input_preprocessing import (Equation_autograd, Equation_mat,
                                                     Equation_NN, Operator_bcond_preproc)


def _sum_terms(D, op, pts):
    total = None
    for term in op.values():
        v = D.take_derivative(term, pts)
        total = v if total is None else total + v
    return total


def _close(a, b):
    a = torch.as_tensor(a).detach().double().reshape(-1)
    b = torch.as_tensor(b).detach().double().reshape(-1)
    np.testing.assert_allclose(a.numpy(), b.numpy(), rtol=1e-12,
                               atol=1e-12 * float(b.abs().max()))


def _net(layers):
    return problems.randomize_biases(problems.make_mlp(layers, dtype=torch.float64), scale=0.3)


@pytest.mark.parametrize('builder,layers', [
    (problems.burgers_nn, [2, 8, 8, 1]),
    (problems.kdv_nn, [2, 8, 8, 1]),
    (problems.legendre_nn, [1, 8, 8, 1]),
])
def test_nn_operator_prepare_and_take_derivative(builder, layers):
    d, eq, b, kw = builder(api, dtype='float64')
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=kw.get('h', 0.001)).set_strategy('NN')
    assert isinstance(pre, Equation_NN)
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=kw.get('h', 0.001))
    net = _net(layers)
    D = Derivative(net, 2).set_strategy('NN')
    ops = pre.operator_prepare()
    cols = [_sum_terms(D, op, None) for op in ops]
    _close(torch.cat(cols, 1) if len(cols) > 1 else cols[0], orc.residual(net))
    # the operator the solver lowers is left as it was
    assert pre.operator is eq.equation_lst or pre.operator == eq.equation_lst


def test_nn_bnd_prepare_by_point_type():
    """The wave problem's u_t condition at t = 0 in 'NN' mode: its points
    fall into two point types ('bf' at x = 1, 'ff' elsewhere on the edge);
    each type's prepared operator (one-sided second-order schemes) gives
    the oracle's values at those points."""
    d, eq, b, kw = problems.wave_mat(api, dtype='float64')
    grid = d.build('NN')
    bconds = b.build(d.variable_dict)
    pre = Operator_bcond_preproc(grid, eq.equation_lst, bconds).set_strategy('NN')
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    net = _net([2, 8, 8, 1])
    D = Derivative(net, 2).set_strategy('NN')
    prepared = pre.bnd_prepare()
    ob, ot, keys, lens = orc.boundary(net)
    col = keys.index('operator')
    ref = ob[:lens[col], col]
    got = []
    n_types = 0
    for bc in prepared:
        if bc['type'] != 'operator':
            continue
        n_types = len(bc['bop'])
        for op in bc['bop']:
            got.append(_sum_terms(D, op, None).reshape(-1))
    assert n_types == 2
    got = torch.cat(got)
    _close(torch.sort(got).values, torch.sort(ref).values)
    # the input conditions are unchanged
    assert all(isinstance(c['bnd'], torch.Tensor) or isinstance(c['bnd'], list) for c in bconds)


@pytest.mark.parametrize('builder,dp', [
    (problems.wave_mat, 2), (problems.wave_mat, 3), (problems.legendre_mat, 2),
    (problems.nonlinear_mat, 2),
])
def test_mat_prepare_and_take_derivative(builder, dp):
    d, eq, b, kw = builder(api, dtype='float64')
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    assert isinstance(pre, Equation_mat)
    orc = O.MatOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                      derivative_points=dp)
    u = problems.mat_init([1] + list(grid.shape[1:]), dtype=torch.float64, std=0.3)
    D = Derivative(u, dp).set_strategy('mat')
    for op, ref_op in zip(pre.operator_prepare(), orc.ops):
        _close(_sum_terms(D, op, grid), orc.field(u, ref_op))
    for bc, oc in zip(pre.bnd_prepare(), orc.conds):
        sets = bc['bnd'] if bc['type'] == 'periodic' else [bc['bnd']]
        for mine, ref in zip(sets, oc['pos']):
            assert len(mine) == len(ref)
            for a, r in zip(mine, ref):
                assert len(a) == len(r) and all(torch.equal(x.reshape(-1), y.reshape(-1))
                                                for x, y in zip(a, r))


@pytest.mark.parametrize('name', ['burgers_ad', 'beam_ad', 'system_ad', 'mixed_ad', 'lotka_ad'])
def test_autograd_prepare_and_take_derivative(name):
    builder, layers = problems.AD_CASES[name]
    d, eq, b, kw = builder(api, dtype='float64')
    grid = d.build('autograd')
    pre = Operator_bcond_preproc(grid, eq.equation_lst,
                                 b.build(d.variable_dict)).set_strategy('autograd')
    assert isinstance(pre, Equation_autograd)
    orc = O.ADOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    net = _net(layers)
    D = Derivative(net, 2).set_strategy('autograd')
    cols = [_sum_terms(D, op, grid.clone()) for op in pre.operator_prepare()]
    _close(torch.cat(cols, 1) if len(cols) > 1 else cols[0], orc.residual(net))
    assert pre.bnd_prepare() is not None


def test_set_strategy_rejects_unknown_modes():
    grid = torch.zeros(3, 1)
    with pytest.raises(NotImplementedError):
        Operator_bcond_preproc(grid, {}, []).set_strategy('fem')
    with pytest.raises(NotImplementedError):
        Derivative(None, 2).set_strategy('fem')
