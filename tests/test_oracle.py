"""The oracle against the reference's own outputs (golden vectors written by
tests/golden/make_golden.py) and against closed forms.

A mistake planted in the oracle (wrong sign, dropped term, wrong padding)
fails these: the goldens pin the full composition, the closed-form tests pin
the stencil tables independently of the reference.
"""
import numpy as np
import pytest
import torch

import problems
from oracle import tedeous_oracle as O
from torch_de_solver_amd import data as api


def _axes(domain):
    return dict(domain.variable_dict)


def _conds(conds):
    return [dict(c) for c in conds.conditions_lst]


def nn_oracle(name, dtype):
    builder, layers = problems.NN_CASES[name]
    d, eq, b, kw = builder(api, dtype=dtype)
    orc = O.NNOracle(_axes(d), eq.equation_lst, _conds(b), h=kw.get('h', 0.001))
    return orc, kw, layers


@pytest.mark.parametrize('name', list(problems.NN_CASES))
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_nn_oracle_matches_reference(name, dtype, golden):
    g = golden(f'{name}_{dtype}')
    orc, kw, layers = nn_oracle(name, dtype)
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    net = problems.make_mlp(layers, dtype=tdt)
    torch.nn.utils.vector_to_parameters(torch.from_numpy(g['params']), net.parameters())
    loss, norm, op, b, t, keys, lens = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])
    # float64 end to end; float32 residual fields bit-identical (same torch ops)
    rtol = 1e-12 if dtype == 'float64' else 1e-6
    np.testing.assert_allclose(op.detach().numpy(), g['op'], rtol=rtol, atol=0 if dtype == 'float32' else 1e-12)
    np.testing.assert_array_equal(t.numpy(), g['true_bval'])
    np.testing.assert_allclose(b.detach().numpy(), g['bval'], rtol=rtol, atol=1e-12)
    assert list(keys) == list(g['keys']) and list(lens) == list(g['lens'])
    np.testing.assert_allclose(loss.detach().numpy(), g['loss'], rtol=rtol)
    np.testing.assert_allclose(norm.numpy(), g['loss_norm'], rtol=rtol)
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    np.testing.assert_allclose(grad, g['grad'], rtol=1e-9 if dtype == 'float64' else 2e-4,
                               atol=1e-9 * np.abs(g['grad']).max())


@pytest.mark.parametrize('name', list(problems.MAT_CASES))
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_mat_oracle_matches_reference(name, dtype, golden):
    g = golden(f'{name}_{dtype}')
    d, eq, b, kw = problems.MAT_CASES[name](api, dtype=dtype)
    orc = O.MatOracle(_axes(d), eq.equation_lst, _conds(b), kw.get('derivative_points', 2))
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    shape = [len(eq.equation_lst)] + list(orc.grid.shape[1:])
    u = torch.from_numpy(g['params']).reshape(shape).to(tdt).requires_grad_()
    loss, norm, op, bv, t, keys, lens = orc.evaluate(u, kw['lambda_operator'], kw['lambda_bound'])
    # the residual field is the same float operations as the reference: exact
    np.testing.assert_array_equal(op.detach().numpy(), g['op'])
    np.testing.assert_array_equal(bv.detach().numpy(), g['bval'])
    np.testing.assert_array_equal(t.numpy(), g['true_bval'])
    assert list(keys) == list(g['keys']) and list(lens) == list(g['lens'])
    np.testing.assert_allclose(loss.detach().numpy(), g['loss'], rtol=1e-12 if dtype == 'float64' else 1e-6)
    loss.backward()
    np.testing.assert_allclose(u.grad.numpy().reshape(-1), g['grad'],
                               rtol=1e-10 if dtype == 'float64' else 1e-5,
                               atol=1e-12 * np.abs(g['grad']).max())


def test_oracle_adam_steps_match_reference(golden):
    g = golden('burgers_nn_float64')
    orc, kw, layers = nn_oracle('burgers_nn', 'float64')
    net = problems.make_mlp(layers, dtype=torch.float64)
    torch.nn.utils.vector_to_parameters(torch.from_numpy(g['params']), net.parameters())
    hist = O.train_adam(list(net.parameters()),
                        lambda: orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])[0],
                        lr=1e-3, epochs=len(g['losses']) + 1)
    np.testing.assert_allclose(hist, g['losses'], rtol=1e-10)


# --- closed forms: these pin the stencil tables without the reference ------

def test_nn_central_schemes_closed_form():
    h = 1e-3
    sh, sg = O.scheme([0], 1, 'central', '1', h)
    assert sh == [[1], [-1]] and np.allclose(sg, [1 / (2 * h), -1 / (2 * h)])
    sh, sg = O.scheme([0, 0], 1, 'central', '1', h)
    assert sh == [[2], [0], [0], [-2]]
    assert np.allclose(sg, np.array([1, -1, -1, 1]) / (4 * h * h))
    sh, sg = O.scheme([0, 0, 0], 1, 'central', '1', h)
    assert sh == [[3], [1], [1], [-1], [1], [-1], [-1], [-3]]
    assert np.allclose(sg, np.array([1, -1, -1, 1, -1, 1, 1, -1]) / (8 * h ** 3))
    sh, sg = O.scheme([0, 1], 2, 'central', '1', h)
    assert sh == [[1, 1], [1, -1], [-1, 1], [-1, -1]]


def test_nn_stencil_on_polynomials():
    """sum_k w_k f(x + s_k h) reproduces f', f'', f''' of a cubic (step 2h)."""
    h = 1e-2
    x = 0.3
    f = lambda z: 2 * z ** 3 - z ** 2 + 5 * z  # noqa: E731
    ders = {1: 6 * x ** 2 - 2 * x + 5, 2: 12 * x - 2, 3: 12.0}
    for k in (1, 2, 3):
        sh, sg = O.scheme([0] * k, 1, 'central', '1', h)
        val = sum(w * f(x + s[0] * h) for s, w in zip(sh, sg))
        # central differences of a cubic: exact for k = 2, 3; O(h^2) for k = 1
        assert abs(val - ders[k]) < (4 * h * h * 3 if k == 1 else 1e-6)


def test_one_sided_schemes():
    h = 1e-2
    f = lambda z: z ** 2  # noqa: E731
    x = 0.5
    for pt, label in (('f', '1'), ('b', '1')):
        sh, sg = O.scheme([0], 1, pt, label, h)
        v = sum(w * f(x + s[0] * h) for s, w in zip(sh, sg))
        assert abs(v - 2 * x) <= h + 1e-9
    # '2': (3u(x+2h) - 4u(x+h) + u(x)) / 2h = u' + 2h u'' + O(h^2): only O(h)
    sh, sg = O.scheme([0], 1, 'f', '2', h)
    assert sh == [[2], [1], [0]]
    v = sum(w * f(x + s[0] * h) for s, w in zip(sh, sg))
    assert abs(v - (2 * x + 2 * h * 2)) < 1e-9


@pytest.mark.parametrize('dp', [2, 3, 4, 5])
def test_mat_alphas_vandermonde(dp):
    back, forw, ab, af = O.mat_alphas(dp)
    for k in range(dp):  # exact on polynomials of degree < dp
        assert abs(sum(a * (l ** k) for a, l in zip(af, forw)) - (1 if k == 1 else 0)) < 1e-12
        assert abs(sum(a * (l ** k) for a, l in zip(ab, back)) - (1 if k == 1 else 0)) < 1e-12
    if dp == 2:
        assert np.allclose(ab, [-1, 1]) and np.allclose(af, [-1, 1])
    if dp == 3:
        assert np.allclose(ab, [0.5, -2, 1.5]) and np.allclose(af, [-1.5, 2, -0.5])


def test_mat_derivative_equals_per_axis_stencil():
    """The flattened-roll derivative equals a clean per-axis stencil with
    one-sided edges (the wrapped entries are all overwritten)."""
    D = O.MatDerivative(2)
    u = torch.randn(7, 9, dtype=torch.float64)
    h = 0.1
    for axis in (0, 1):
        du = D.d(u, torch.tensor(h, dtype=torch.float64), axis).numpy()
        v = np.moveaxis(u.numpy(), axis, -1)
        ref = np.empty_like(v)
        ref[..., 1:-1] = (v[..., 2:] - v[..., :-2]) / (2 * h)
        ref[..., 0] = (v[..., 1] - v[..., 0]) / h
        ref[..., -1] = (v[..., -1] - v[..., -2]) / h
        np.testing.assert_allclose(du, np.moveaxis(ref, -1, axis), rtol=1e-12)


@pytest.mark.parametrize('dp', [2, 3])
def test_mat_derivative_3d_equals_per_axis_stencil(dp):
    """On a 3-D field (the reference's pos == 2 branch) the derivative along
    each axis is the clean per-axis stencil with one-sided ends, against a
    numpy restatement with the Vandermonde coefficients."""
    D = O.MatDerivative(dp)
    back, forw, ab, af = O.mat_alphas(dp)
    u = torch.randn(6, 7, 8, dtype=torch.float64)
    h = 0.1
    for axis in (0, 1, 2):
        du = D.d(u, torch.tensor(h, dtype=torch.float64), axis).numpy()
        v = np.moveaxis(u.numpy(), axis, -1)
        n = v.shape[-1]
        ref = np.empty_like(v)
        for i in range(n):
            b = sum(a * v[..., i + o] for a, o in zip(ab, back)) if i + back[0] >= 0 else None
            f = sum(a * v[..., i + o] for a, o in zip(af, forw)) if i + forw[-1] < n else None
            if i < dp - 1:
                ref[..., i] = f / h
            elif i >= n - dp + 1:
                ref[..., i] = b / h
            else:
                ref[..., i] = (b + f) / (2 * h)
        np.testing.assert_allclose(du, np.moveaxis(ref, -1, axis), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize('dp', [4, 5])
def test_mat_fourth_order_chain_is_exact_on_cubics(dp):
    """Four successive applications (a factor written [0, 0, 0, 0]) on a
    cubic in x, every stencil exact for degree < dp: the third application
    is 6 everywhere and the fourth 0; three axes, the cubic along axis 1."""
    d = O.MatDerivative(dp)
    n = 3 * dp
    x = torch.linspace(0, 1, n, dtype=torch.float64)
    grid = torch.stack(torch.meshgrid(torch.linspace(0, 1, 5, dtype=torch.float64), x,
                                      torch.linspace(0, 1, 4, dtype=torch.float64),
                                      indexing='ij'))
    u = (grid[1] ** 3 - 2 * grid[1] ** 2 + grid[0] * grid[2]).unsqueeze(0)
    t3 = d.term(u, {'coeff': 1, 'd3': [[1, 1, 1]], 'pow': [1], 'var': [0]}, grid)
    t4 = d.term(u, {'coeff': 1, 'd4': [[1, 1, 1, 1]], 'pow': [1], 'var': [0]}, grid)
    # the step is the float32 difference of the axis values (as the
    # reference takes it), 1e-8 relative off 1/(n-1): a relative error of
    # about 3e-8 per power of 1/h
    np.testing.assert_allclose(t3.numpy(), 6.0, rtol=1e-6)
    np.testing.assert_allclose(t4.numpy(), 0.0, atol=1e-7)


def test_padded_bc_mean_divides_by_max_len():
    b = {'dirichlet': torch.tensor([1., 2., 3., 4.]), 'operator': torch.tensor([2.])}
    t = {'dirichlet': torch.zeros(4), 'operator': torch.zeros(1)}
    m, tm, keys, lens = O.pad_matrix(b, t)
    assert m.shape == (4, 2) and lens == [4, 1]
    mse = torch.mean((m - tm) ** 2, 0)
    assert torch.allclose(mse, torch.tensor([30 / 4, 4 / 4]))


# ---------------------------------------------------------------------------
# condition types beyond the reference's golden vectors: pinned by the
# identities that tie them to the operator and Dirichlet paths pinned above
# ---------------------------------------------------------------------------

def _wave_periodic_nn():
    """The reference's example_wave_1d_basic_periodic.py setup, small: u
    periodic in x, du/dx periodic in x (an operator on each point set)."""
    d = api.Domain()
    d.variable('x', [0, 1], 12, dtype='float64')
    d.variable('t', [0, 1], 8, dtype='float64')
    axes = dict(d.variable_dict)
    bop = {'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1, 'var': 0}}
    sets = [{'x': 0, 't': [0, 1]}, {'x': 1, 't': [0, 1]}]
    return axes, sets, bop


def test_oracle_nn_periodic_operator_is_difference_of_operator_conditions():
    axes, sets, bop = _wave_periodic_nn()
    eq = [problems.wave_eq()]
    per = O.NNOracle(axes, eq, [{'bnd': sets, 'bop': bop, 'bval': torch.tensor([0.]),
                                 'var': 0, 'type': 'periodic'}], h=0.01)
    ops = [O.NNOracle(axes, eq, [{'bnd': s, 'bop': bop, 'bval': 0.0, 'var': 0,
                                  'type': 'operator'}], h=0.01) for s in sets]
    net = problems.make_mlp([2, 16, 16, 1], seed=3, dtype=torch.float64)
    problems.randomize_biases(net, seed=4)
    with torch.no_grad():
        vp = per.boundary(net)[0][:, 0]
        v0 = ops[0].boundary(net)[0][:, 0]
        v1 = ops[1].boundary(net)[0][:, 0]
    # the two sets have points of different types ('f.' at x = 0, 'b.' at x = 1)
    assert torch.all(v0 != 0)
    torch.testing.assert_close(vp, v0 - v1, rtol=0, atol=0)
    assert not torch.allclose(vp, v0 + v1)


def test_oracle_nn_data_conditions_equal_dirichlet_and_operator():
    axes, sets, bop = _wave_periodic_nn()
    eq = [problems.wave_eq()]
    pts = {'x': [0, 1], 't': 0.5}
    net = problems.make_mlp([2, 16, 16, 1], seed=5, dtype=torch.float64)
    problems.randomize_biases(net, seed=6)
    for b in (None, bop):
        data = O.NNOracle(axes, eq, [{'bnd': pts, 'bop': b, 'bval': 0.25, 'var': 0,
                                      'type': 'data'}], h=0.01)
        same = O.NNOracle(axes, eq, [{'bnd': pts, 'bop': b, 'bval': 0.25, 'var': 0,
                                      'type': 'dirichlet' if b is None else 'operator'}], h=0.01)
        with torch.no_grad():
            bd, td, kd, _ = data.boundary(net)
            bs, ts, ks, _ = same.boundary(net)
        assert kd == ['data']
        torch.testing.assert_close(bd, bs, rtol=0, atol=0)
        torch.testing.assert_close(td, ts, rtol=0, atol=0)


def test_oracle_mat_robin_is_alpha_u_plus_beta_operator():
    d = api.Domain()
    d.variable('x', [0, 1], 10, dtype='float64')
    d.variable('t', [0, 1], 8, dtype='float64')
    axes = dict(d.variable_dict)
    eq = [problems.wave_eq()]
    bop = {'u': {'coeff': 2.0, 'u': [None], 'pow': 1, 'var': 0},
           'du/dt': {'coeff': 0.5, 'du/dt': [1], 'pow': 1, 'var': 0},
           'du/dx': {'coeff': -3.0, 'du/dx': [0], 'pow': 1, 'var': 0}}
    pts = {'x': 1, 't': [0, 1]}
    rob = O.MatOracle(axes, eq, [{'bnd': pts, 'bop': bop, 'bval': 0.0, 'var': 0, 'type': 'robin'}])
    opc = O.MatOracle(axes, eq, [{'bnd': pts, 'bop': bop, 'bval': 0.0, 'var': 0,
                                  'type': 'operator'}])
    dic = O.MatOracle(axes, eq, [{'bnd': pts, 'bop': None, 'bval': 0.0, 'var': 0,
                                  'type': 'dirichlet'}])
    u = problems.mat_init([1, 11, 9], seed=7, dtype=torch.float64, std=1.0)
    vr = rob.boundary(u)[0][:, 0]
    vo = opc.boundary(u)[0][:, 0]
    vd = dic.boundary(u)[0][:, 0]
    # alpha = 2 (first term), betas = 0.5, -3: the whole operator (all three
    # terms) is scaled by each beta
    torch.testing.assert_close(vr, 2.0 * vd + (0.5 * vo + -3.0 * vo), rtol=1e-15, atol=1e-15)
    assert not torch.allclose(vr, vo)


def test_oracle_causal_loss_weights_by_loops():
    """The causal loss (tedeous/losses.py:137-182) against explicit loops:
    w[i][k] = exp(-tol * sum_{j<i} res[j][k]) over slices i of the points,
    res = sum over equations of op^2, loss = mean(w res) + boundary part,
    no gradient through w."""
    g = torch.Generator().manual_seed(3)
    n_t, K, n_eq = 5, 4, 2
    op = torch.randn(n_t * K, n_eq, generator=g, dtype=torch.float64).requires_grad_()
    bval = torch.randn(6, 2, generator=g, dtype=torch.float64)
    true = torch.randn(6, 2, generator=g, dtype=torch.float64)
    lam_b = torch.tensor([[10.0, 3.0]], dtype=torch.float64)
    tol = 0.7
    loss, norm = O.causal_loss(op, bval, true, lam_b, tol, n_t)
    loss.backward()
    o = op.detach().numpy()
    res = (o ** 2).sum(1).reshape(n_t, K)
    w = np.ones_like(res)
    for i in range(n_t):
        for k in range(K):
            w[i, k] = np.exp(-tol * sum(res[j, k] for j in range(i)))
    loss_oper = (w * res).mean()
    bd = ((bval - true) ** 2).mean(0).numpy()
    np.testing.assert_allclose(float(loss), loss_oper + (bd * [10.0, 3.0]).sum(), rtol=1e-14)
    np.testing.assert_allclose(float(norm), loss_oper + bd.sum(), rtol=1e-14)
    # gradient: 2 w op / N per equation column (w is a constant)
    gw = 2 * np.repeat(w.reshape(-1, 1), n_eq, 1) * o / (n_t * K)
    np.testing.assert_allclose(op.grad.numpy(), gw, rtol=1e-13)
    # slices that do not divide the points: one slice per point
    l2, _ = O.causal_loss(op.detach(), bval, true, lam_b, tol, 3)
    r1 = (o ** 2).sum(1)
    w1 = np.exp(-tol * np.concatenate([[0.0], np.cumsum(r1)[:-1]]))
    np.testing.assert_allclose(float(l2), (w1 * r1).mean() + (bd * [10.0, 3.0]).sum(), rtol=1e-13)


def test_oracle_causal_loss_small_tol_is_default_loss():
    g = torch.Generator().manual_seed(4)
    op = torch.randn(12, 1, generator=g, dtype=torch.float64)
    bval = torch.randn(5, 1, generator=g, dtype=torch.float64)
    true = torch.zeros(5, 1, dtype=torch.float64)
    lam_b = torch.tensor([[100.0]], dtype=torch.float64)
    lc, nc = O.causal_loss(op, bval, true, lam_b, 1e-14, 4)
    ld, nd = O.default_loss(op, bval, true, torch.ones(1, 1, dtype=torch.float64), lam_b)
    np.testing.assert_allclose(float(lc), float(ld), rtol=1e-12)
    np.testing.assert_allclose(float(nc), float(nd), rtol=1e-12)


# ---------------------------------------------------------------------------
# 'autograd' mode: the nested-autograd restatement pinned by closed forms
# ---------------------------------------------------------------------------

class _Analytic(torch.nn.Module):
    """u0 = sin(2x) e^-t + x^3 t^2, u1 = cos(x t): a stand-in network whose
    partial derivatives are known in closed form."""

    def forward(self, p):
        x, t = p[:, 0], p[:, 1]
        return torch.stack([torch.sin(2 * x) * torch.exp(-t) + x ** 3 * t ** 2,
                            torch.cos(x * t)], 1)


def _u0_derivs(x, t):
    s, c, e = np.sin(2 * x), np.cos(2 * x), np.exp(-t)
    return {
        (): s * e + x ** 3 * t ** 2,
        (0,): 2 * c * e + 3 * x ** 2 * t ** 2,
        (1,): -s * e + 2 * x ** 3 * t,
        (0, 0): -4 * s * e + 6 * x * t ** 2,
        (0, 1): -2 * c * e + 6 * x ** 2 * t,
        (1, 1): s * e + 2 * x ** 3,
        (0, 0, 0): -8 * c * e + 6 * t ** 2,
        (0, 0, 1): 4 * s * e + 12 * x * t,
        (0, 0, 0, 0): 16 * s * e,
    }


def test_ad_oracle_derivatives_closed_form():
    """Every mixed partial up to order 4, in any axis order, and a second
    output: the nested autograd chain equals the closed form."""
    rng = np.random.default_rng(3)
    pts = torch.from_numpy(rng.uniform(-1, 1, size=(17, 2)))
    x, t = pts[:, 0].numpy(), pts[:, 1].numpy()
    want = _u0_derivs(x, t)
    net = _Analytic()
    for axes in ([0], [1], [0, 0], [0, 1], [1, 0], [1, 1], [0, 0, 0], [0, 1, 0], [1, 0, 0],
                 [0, 0, 0, 0]):
        got = O.ad_derivative(net, pts, 0, axes).detach().numpy().reshape(-1)
        np.testing.assert_allclose(got, want[tuple(sorted(axes))], rtol=1e-12, atol=1e-12)
    # u1 = cos(xt): d2/dxdt = -sin(xt) - xt cos(xt)
    got = O.ad_derivative(net, pts, 1, [0, 1]).detach().numpy().reshape(-1)
    np.testing.assert_allclose(got, -np.sin(x * t) - x * t * np.cos(x * t), rtol=1e-12, atol=1e-12)


def test_ad_oracle_residual_and_conditions_closed_form():
    """Residual, operator, Robin and data values and the padded loss of the
    'autograd' oracle on two problems, with the analytic stand-in network."""
    net = _Analytic()
    # wave: u_tt - 4 u_xx; operator u_t at t = 0; Robin 2u + 0.5 u_x at x = 1
    d, eq, b, kw = problems.wave_ad(api, n=6)
    orc = O.ADOracle(_axes(d), eq.equation_lst, _conds(b))
    loss, norm, op, bv, tv, keys, lens = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])
    g = orc.grid.numpy()
    D = _u0_derivs(g[:, 0], g[:, 1])
    np.testing.assert_allclose(op.detach().numpy().reshape(-1), D[(1, 1)] - 4 * D[(0, 0)],
                               rtol=1e-12, atol=1e-12)
    assert keys == ['dirichlet', 'operator', 'robin'] and lens == [14, 7, 7]
    xs = np.linspace(0, 1, 7)
    np.testing.assert_allclose(bv[:7, 1].detach().numpy(), _u0_derivs(xs, 0 * xs)[(1,)],
                               rtol=1e-12, atol=1e-12)
    ts = np.linspace(0, 1, 7)
    Dr = _u0_derivs(1 + 0 * ts, ts)
    # the reference's Robin: alpha u + beta * (alpha u + beta u_x)
    np.testing.assert_allclose(bv[:7, 2].detach().numpy(),
                               2 * Dr[()] + 0.5 * (2 * Dr[()] + 0.5 * Dr[(0,)]),
                               rtol=1e-12, atol=1e-12)
    # padded mean over max_len = 14 for every column
    r = D[(1, 1)] - 4 * D[(0, 0)]
    bdiff = (bv - tv).detach().numpy()
    want = np.mean(r ** 2) + 100 * np.sum(bdiff ** 2) / 14
    np.testing.assert_allclose(float(loss), want, rtol=1e-12)
    np.testing.assert_allclose(float(norm), np.mean(r ** 2) + np.sum(bdiff ** 2) / 14, rtol=1e-12)

    # mixed partials, powers, a three-factor product, tensor and callable
    # coefficients, 'data' with and without an operator
    d, eq, b, kw = problems.mixed_ad(api, n=5)
    orc = O.ADOracle(_axes(d), eq.equation_lst, _conds(b))
    loss, norm, op, bv, tv, keys, lens = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])
    g = orc.grid.numpy()
    x, t = g[:, 0], g[:, 1]
    D = _u0_derivs(x, t)
    want_r = (1 + x * t) * D[(0, 1)] + 0.5 * D[(0, 1)] ** 2 - D[()] * D[(0,)] * D[(1,)] ** 2 \
        + (x - 0.5) * D[(0, 0, 1)]
    np.testing.assert_allclose(op.detach().numpy().reshape(-1), want_r, rtol=1e-11, atol=1e-11)
    assert keys == ['dirichlet', 'data'] and lens == [6, 6]
    px, pt = np.array([0.25, 0.5, 0.75]), np.array([0.5, 1.0, 1.5])
    Dp = _u0_derivs(px, pt)
    np.testing.assert_allclose(bv[:, 1].detach().numpy(), np.concatenate([Dp[()], Dp[(0, 1)]]),
                               rtol=1e-12, atol=1e-12)


def test_ad_oracle_callable_pow_closed_form():
    """Callable powers act on the term's running product
    (tedeous/derivative.py:124-129): with the analytic stand-in network the
    residual of sin_source_ad is u_t - 0.01 u_xx + (1 + x t) sin(3 u^2)
    + 0.5 (u u_x)^2 + exp(-u_x / 2) -- the callable of [u, u_x] sees u * u_x,
    not u_x alone -- and the engine's operation chains of the three callables
    are the traced operations in order."""
    from torch_de_solver_amd.compiler import AD_OP, callable_chain
    net = _Analytic()
    d, eq, b, kw = problems.sin_source_ad(api, n=6)
    orc = O.ADOracle(_axes(d), eq.equation_lst, _conds(b))
    _, _, op, *_ = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])
    g = orc.grid.numpy()
    x, t = g[:, 0], g[:, 1]
    D = _u0_derivs(x, t)
    want = D[(1,)] - 0.01 * D[(0, 0)] + (1 + x * t) * np.sin(3 * D[()] ** 2) \
        + 0.5 * (D[()] * D[(0,)]) ** 2 + np.exp(-D[(0,)] / 2)
    np.testing.assert_allclose(op.detach().numpy().reshape(-1), want, rtol=1e-12, atol=1e-12)
    assert callable_chain(lambda u: torch.sin(3.0 * u ** 2)) == \
        [(AD_OP['POWI'], 2.0), (AD_OP['MUL'], 3.0), (AD_OP['SIN'], 0.0)]
    assert callable_chain(lambda v: torch.exp(-v / 2)) == \
        [(AD_OP['NEG'], 0.0), (AD_OP['DIV'], 2.0), (AD_OP['EXP'], 0.0)]
    assert callable_chain(lambda v: 1 - 2 / v) == [(AD_OP['RDIV'], 2.0), (AD_OP['RSUB'], 1.0)]
    with pytest.raises(NotImplementedError):
        callable_chain(lambda v: v * torch.sin(v))


def test_ad_oracle_system_exact_solution():
    """u' = v, v' = -u is solved exactly by (cos t, -sin t): residuals vanish
    and the loss is the boundary misfit alone."""
    class Osc(torch.nn.Module):
        def forward(self, p):
            return torch.cat([torch.cos(p), -torch.sin(p)], 1)

    d, eq, b, kw = problems.system_ad(api, n=12)
    orc = O.ADOracle(_axes(d), eq.equation_lst, _conds(b))
    loss, norm, op, bv, tv, keys, lens = orc.evaluate(Osc(), kw['lambda_operator'], kw['lambda_bound'])
    assert float(op.abs().max()) < 1e-14
    assert float(loss) < 1e-28


# ---------------------------------------------------------------------------
# weak form: the restated trapezoid loop pinned by numpy's trapezoid
# ---------------------------------------------------------------------------

def test_weak_integration_matches_numpy_trapezoid():
    """One level on a 1-D grid; two levels on a 2-D grid with the integrand
    squared at each level and no pairs across runs of the first coordinate
    (tedeous/eval.py:13-52)."""
    rng = np.random.default_rng(5)
    x = np.sort(rng.uniform(0, 1, 9))
    f = rng.standard_normal(9)
    u, _ = O.integration(torch.from_numpy(f), torch.from_numpy(x).reshape(-1, 1))
    np.testing.assert_allclose(float(u), np.trapezoid(f ** 2, x), rtol=1e-13)
    xs, ts = np.array([0.1, 0.3, 0.35, 0.8]), np.array([0.0, 0.2, 0.5, 0.55, 1.0])
    grid = torch.cartesian_prod(torch.from_numpy(xs), torch.from_numpy(ts))
    s = rng.standard_normal(len(grid))
    op = torch.from_numpy(s).reshape(-1, 1)
    w = O.weak_operator(op, grid, [])
    R = np.array([np.trapezoid(s.reshape(4, 5)[k] ** 2, ts) for k in range(4)])
    np.testing.assert_allclose(float(w), np.trapezoid(R ** 2, xs), rtol=1e-13)
    # the basis functions multiply the residual before squaring
    fb = lambda g: 1 + g[:, 0] * g[:, 1]  # noqa: E731
    w2 = O.weak_operator(op, grid, [fb])
    sb = s * (1 + grid[:, 0] * grid[:, 1]).numpy()
    R2 = np.array([np.trapezoid(sb.reshape(4, 5)[k] ** 2, ts) for k in range(4)])
    np.testing.assert_allclose(float(w2), np.trapezoid(R2 ** 2, xs), rtol=1e-13)


# ---------------------------------------------------------------------------
# natural gradient: the restated NGD pieces pinned by closed forms
# ---------------------------------------------------------------------------

def test_ngd_gram_closed_forms():
    """Per-residual autograd rows: a linear model r = A th - b has J = A, so
    G = A^T A / N; r = sin(A th) has J = diag(cos(A th)) A."""
    rng = np.random.default_rng(11)
    A = torch.from_numpy(rng.standard_normal((13, 5)))
    b = torch.from_numpy(rng.standard_normal(13))
    th = torch.from_numpy(rng.standard_normal(5)).requires_grad_()
    G = O.ngd_gram(A @ th - b, [th])
    np.testing.assert_allclose(G.detach().numpy(), (A.T @ A).numpy() / 13, rtol=1e-14, atol=1e-15)
    G2 = O.ngd_gram(torch.sin(A @ th), [th])
    J = torch.cos(A @ th).detach().reshape(-1, 1) * A
    np.testing.assert_allclose(G2.detach().numpy(), (J.T @ J).numpy() / 13, rtol=1e-13,
                               atol=1e-15)
    # a parameter that does not enter gives zero rows and columns
    z = torch.zeros(2, dtype=torch.float64, requires_grad=True)
    G3 = O.ngd_gram(A @ th, [th, z])
    assert G3.shape == (7, 7) and float(G3[5:].abs().max()) == 0.0


def test_ngd_lstsq_is_the_pseudo_inverse():
    """SVD solution against numpy's pseudo-inverse: full rank (the solve),
    and exactly rank-deficient (singular values below the absolute machine
    epsilon dropped)."""
    rng = np.random.default_rng(12)
    A = rng.standard_normal((6, 6)) + 6 * np.eye(6)
    B = rng.standard_normal(6)
    x = O.ngd_lstsq(torch.from_numpy(A), torch.from_numpy(B)).numpy()
    np.testing.assert_allclose(x, np.linalg.solve(A, B), rtol=1e-12)
    # exactly rank 4 (a permuted diagonal: its zero singular values are
    # exact, so none is inverted)
    Pm = np.eye(6)[[2, 0, 5, 1, 4, 3]]
    S = np.array([3.0, 2.0, 1.0, 0.5, 0.0, 0.0])
    Ar = Pm @ np.diag(S) @ Pm.T
    xr = O.ngd_lstsq(torch.from_numpy(Ar), torch.from_numpy(B)).numpy()
    np.testing.assert_allclose(xr, np.linalg.pinv(Ar, rcond=1e-10) @ B, rtol=1e-10, atol=1e-12)
    # a right-hand side with several columns
    Bm = rng.standard_normal((6, 3))
    xm = O.ngd_lstsq(torch.from_numpy(A), torch.from_numpy(Bm)).numpy()
    np.testing.assert_allclose(xm, np.linalg.solve(A, Bm), rtol=1e-12)


def test_ngd_line_search_takes_the_first_minimum():
    steps = O.ngd_steps(20)
    assert steps.dtype == torch.float32 and len(steps) == 21
    np.testing.assert_array_equal(steps.numpy(), 0.5 ** np.arange(21, dtype=np.float32))
    th = torch.tensor([1.0, -2.0], dtype=torch.float64, requires_grad=True)
    f = torch.tensor([0.75, -1.5], dtype=torch.float64)
    target = torch.tensor([0.625, -1.25], dtype=torch.float64)  # th - 0.5 f

    def loss():
        return ((th - target) ** 2).sum()

    losses, k = O.ngd_line_search([th], f, loss, steps)
    assert k == 1 and float(losses[1]) == 0.0
    np.testing.assert_array_equal(th.detach().numpy(), target.numpy())
    # ties: the first candidate
    th2 = torch.tensor([3.0], dtype=torch.float64, requires_grad=True)
    losses, k = O.ngd_line_search([th2], torch.tensor([1.0], dtype=torch.float64),
                                  lambda: torch.zeros((), dtype=torch.float64), steps)
    assert k == 0


def test_ngd_step_closed_form_on_a_linear_model():
    """Legendre's operator on the linear network u = w x + b: every central
    difference of a linear function is exact, so the residual is
    r = 10 x w + 12 b (J row [10 x, 12]) and the Dirichlet values are
    u(-1) = b - w, u(1) = w + b (J rows [-1, 1], [1, 1], one 'dirichlet'
    column of length 2): G_int = sum [10x, 12]^T [10x, 12] / N and
    G_b = I."""
    d, eq, b, kw = problems.legendre_nn(api, n=20, dtype='float64')
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    net = torch.nn.Linear(1, 1).double()
    with torch.no_grad():
        net.weight.fill_(0.3)
        net.bias.fill_(-0.2)
    out = O.ngd_step(net, orc, 1, 10, grid_steps_number=20)
    x = d.variable_dict['x'].double().numpy()
    rows = np.stack([10 * x, 12 * np.ones_like(x)], 1)
    np.testing.assert_allclose(out['G_int'].numpy(), rows.T @ rows / len(x), rtol=1e-8,
                               atol=1e-12)
    np.testing.assert_allclose(out['G_b'].numpy(), np.eye(2), rtol=1e-12, atol=1e-12)
    # the natural gradient solves G f = grad
    np.testing.assert_allclose(out['G'].numpy() @ out['f'].numpy(), out['grad'].numpy(),
                               rtol=1e-8, atol=1e-10)


# ---------------------------------------------------------------------------
# 'mat' causal loss: the slices are the operator matrix's columns
# ---------------------------------------------------------------------------

def test_mat_causal_loss_restatement():
    """With one equation the 'mat' operator matrix has one column, so the
    causal loss has one slice, w = 1, and equals the default loss with
    lambda_operator = 1 whatever lambda_operator was; with two equations
    (each field broadcast over two copies, [2N, 2]) the operator term is
    (1/(2N)) sum_i R_i (1 + exp(-tol R_i)), R_i = r_0[i]^2 + r_1[i]^2."""
    d, eq, b, kw = problems.wave_mat(api, dtype='float64')
    orc = O.MatOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst])
    u = problems.mat_init([1, 12, 12], dtype=torch.float64, std=0.3)
    lc, nc = orc.evaluate(u, 7, 100, tol=0.5)[:2]
    ld, nd = orc.evaluate(u, 1, 100)[:2]
    assert float(lc) == float(ld) and float(nc) == float(nd)
    # two equations: explicit loops over copies and points
    rng = np.random.default_rng(2)
    N = 9
    r = torch.from_numpy(rng.standard_normal((2, N)))
    op = torch.stack([r[0].repeat(2), r[1].repeat(2)], 1)      # [2N, 2]
    zb = torch.zeros(3, 1, dtype=torch.float64)
    tol = 0.3
    loss, _ = O.causal_loss(op, zb, zb, torch.ones(1, 1, dtype=torch.float64), tol, op.shape[1])
    R = (r ** 2).sum(0).numpy()
    ref = sum(R[i] * (1 + np.exp(-tol * R[i])) for i in range(N)) / (2 * N)
    np.testing.assert_allclose(float(loss), ref, rtol=1e-14)
