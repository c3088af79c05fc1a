"""Small problem definitions shared by the golden-vector generator, the oracle
tests and the GPU parity tests.

Each builder takes a namespace ``api`` providing ``Domain``, ``Conditions``
and ``Equation`` (the reference's classes or this package's, which have the
same interface) and returns ``(domain, equation, conditions, compile_kwargs)``.
Network / grid-tensor initialisation is seeded and independent of ``api``.
"""

import math

import numpy as np
import torch

MU = 0.01 / math.pi


def legendre_eq(grid_coeffs=None, mode='NN'):
    """(1 - x^2) u'' - 2 x u' + 12 u = 0 (n = 3)."""
    if mode == 'NN':
        c1, c2 = grid_coeffs
    else:
        c1 = lambda g: 1 - g ** 2  # noqa: E731
        c2 = lambda g: -2 * g      # noqa: E731
    return {
        '(1-t^2)*d2u/dt2': {'coeff': c1, 'd2u/dt2': [0, 0], 'pow': 1, 'var': 0},
        '-2t*du/dt': {'coeff': c2, 'du/dt': [0], 'pow': 1, 'var': 0},
        'n*(n+1)*u': {'coeff': 12, 'u': [None], 'pow': 1, 'var': 0},
    }


def burgers_eq(mu=MU):
    return {
        'du/dt': {'coeff': 1., 'du/dt': [1], 'pow': 1, 'var': 0},
        '+u*du/dx': {'coeff': 1, 'u*du/dx': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        '-mu*d2u/dx2': {'coeff': -mu, 'd2u/dx2': [0, 0], 'pow': 1, 'var': 0},
    }


def kdv_eq():
    return {
        'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1, 'var': 0},
        '6u*du/dx': {'coeff': 6, 'u*du/dx': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        'd3u/dx3': {'coeff': 1, 'd3u/dx3': [0, 0, 0], 'pow': 1, 'var': 0},
    }


def wave_eq():
    return {
        'd2u/dt2': {'coeff': 1, 'd2u/dt2': [1, 1], 'pow': 1},
        '-4*d2u/dx2': {'coeff': -4, 'd2u/dx2': [0, 0], 'pow': 1},
    }


def legendre_nn(api, n=20, dtype='float32'):
    d = api.Domain()
    d.variable('x', [-1, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': -1}, value=-1)
    b.dirichlet({'x': 1}, value=1)
    grid = d.build('NN')
    eq = api.Equation()
    eq.add(legendre_eq((1 - grid ** 2, -2 * grid), 'NN'))
    return d, eq, b, dict(mode='NN', lambda_operator=1, lambda_bound=10)


def legendre_nn_opbc(api, n=20, dtype='float32'):
    """t in [0, 1], u(0) = P3(0) = 0, u'(1) = P3'(1) = 6 (an operator condition)."""
    d = api.Domain()
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'t': 0}, value=0.0)
    bop = {'du/dt': {'coeff': 1, 'du/dt': [0], 'pow': 1, 'var': 0}}
    b.operator({'t': 1}, operator=bop, value=6.0)
    grid = d.build('NN')
    eq = api.Equation()
    eq.add(legendre_eq((1 - grid ** 2, -2 * grid), 'NN'))
    return d, eq, b, dict(mode='NN', lambda_operator=1, lambda_bound=10)


def burgers_nn(api, nx=15, nt=9, dtype='float32', mu=MU):
    d = api.Domain()
    d.variable('x', [-1, 1], nx, dtype=dtype)
    d.variable('t', [0, 1], nt, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': [-1, 1], 't': 0}, value=lambda g: -torch.sin(np.pi * g[:, 0]))
    b.dirichlet({'x': -1, 't': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    eq = api.Equation()
    eq.add(burgers_eq(mu))
    return d, eq, b, dict(mode='NN', lambda_operator=1, lambda_bound=10)


def kdv_nn(api, nx=24, nt=7, dtype='float32', L=20.0):
    d = api.Domain()
    tdt = torch.float32 if dtype == 'float32' else torch.float64
    x = L * torch.arange(nx, dtype=tdt) / nx
    d.variable('x', x, None, dtype=dtype)
    d.variable('t', [0, 1], nt, dtype=dtype)
    b = api.Conditions()
    b.periodic([{'x': 0, 't': [0, 1]}, {'x': L, 't': [0, 1]}])
    u0 = 0.5 / torch.cosh(0.5 * (x - L / 2)) ** 2
    b.dirichlet({'x': [0, L], 't': 0}, value=u0)
    eq = api.Equation()
    eq.add(kdv_eq())
    return d, eq, b, dict(mode='NN', lambda_operator=1, lambda_bound=100, h=0.01)


def wave_mat(api, n=11, dtype='float32', dp=2):
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    bop = {'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1}}
    b.operator({'x': [0, 1], 't': 0}, operator=bop, value=0)
    eq = api.Equation()
    eq.add(wave_eq())
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=100, derivative_points=dp)


def legendre_mat(api, n=20, dtype='float32'):
    d = api.Domain()
    d.variable('x', [-1, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': -1}, value=-1)
    bop = {'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1, 'var': 0}}
    b.operator({'x': 1}, operator=bop, value=6.0)
    eq = api.Equation()
    eq.add(legendre_eq(mode='mat'))
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=10)


def nonlinear_mat(api, n=14, dtype='float32'):
    """A 2-D 'mat' problem with a product term and a squared factor:
    u_t + u*u_x - 0.1*(u_xx)^2 = 0, Dirichlet on t = 0 and a periodic pair."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 0.5], n - 3, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.cos(np.pi * x))
    b.periodic([{'x': 0, 't': [0, 0.5]}, {'x': 1, 't': [0, 0.5]}])
    eq = api.Equation()
    eq.add({
        'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1, 'var': 0},
        'u*u_x': {'coeff': 1, 'uux': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        'uxx^2': {'coeff': -0.1, 'uxx': [0, 0], 'pow': 2, 'var': 0},
    })
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=10)


def make_mlp(layers, seed=0, dtype=torch.float32):
    """Linear/Tanh stack with Xavier-normal weights and zero biases."""
    torch.manual_seed(seed)
    mods = []
    for i in range(len(layers) - 1):
        lin = torch.nn.Linear(layers[i], layers[i + 1])
        torch.nn.init.xavier_normal_(lin.weight)
        torch.nn.init.zeros_(lin.bias)
        mods.append(lin)
        if i < len(layers) - 2:
            mods.append(torch.nn.Tanh())
    return torch.nn.Sequential(*mods).to(dtype)


def randomize_biases(net, seed=1, scale=0.1):
    """Give the biases non-zero values so that their gradients are exercised."""
    g = torch.Generator().manual_seed(seed)
    with torch.no_grad():
        for m in net:
            if isinstance(m, torch.nn.Linear):
                m.bias.copy_(scale * torch.randn(m.bias.shape, generator=g, dtype=torch.float64)
                             .to(m.bias.dtype))
    return net


NN_CASES = {
    # name: (builder, mlp layers)
    'legendre_nn': (legendre_nn, [1, 16, 16, 1]),
    'legendre_nn_opbc': (legendre_nn_opbc, [1, 16, 16, 1]),
    'burgers_nn': (burgers_nn, [2, 12, 12, 12, 1]),
    'kdv_nn': (kdv_nn, [2, 10, 10, 10, 1]),
}

MAT_CASES = {
    'wave_mat': wave_mat,
    'wave_mat_dp3': lambda api, dtype='float32': wave_mat(api, dtype=dtype, dp=3),
    'legendre_mat': legendre_mat,
    'nonlinear_mat': nonlinear_mat,
}


def heat3_mat(api, n=(8, 7, 6), dtype='float32', dp=2):
    """3-D 'mat' grid (x, y, t): u_t - 0.3 (u_xx + u_yy) - 0.1 u_xyt + 0.5 u u_x
    = 0, u = sin(pi x) sin(pi y) at t = 0, u = 0 at x = 1, a periodic pair in
    y, an operator condition u_x + u_y = 0 at x = 0."""
    d = api.Domain()
    d.variable('x', [0, 1], n[0] - 1, dtype=dtype)
    d.variable('y', [0, 1], n[1] - 1, dtype=dtype)
    d.variable('t', [0, 1], n[2] - 1, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': [0, 1], 'y': [0, 1], 't': 0},
                value=lambda g: torch.sin(np.pi * g[:, 0]) * torch.sin(np.pi * g[:, 1]))
    b.dirichlet({'x': 1, 'y': [0, 1], 't': [0, 1]}, value=0.0)
    b.periodic([{'x': [0, 1], 'y': 0, 't': [0, 1]}, {'x': [0, 1], 'y': 1, 't': [0, 1]}])
    b.operator({'x': 0, 'y': [0, 1], 't': [0, 1]},
               operator={'u_x': {'coeff': 1, 'u_x': [0], 'pow': 1},
                         'u_y': {'coeff': 1, 'u_y': [1], 'pow': 1}}, value=0.0)
    eq = api.Equation()
    eq.add({'u_t': {'coeff': 1, 'u_t': [2], 'pow': 1},
            '-0.3u_xx': {'coeff': -0.3, 'u_xx': [0, 0], 'pow': 1},
            '-0.3u_yy': {'coeff': -0.3, 'u_yy': [1, 1], 'pow': 1},
            '-0.1u_xyt': {'coeff': -0.1, 'u_xyt': [0, 1, 2], 'pow': 1},
            '0.5uu_x': {'coeff': 0.5, 'uux': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]}})
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=10, derivative_points=dp)


def beam_mat(api, n=24, dtype='float32', dp=2):
    """Fourth order on one axis ('mat'): u_xxxx = 1 with u = u_x = 0 at x = 0
    and u_xx = u_xxx = 0 at x = 1."""
    d = api.Domain()
    d.variable('x', [0, 1], n - 1, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': 0}, value=0.0)
    b.operator({'x': 0}, operator={'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1}}, value=0.0)
    b.operator({'x': 1}, operator={'d2u/dx2': {'coeff': 1, 'd2u/dx2': [0, 0], 'pow': 1}},
               value=0.0)
    b.operator({'x': 1}, operator={'d3u/dx3': {'coeff': 1, 'd3u/dx3': [0, 0, 0], 'pow': 1}},
               value=0.0)
    eq = api.Equation()
    eq.add({'d4u/dx4': {'coeff': 1, 'd4u/dx4': [0, 0, 0, 0], 'pow': 1},
            '-1': {'coeff': -1, 'u': [None], 'pow': 0}})
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=10, derivative_points=dp)


def plate_mat(api, n=(16, 12), dtype='float32', dp=2):
    """Fourth order in 2-D ('mat'): u_tt + 0.01 u_xxxx = 0 on [0, 1]^2 with
    u = sin(pi x) and u_t = 0 at t = 0, u = 0 and u_xx = 0 at x = 0 and 1."""
    d = api.Domain()
    d.variable('x', [0, 1], n[0] - 1, dtype=dtype)
    d.variable('t', [0, 1], n[1] - 1, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.operator({'x': [0, 1], 't': 0}, operator={'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1}},
               value=0.0)
    for xb in (0, 1):
        b.dirichlet({'x': xb, 't': [0, 1]}, value=0.0)
        b.operator({'x': xb, 't': [0, 1]},
                   operator={'u_xx': {'coeff': 1, 'u_xx': [0, 0], 'pow': 1}}, value=0.0)
    eq = api.Equation()
    eq.add({'u_tt': {'coeff': 1, 'u_tt': [1, 1], 'pow': 1},
            '0.01u_xxxx': {'coeff': 0.01, 'u_xxxx': [0, 0, 0, 0], 'pow': 1}})
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=10, derivative_points=dp)


# 'mat' problems served by the full-grid passes: 3-D grids, chains of 4
MAT_ND_CASES = {
    'heat3_mat': heat3_mat,
    'heat3_mat_dp3': lambda api, dtype='float32': heat3_mat(api, n=(9, 8, 7), dtype=dtype, dp=3),
    'beam_mat': beam_mat,
    'beam_mat_dp3': lambda api, dtype='float32': beam_mat(api, dtype=dtype, dp=3),
    'plate_mat': plate_mat,
}


def mat_init(shape, seed=0, dtype=torch.float32, std=0.01):
    g = torch.Generator().manual_seed(seed)
    return (std * torch.randn(shape, generator=g, dtype=torch.float64)).to(dtype)


# 2-D float 'mat' programs for the kernel a plan generates for its own
# program: builder(api, nx, nt, dp) -> (domain, equation, conditions, dp,
# number of equations).


def spec_wave(api, nx, nt, dp):
    d = api.Domain()
    d.variable('x', [0, 1], nx)
    d.variable('t', [0, 1], nt)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    b.operator({'x': [0, 1], 't': 0}, operator={'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1}},
               value=0)
    eq = api.Equation()
    eq.add(wave_eq())
    return d, eq, b, dp, 1


def spec_nonlinear(api, nx, nt, dp):
    d = api.Domain()
    d.variable('x', [0, 1], nx)
    d.variable('t', [0, 0.5], nt)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.cos(np.pi * x))
    b.periodic([{'x': 0, 't': [0, 0.5]}, {'x': 1, 't': [0, 0.5]}])
    eq = api.Equation()
    eq.add({
        'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1, 'var': 0},
        'u*u_x': {'coeff': 1, 'uux': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        'uxx^2': {'coeff': -0.1, 'uxx': [0, 0], 'pow': 2, 'var': 0},
    })
    return d, eq, b, dp, 1


def spec_mixed(api, nx, nt, dp):
    """u_t + 0.01 u_xxx + c(x, t) u u_x - 0.5 u^1.5 + 0.2 u_xt + u_tt^3 = 0,
    c = 1 + x t (a grid field); u near 1 keeps the real power defined."""
    d = api.Domain()
    d.variable('x', [0, 1], nx)
    d.variable('t', [0, 1], nt)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=1 + 0.1 * torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 1]}, value=1)
    eq = api.Equation()
    eq.add({
        'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1, 'var': 0},
        'u_xxx': {'coeff': 0.01, 'u_xxx': [0, 0, 0], 'pow': 1, 'var': 0},
        'c u u_x': {'coeff': lambda g: 1 + g[0] * g[1], 'uux': [[None], [0]], 'pow': [1, 1],
                    'var': [0, 0]},
        'u^1.5': {'coeff': -0.5, 'u': [None], 'pow': 1.5, 'var': 0},
        'u_xt': {'coeff': 0.2, 'u_xt': [0, 1], 'pow': 1, 'var': 0},
        'u_tt^3': {'coeff': 1e-3, 'u_tt': [1, 1], 'pow': 3, 'var': 0},
    })
    return d, eq, b, dp, 1


def spec_system(api, nx, nt, dp):
    """u_t - v_x = 0, v_t - u_x + c(x, t) u = 0 (two planes)."""
    d = api.Domain()
    d.variable('x', [0, 1], nx)
    d.variable('t', [0, 1], nt)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x), var=0)
    b.dirichlet({'x': [0, 1], 't': 0}, value=0, var=1)
    eq = api.Equation()
    eq.add({
        'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1, 'var': 0},
        '-v_x': {'coeff': -1, 'v_x': [0], 'pow': 1, 'var': 1},
    })
    eq.add({
        'v_t': {'coeff': 1, 'v_t': [1], 'pow': 1, 'var': 1},
        '-u_x': {'coeff': -1, 'u_x': [0], 'pow': 1, 'var': 0},
        'c u': {'coeff': lambda g: 0.5 * g[0] - g[1], 'u': [None], 'pow': 1, 'var': 0},
    })
    return d, eq, b, dp, 2


SPECIALIZED_CASES = {
    # name: (builder, nx, nt, dp, initial mean)
    'wave_dp2': (spec_wave, 300, 262, 2, 0.0),
    'wave_dp3': (spec_wave, 256, 300, 3, 0.0),
    'nonlinear_dp2': (spec_nonlinear, 279, 230, 2, 0.0),
    'mixed_dp2': (spec_mixed, 260, 211, 2, 1.0),
    'mixed_dp3': (spec_mixed, 199, 240, 3, 1.0),
    'system_dp2': (spec_system, 230, 250, 2, 0.0),
}


def schrodinger_mat(api, nx=40, nt=40, dtype='float32'):
    """The reference's example_schrodinger_matrix.py problem ('mat' mode,
    derivative_points=5): i h_t + h_xx / 2 + |h|^2 h = 0 as two real
    equations (u = Re h, v = Im h) on x in [-5, 5], t in [0, pi/2];
    h(x, 0) = 2 / cosh(x); u, v, u_x and v_x periodic in x."""
    d = api.Domain()
    d.variable('x', [-5, 5], nx, dtype=dtype)
    d.variable('t', [0, np.pi / 2], nt, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [-5, 5], 't': 0}, value=2 / torch.cosh(x), var=0)
    b.dirichlet({'x': [-5, 5], 't': 0}, value=0, var=1)
    per = [{'x': -5, 't': [0, np.pi / 2]}, {'x': 5, 't': [0, np.pi / 2]}]
    b.periodic(per, var=0)
    b.periodic(per, var=1)
    b.periodic(per, operator={'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1, 'var': 0}}, var=0)
    b.periodic(per, operator={'dv/dx': {'coeff': 1, 'dv/dx': [0], 'pow': 1, 'var': 1}}, var=1)
    eq = api.Equation()
    eq.add({
        'du/dt': {'coeff': 1, 'term': [1], 'pow': 1, 'var': 0},
        '1/2*d2v/dx2': {'coeff': 1 / 2, 'term': [0, 0], 'pow': 1, 'var': 1},
        'v * u**2': {'coeff': 1, 'term': [[None], [None]], 'pow': [1, 2], 'var': [1, 0]},
        'v**3': {'coeff': 1, 'term': [None], 'pow': 3, 'var': 1},
    })
    eq.add({
        'dv/dt': {'coeff': 1, 'term': [1], 'pow': 1, 'var': 1},
        '-1/2*d2u/dx2': {'coeff': -1 / 2, 'term': [0, 0], 'pow': 1, 'var': 0},
        '-u * v ** 2': {'coeff': -1, 'term': [[None], [None]], 'pow': [1, 2], 'var': [0, 1]},
        '-u ** 3': {'coeff': -1, 'term': [None], 'pow': 3, 'var': 0},
    })
    return d, eq, b, dict(mode='mat', lambda_operator=1, lambda_bound=1, derivative_points=5)


# ---------------------------------------------------------------------------
# 'autograd' mode (exact derivatives of the network)
# ---------------------------------------------------------------------------

def burgers_ad(api, nx=12, nt=8, dtype='float64', mu=MU):
    d, eq, b, kw = burgers_nn(api, nx=nx, nt=nt, dtype=dtype, mu=mu)
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=10)


def kdv_ad(api, nx=12, nt=6, dtype='float64', L=20.0):
    """Third order in x; plain periodic values and periodic u_x (a periodic
    condition with an operator)."""
    d = api.Domain()
    d.variable('x', [0, L], nx, dtype=dtype)
    d.variable('t', [0, 1], nt, dtype=dtype)
    b = api.Conditions()
    per = [{'x': 0, 't': [0, 1]}, {'x': L, 't': [0, 1]}]
    b.periodic(per)
    b.periodic(per, operator={'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1, 'var': 0}})
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, L], 't': 0}, value=0.5 / torch.cosh(0.5 * (x - L / 2)) ** 2)
    eq = api.Equation()
    eq.add(kdv_eq())
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=100)


def wave_ad(api, n=10, dtype='float64'):
    """Second order in both axes, an operator condition u_t = 0 at t = 0
    and a Robin condition 2 u + 0.5 u_x at x = 1."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 1]}, value=0)
    b.operator({'x': [0, 1], 't': 0}, operator={'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1}},
               value=0)
    b.robin({'x': 1, 't': [0, 1]}, value=0.25,
            operator={'u': {'coeff': 2.0, 'u': [None], 'pow': 1, 'var': 0},
                      'du/dx': {'coeff': 0.5, 'du/dx': [0], 'pow': 1, 'var': 0}})
    eq = api.Equation()
    eq.add(wave_eq())
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=100)


def mixed_ad(api, n=9, dtype='float64'):
    """Mixed partials in either order, powers, a product of three factors, a
    tensor coefficient and 'data' conditions with and without an operator."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 2], n, dtype=dtype)
    grid = d.build('autograd')
    cx = (1 + grid[:, 0] * grid[:, 1]).reshape(-1, 1)
    b = api.Conditions()
    b.dirichlet({'x': 0, 't': [0, 2]}, value=lambda g: torch.sin(g[:, 1]))
    pts = torch.tensor([[0.25, 0.5], [0.5, 1.0], [0.75, 1.5]], dtype=grid.dtype)
    b.data(pts, operator=None, value=torch.tensor([0.1, 0.2, 0.3], dtype=grid.dtype), var=0)
    b.data(pts, value=torch.tensor([0.0, -0.1, 0.1], dtype=grid.dtype),
           operator={'d2u/dxdt': {'coeff': 1, 'd2u/dxdt': [0, 1], 'pow': 1, 'var': 0}})
    eq = api.Equation()
    eq.add({
        'd2u/dxdt': {'coeff': cx, 'd2u/dxdt': [0, 1], 'pow': 1, 'var': 0},
        'd2u/dtdx^2': {'coeff': 0.5, 'd2u/dtdx': [1, 0], 'pow': 2, 'var': 0},
        'u*ux*ut': {'coeff': -1, 'u*ux*ut': [[None], [0], [1]], 'pow': [1, 1, 2],
                    'var': [0, 0, 0]},
        'd3u/dx2dt': {'coeff': lambda g: g[:, 0:1] - 0.5, 'd3u/dx2dt': [0, 1, 0], 'pow': 1,
                      'var': 0},
    })
    return d, eq, b, dict(mode='autograd', lambda_operator=2, lambda_bound=[10, 3])


def system_ad(api, n=30, dtype='float64'):
    """Two outputs on one axis: u' = v, v' = -u, with u(0) = 1, v(0) = 0."""
    d = api.Domain()
    d.variable('t', [0, 3], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'t': 0}, value=1.0, var=0)
    b.dirichlet({'t': 0}, value=0.0, var=1)
    eq = api.Equation()
    eq.add({'du/dt': {'coeff': 1, 'du/dt': [0], 'pow': 1, 'var': 0},
            '-v': {'coeff': -1, 'v': [None], 'pow': 1, 'var': 1}})
    eq.add({'dv/dt': {'coeff': 1, 'dv/dt': [0], 'pow': 1, 'var': 1},
            'u': {'coeff': 1, 'u': [None], 'pow': 1, 'var': 0}})
    return d, eq, b, dict(mode='autograd', lambda_operator=[1, 2], lambda_bound=10)


def lotka_ad(api, n=40, dtype='float64'):
    """Lotka-Volterra in the reference's NGD-example notation
    (examples/examples_lotka_volterra/example_Lotka_Volterra_paper_NGD.py:
    175-224): one-factor terms written with 'term': [0] and 'var': [k],
    two-factor terms with 'var': [0, 1]."""
    d = api.Domain()
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'t': 0}, value=4.0, var=0)
    b.dirichlet({'t': 0}, value=2.0, var=1)
    eq = api.Equation()
    eq.add({'dx/dt': {'coeff': 1, 'term': [0], 'pow': 1, 'var': [0]},
            '-x*alpha': {'coeff': -2., 'term': [None], 'pow': 1, 'var': [0]},
            '+beta*x*y': {'coeff': 2., 'term': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]}})
    eq.add({'dy/dt': {'coeff': 1, 'term': [0], 'pow': 1, 'var': [1]},
            '+y*delta': {'coeff': 2., 'term': [None], 'pow': 1, 'var': [1]},
            '-gamma*x*y': {'coeff': -2., 'term': [[None], [None]], 'pow': [1, 1],
                           'var': [0, 1]}})
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=100)


def beam_ad(api, n=20, dtype='float64'):
    """Fourth order: u'''' = 1 on [0, 1], u(0) = u'(0) = 0, u''(1) = u'''(1) = 0."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': 0}, value=0.0)
    b.operator({'x': 0}, operator={'du/dx': {'coeff': 1, 'du/dx': [0], 'pow': 1}}, value=0.0)
    b.operator({'x': 1}, operator={'d2u/dx2': {'coeff': 1, 'd2u/dx2': [0, 0], 'pow': 1}},
               value=0.0)
    b.operator({'x': 1}, operator={'d3u/dx3': {'coeff': 1, 'd3u/dx3': [0, 0, 0], 'pow': 1}},
               value=0.0)
    eq = api.Equation()
    eq.add({'d4u/dx4': {'coeff': 1, 'd4u/dx4': [0, 0, 0, 0], 'pow': 1, 'var': 0},
            '-1': {'coeff': -1, 'u': [None], 'pow': 0, 'var': 0}})
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=10)


def legendre_ad(api, n=20, dtype='float64'):
    d = api.Domain()
    d.variable('x', [-1, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': -1}, value=-1)
    b.dirichlet({'x': 1}, value=1)
    eq = api.Equation()
    eq.add(legendre_eq(mode='autograd'))
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=10)


# builder, network layers
AD_CASES = {
    'burgers_ad': (burgers_ad, [2, 16, 16, 16, 1]),
    'kdv_ad': (kdv_ad, [2, 20, 20, 1]),
    'wave_ad': (wave_ad, [2, 12, 12, 1]),
    'mixed_ad': (mixed_ad, [2, 24, 24, 1]),
    'system_ad': (system_ad, [1, 16, 16, 2]),
    'beam_ad': (beam_ad, [1, 10, 10, 10, 1]),
    'legendre_ad': (legendre_ad, [1, 32, 32, 1]),
    'lotka_ad': (lotka_ad, [1, 12, 12, 2]),
}


# ---------------------------------------------------------------------------
# 'NN' weak form (tedeous/eval.py:195-221)
# ---------------------------------------------------------------------------

def lotka_weak_nn(api, n=40, dtype='float64'):
    """Lotka-Volterra on t in [0, 1] in weak form with one basis function
    (two equations, 1-D: one trapezoid level)."""
    d = api.Domain()
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'t': 0}, value=1.0, var=0)
    b.dirichlet({'t': 0}, value=0.5, var=1)
    eq = api.Equation()
    eq.add({'du/dt': {'coeff': 1, 'du/dt': [0], 'pow': 1, 'var': 0},
            '-a*u': {'coeff': -0.8, 'u': [None], 'pow': 1, 'var': 0},
            '+b*u*v': {'coeff': 0.5, 'u*v': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]}})
    eq.add({'dv/dt': {'coeff': 1, 'dv/dt': [0], 'pow': 1, 'var': 1},
            '+c*v': {'coeff': 0.6, 'v': [None], 'pow': 1, 'var': 1},
            '-d*u*v': {'coeff': -0.3, 'u*v': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]}})
    weak = [lambda g: torch.sin(np.pi * g[:, 0]) + 1.5]
    return d, eq, b, dict(mode='NN', lambda_operator=[1, 2], lambda_bound=10, h=1e-3), weak


def burgers_weak_nn(api, dtype='float64'):
    """Burgers in weak form with two basis functions (2-D: two trapezoid
    levels)."""
    d, eq, b, kw = burgers_nn(api, dtype=dtype)
    weak = [lambda g: torch.cos(0.5 * np.pi * g[:, 0]) + 0.2,
            lambda g: 1 + g[:, 1]]
    return d, eq, b, kw, weak


WEAK_CASES = {
    'lotka_weak_nn': (lotka_weak_nn, [1, 16, 16, 2]),
    'burgers_weak_nn': (burgers_weak_nn, [2, 12, 12, 12, 1]),
}


def coupled_weak_ad(api, n=10, dtype='float64'):
    """Two coupled equations on a 2-D grid in weak form ('autograd'):
    u_t - v_x = 0, v_t - u_x + u v = 0."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x), var=0)
    b.dirichlet({'x': [0, 1], 't': 0}, value=0.0, var=1)
    eq = api.Equation()
    eq.add({'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1, 'var': 0},
            '-dv/dx': {'coeff': -1, 'dv/dx': [0], 'pow': 1, 'var': 1}})
    eq.add({'dv/dt': {'coeff': 1, 'dv/dt': [1], 'pow': 1, 'var': 1},
            '-du/dx': {'coeff': -1, 'du/dx': [0], 'pow': 1, 'var': 0},
            'u*v': {'coeff': 1, 'u*v': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]}})
    weak = [lambda g: torch.sin(np.pi * g[:, 0]) + 0.5, lambda g: 1 + g[:, 1] ** 2]
    return d, eq, b, dict(mode='autograd', lambda_operator=[1, 3], lambda_bound=10), weak


def coupled_weak_nn(api, dtype='float64'):
    """The coupled two-equation system in weak form on a 2-D grid, 'NN' mode
    (each equation's seeds carry its own run sums: a second forward pass)."""
    d, eq, b, kw, weak = coupled_weak_ad(api, dtype=dtype)
    return d, eq, b, dict(mode='NN', lambda_operator=[1, 3], lambda_bound=10, h=1e-3), weak


WEAK_CASES['coupled_weak_nn'] = (coupled_weak_nn, [2, 16, 16, 2])


def heat3_ad(api, n=5, dtype='float64'):
    """Three inputs: u_t = 0.3 (u_xx + u_yy) + 0.1 u_xyt on the unit cube,
    u = sin(pi x) sin(pi y) at t = 0, an operator condition u_x + u_y at
    x = 0 ('autograd')."""
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('y', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 1], n, dtype=dtype)
    b = api.Conditions()
    b.dirichlet({'x': [0, 1], 'y': [0, 1], 't': 0},
                value=lambda g: torch.sin(np.pi * g[:, 0]) * torch.sin(np.pi * g[:, 1]))
    b.operator({'x': 0, 'y': [0, 1], 't': [0, 1]},
               operator={'u_x': {'coeff': 1, 'u_x': [0], 'pow': 1},
                         'u_y': {'coeff': 1, 'u_y': [1], 'pow': 1}}, value=0.0)
    eq = api.Equation()
    eq.add({'u_t': {'coeff': 1, 'u_t': [2], 'pow': 1},
            '-0.3u_xx': {'coeff': -0.3, 'u_xx': [0, 0], 'pow': 1},
            '-0.3u_yy': {'coeff': -0.3, 'u_yy': [1, 1], 'pow': 1},
            '-0.1u_xyt': {'coeff': -0.1, 'u_xyt': [0, 1, 2], 'pow': 1}})
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=10)


AD_CASES['heat3_ad'] = (heat3_ad, [3, 20, 20, 1])


def sin_source_ad(api, n=10, dtype='float64'):
    """Callable powers (tedeous/derivative.py:124-129: the callable applied
    to the term's running product), after the reference's long-time heat
    example (examples/examples_heat/example_heat_2d_long_time.py:70-100):
    u_t - 0.01 u_xx + c(x, t) sin(3 u^2) + 0.5 (u u_x)^2 + exp(-u_x / 2) = 0,
    the second product term written as [u, u_x] with pow [1, v -> v * v]
    (the callable sees u * u_x), the last as u_x with pow v -> exp(-v / 2)."""
    k = 3.0
    d = api.Domain()
    d.variable('x', [0, 1], n, dtype=dtype)
    d.variable('t', [0, 0.5], n, dtype=dtype)
    b = api.Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 0.5]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 0.5]}, value=0)
    eq = api.Equation()
    eq.add({
        'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1, 'var': 0},
        '-0.01 d2u/dx2': {'coeff': -0.01, 'd2u/dx2': [0, 0], 'pow': 1, 'var': 0},
        'c sin(k u^2)': {'coeff': lambda g: 1 + g[:, 0:1] * g[:, 1:2], 'u': [None],
                         'pow': lambda u: torch.sin(k * u ** 2), 'var': 0},
        '(u u_x)^2': {'coeff': 0.5, 'u*u_x': [[None], [0]], 'pow': [1, lambda v: v * v],
                      'var': [0, 0]},
        'exp(-u_x/2)': {'coeff': 1, 'du/dx': [0], 'pow': lambda v: torch.exp(-v / 2),
                        'var': 0},
    })
    return d, eq, b, dict(mode='autograd', lambda_operator=1, lambda_bound=10)


AD_CASES['sin_source_ad'] = (sin_source_ad, [2, 16, 16, 1])
