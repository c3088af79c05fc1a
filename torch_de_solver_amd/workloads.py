"""The standard workloads: Legendre ('NN', CPU-sized), 2-D wave ('mat'),
Burgers ('NN'), KdV ('NN') and a sweep of independent Burgers solves.

Each builder returns ``(domain, equation, conditions, net_or_grid,
compile_kwargs)`` written against this package's reference-compatible API,
with seeded initialisation.  Sizes follow the problem statements:

* Legendre: x in [-1, 1], 100 points, u(-1) = -1, u(1) = 1, MLP 1-100-100-1.
* wave: u_tt = 4 u_xx on [0, 1]^2, 128 x 128, u(x,0) = sin(pi x),
  u_t(x,0) = 0, u(0,t) = u(1,t) = 0; grid tensor N(0, 0.01^2) (std 0.01).
* Burgers: u_t + u u_x = nu u_xx, nu = 0.01/pi, x in [-1, 1] (256 points),
  t in [0, 1] (100 points), u(x,0) = -sin(pi x), u(+-1,t) = 0, MLP
  2-100x4-1, h = 1e-3.
* KdV: u_t + 6 u u_x + u_xxx = 0, x in [0, 20) periodic (512 points),
  t in [0, 1] (128 points), u(x,0) = 0.5 sech^2(0.5(x-10)), MLP 2-64x4-1,
  h = 1e-2.
* Poisson ('autograd'): -u_xx - u_yy = 2 pi^2 sin(pi x) sin(pi y) on
  [0, 1]^2, u = 0 on the sides, 201 x 201 points, MLP 2-100x5-1.
* sweep: 256 Burgers solves, nu = 10^U(-3,-1), nx in [64, 512],
  nt in [32, 128], MLP 2-32-32-32-1 each, seeded per instance.
* Lotka-Volterra ('autograd', the natural-gradient workload): x' = x(a - b y),
  y' = y(-d + g x), a = b = d = g = 20, x(0) = 4, y(0) = 2, t in [0, 1]
  (501 points), MLP 1-32-32-2, lambda_bound 100.
"""

import math
from dataclasses import dataclass
from typing import List, Tuple

import numpy as np
import torch

from .data import Conditions, Domain, Equation


def tanh_mlp(layers: List[int], seed: int, dtype=torch.float32) -> torch.nn.Sequential:
    """Linear/Tanh stack, Xavier-normal weights, zero biases, seeded."""
    g = torch.Generator().manual_seed(seed)
    mods = []
    for i in range(len(layers) - 1):
        lin = torch.nn.Linear(layers[i], layers[i + 1])
        std = math.sqrt(2.0 / (layers[i] + layers[i + 1]))
        with torch.no_grad():
            lin.weight.copy_(torch.randn(lin.weight.shape, generator=g) * std)
            lin.bias.zero_()
        mods.append(lin)
        if i < len(layers) - 2:
            mods.append(torch.nn.Tanh())
    return torch.nn.Sequential(*mods).to(dtype)


def legendre(n_points: int = 99, seed: int = 0):
    d = Domain()
    d.variable('x', [-1, 1], n_points)
    b = Conditions()
    b.dirichlet({'x': -1}, value=-1)
    b.dirichlet({'x': 1}, value=1)
    grid = d.build('NN')
    eq = Equation()
    eq.add({
        '(1-x^2)u_xx': {'coeff': 1 - grid ** 2, 'u_xx': [0, 0], 'pow': 1, 'var': 0},
        '-2x u_x': {'coeff': -2 * grid, 'u_x': [0], 'pow': 1, 'var': 0},
        '12u': {'coeff': 12, 'u': [None], 'pow': 1, 'var': 0},
    })
    net = tanh_mlp([1, 100, 100, 1], seed)
    return d, eq, b, net, dict(mode='NN', lambda_operator=1, lambda_bound=10)


def wave(n: int = 128, seed: int = 0):
    d = Domain()
    d.variable('x', [0, 1], n - 1)
    d.variable('t', [0, 1], n - 1)
    b = Conditions()
    x = d.variable_dict['x']
    b.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b.dirichlet({'x': 0, 't': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    b.operator({'x': [0, 1], 't': 0}, operator={'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1}},
               value=0)
    eq = Equation()
    eq.add({
        'u_tt': {'coeff': 1, 'u_tt': [1, 1], 'pow': 1},
        '-4u_xx': {'coeff': -4, 'u_xx': [0, 0], 'pow': 1},
    })
    g = torch.Generator().manual_seed(seed)
    u0 = 0.01 * torch.randn((1, n, n), generator=g)
    return d, eq, b, u0, dict(mode='mat', lambda_operator=1, lambda_bound=100)


def poisson_ad(grid_res: int = 200, seed: int = 0):
    """'autograd' mode: -u_xx - u_yy = 2 pi^2 sin(pi x) sin(pi y) on [0, 1]^2
    with u = 0 on the sides (u = sin(pi x) sin(pi y)), (grid_res + 1)^2
    points, MLP 2-100x5-1 (the network of the reference's 2-D Poisson
    autograd examples)."""
    d = Domain()
    d.variable('x', [0, 1], grid_res)
    d.variable('y', [0, 1], grid_res)
    b = Conditions()
    b.dirichlet({'x': 0, 'y': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 'y': [0, 1]}, value=0)
    b.dirichlet({'x': [0, 1], 'y': 0}, value=0)
    b.dirichlet({'x': [0, 1], 'y': 1}, value=0)
    eq = Equation()
    eq.add({
        '-u_xx': {'coeff': -1., 'u_xx': [0, 0], 'pow': 1, 'var': 0},
        '-u_yy': {'coeff': -1., 'u_yy': [1, 1], 'pow': 1, 'var': 0},
        '-f': {'coeff': lambda g: -2 * np.pi ** 2 * torch.sin(np.pi * g[:, 0]) *
               torch.sin(np.pi * g[:, 1]), 'u': [None], 'pow': 0, 'var': 0},
    })
    net = tanh_mlp([2, 100, 100, 100, 100, 100, 1], seed)
    return d, eq, b, net, dict(mode='autograd', lambda_operator=1, lambda_bound=100)


def lotka_ad(grid_res: int = 500, seed: int = 0, dtype: str = 'float32'):
    """'autograd' Lotka-Volterra system as trained by natural-gradient
    descent in the reference's examples (examples/examples_lotka_volterra/
    example_Lotka_Volterra_paper_NGD.py:30-40, 159-228: the equations,
    constants, conditions and lambda_bound = 100) with the plain 1-32-32-2
    tanh network that example keeps as its alternative (lines 131-137)."""
    al = be = de = ga = 20.
    d = Domain()
    d.variable('t', [0, 1], grid_res, dtype=dtype)
    b = Conditions()
    b.dirichlet({'t': 0}, value=4., var=0)
    b.dirichlet({'t': 0}, value=2., var=1)
    eq = Equation()
    eq.add({
        'dx/dt': {'coeff': 1, 'term': [0], 'pow': 1, 'var': [0]},
        '-x*alpha': {'coeff': -al, 'term': [None], 'pow': 1, 'var': [0]},
        '+beta*x*y': {'coeff': be, 'term': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]},
    })
    eq.add({
        'dy/dt': {'coeff': 1, 'term': [0], 'pow': 1, 'var': [1]},
        '+y*delta': {'coeff': de, 'term': [None], 'pow': 1, 'var': [1]},
        '-gamma*x*y': {'coeff': -ga, 'term': [[None], [None]], 'pow': [1, 1], 'var': [0, 1]},
    })
    net = tanh_mlp([1, 32, 32, 2], seed, torch.float32 if dtype == 'float32' else torch.float64)
    return d, eq, b, net, dict(mode='autograd', lambda_operator=1, lambda_bound=100)


def burgers_equation(nu: float) -> dict:
    return {
        'u_t': {'coeff': 1., 'u_t': [1], 'pow': 1, 'var': 0},
        'u*u_x': {'coeff': 1, 'u*u_x': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        '-nu*u_xx': {'coeff': -nu, 'u_xx': [0, 0], 'pow': 1, 'var': 0},
    }


def burgers(nx: int = 256, nt: int = 100, nu: float = 0.01 / math.pi,
            layers=(2, 100, 100, 100, 100, 1), seed: int = 0, h: float = 1e-3,
            dtype: str = 'float32'):
    d = Domain()
    d.variable('x', [-1, 1], nx - 1, dtype=dtype)
    d.variable('t', [0, 1], nt - 1, dtype=dtype)
    b = Conditions()
    b.dirichlet({'x': [-1, 1], 't': 0}, value=lambda g: -torch.sin(np.pi * g[:, 0]))
    b.dirichlet({'x': -1, 't': [0, 1]}, value=0)
    b.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    eq = Equation()
    eq.add(burgers_equation(nu))
    net = tanh_mlp(list(layers), seed, dtype=torch.float64 if dtype == 'float64' else torch.float32)
    return d, eq, b, net, dict(mode='NN', lambda_operator=1, lambda_bound=10, h=h)


def kdv(nx: int = 512, nt: int = 128, L: float = 20.0, seed: int = 0, dtype: str = 'float32'):
    d = Domain()
    x = L * torch.arange(nx, dtype=torch.float32) / nx
    d.variable('x', x, None, dtype=dtype)
    d.variable('t', [0, 1], nt - 1, dtype=dtype)
    b = Conditions()
    b.periodic([{'x': 0, 't': [0, 1]}, {'x': L, 't': [0, 1]}])
    b.dirichlet({'x': [0, L], 't': 0}, value=0.5 / torch.cosh(0.5 * (x - L / 2)) ** 2)
    eq = Equation()
    eq.add({
        'u_t': {'coeff': 1, 'u_t': [1], 'pow': 1, 'var': 0},
        '6u*u_x': {'coeff': 6, 'u*u_x': [[None], [0]], 'pow': [1, 1], 'var': [0, 0]},
        'u_xxx': {'coeff': 1, 'u_xxx': [0, 0, 0], 'pow': 1, 'var': 0},
    })
    net = tanh_mlp([2, 64, 64, 64, 64, 1], seed,
                   dtype=torch.float64 if dtype == 'float64' else torch.float32)
    return d, eq, b, net, dict(mode='NN', lambda_operator=1, lambda_bound=100, h=0.01)


@dataclass
class SweepInstance:
    nu: float
    nx: int
    nt: int
    seed: int


def sweep_instances(n: int = 256, seed: int = 0) -> List[SweepInstance]:
    """nu = 10^U(-3,-1), nx ~ U_int[64, 512], nt ~ U_int[32, 128], drawn in
    that order per instance from one generator seeded with ``seed``."""
    g = torch.Generator().manual_seed(seed)
    out = []
    for i in range(n):
        nu = float(10 ** (-3 + 2 * torch.rand(1, generator=g, dtype=torch.float64)))
        nx = int(torch.randint(64, 513, (1,), generator=g))
        nt = int(torch.randint(32, 129, (1,), generator=g))
        out.append(SweepInstance(nu, nx, nt, seed + i))
    return out


def sweep_problem(inst: SweepInstance, dtype: str = 'float32'):
    return burgers(nx=inst.nx, nt=inst.nt, nu=inst.nu, layers=(2, 32, 32, 32, 1), seed=inst.seed,
                   dtype=dtype)


def collocation_points(nx: int, nt: int) -> int:
    """'central' (strict interior) points of an nx x nt grid."""
    return (nx - 2) * (nt - 2)
