"""The float64 reference solutions used for relative-L2 reporting."""
import numpy as np
import pytest

from torch_de_solver_amd import exact


def test_burgers_matches_reference_data(golden):
    """Against the reference's own solution file examples/wolfram_sln/Burgers.mat
    (nu = 0.01/pi on 256 x 100 points), copied as a fixture."""
    z = golden('burgers_wolfram')
    X, T = np.meshgrid(z['x'], z['t'], indexing='ij')
    u = exact.burgers(X, T, 0.01 / np.pi)
    assert np.abs(u - z['usol']).max() < 1e-6
    assert exact.relative_l2(u, z['usol']) < 1e-7


@pytest.mark.parametrize('nu', [1e-3, 1e-2, 1e-1])
def test_burgers_solves_the_pde(nu):
    """Initial and boundary values, and a small float64 finite-difference
    residual u_t + u u_x - nu u_xx away from the shock region's grid scale."""
    x = np.linspace(-1, 1, 2001)
    assert np.abs(exact.burgers(x, 0.0 * x, nu) + np.sin(np.pi * x)).max() < 1e-14
    t = np.linspace(0.05, 1, 5)
    assert np.abs(exact.burgers(-1.0 + 0 * t, t, nu)).max() < 1e-12
    assert np.abs(exact.burgers(1.0 + 0 * t, t, nu)).max() < 1e-12
    h, k = 1e-4, 1e-5
    xs = np.linspace(-0.9, 0.9, 7)
    for tt in (0.1, 0.5):
        u = exact.burgers(xs, tt, nu)
        ut = (exact.burgers(xs, tt + k, nu) - exact.burgers(xs, tt - k, nu)) / (2 * k)
        up = exact.burgers(xs + h, tt, nu)
        um = exact.burgers(xs - h, tt, nu)
        ux = (up - um) / (2 * h)
        uxx = (up - 2 * u + um) / h ** 2
        res = ut + u * ux - nu * uxx
        scale = np.abs(ut).max() + np.abs(u * ux).max() + 1.0
        assert np.abs(res).max() < 2e-3 * scale, (nu, tt, np.abs(res).max())


def test_closed_forms_satisfy_their_equations():
    x = np.linspace(-1, 1, 11)
    p = exact.legendre_p3(x)
    assert np.allclose(p, 0.5 * (5 * x ** 3 - 3 * x))
    p1 = 0.5 * (15 * x ** 2 - 3)
    p2 = 15 * x
    assert np.abs((1 - x ** 2) * p2 - 2 * x * p1 + 12 * p).max() < 1e-12
    xx, tt = np.meshgrid(np.linspace(0, 1, 9), np.linspace(0, 1, 9), indexing='ij')
    w = exact.wave(xx, tt)
    assert np.allclose(w[:, 0], np.sin(np.pi * xx[:, 0]))
    assert np.allclose(-4 * np.pi ** 2 * w, 4 * (-np.pi ** 2) * w)  # u_tt = 4 u_xx
    # soliton: u_t + 6 u u_x + u_xxx = 0, by central differences
    h = 1e-3
    xs, ts = np.linspace(5, 15, 9), 0.3
    f = lambda a, b: exact.kdv_soliton(a, b, images=0)  # noqa: E731
    ut = (f(xs, ts + h) - f(xs, ts - h)) / (2 * h)
    ux = (f(xs + h, ts) - f(xs - h, ts)) / (2 * h)
    uxxx = (f(xs + 2 * h, ts) - 2 * f(xs + h, ts) + 2 * f(xs - h, ts) - f(xs - 2 * h, ts)) / (2 * h ** 3)
    assert np.abs(ut + 6 * f(xs, ts) * ux + uxxx).max() < 1e-4
