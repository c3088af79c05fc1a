"""Derivative strategies of the reference's plugin interface
(tedeous/derivative.py:10-363).

``Derivative(model, derivative_points).set_strategy(mode)`` returns an
object whose ``take_derivative(term, grid_points)`` gives the value of one
prepared term (``Equation_*.operator_prepare()`` output) as ``[n, 1]``
('NN', 'autograd') or as a field shaped like the grid ('mat'), computed
with torch from the caller's model:

* 'NN': ``coeff * prod_j (sum_k w_jk * model(grid_jk)[:, var_j]) ** pow_j``
  over the materialised shifted grids (tedeous/derivative.py:30-58);
* 'autograd': each factor's derivative by repeated ``torch.autograd.grad``
  of the summed output along its axes (tedeous/derivative.py:73-132);
* 'mat': each factor's axis applications of the ``derivative_points``
  stencil, one-sided ``/h`` on the first and last ``dp - 1`` positions of
  the axis, central ``(back + forw)/(2h)`` inside (tedeous/derivative.py:
  135-323).

A callable ``pow`` is applied to the running product, as the reference
does.  These objects serve code written against the plugin interface; the
training path (``Solution.evaluate``) runs the native engines instead.
"""

from typing import Callable, List, Union

import torch

from .finite_diffs import mat_alphas


def _apply_pow(prod, value, p):
    if isinstance(p, (int, float)):
        return prod * value ** p
    if callable(p):
        return p(prod * value)
    raise TypeError(f"pow {p!r}")


class DerivativeInt:
    def take_derivative(self, value):
        raise NotImplementedError


class Derivative_NN(DerivativeInt):
    def __init__(self, model: torch.nn.Module):
        self.model = model

    def take_derivative(self, term: dict, *args) -> torch.Tensor:
        key = list(term.keys())[1]
        coeff = term['coeff']
        if isinstance(coeff, tuple):
            coeff = coeff[0](coeff[1]).reshape(-1, 1)
        grids, weights = term[key]
        out = 1.
        for j, (gl, wl) in enumerate(zip(grids, weights)):
            s = 0.
            for g, w in zip(gl, wl):
                s = s + self.model(g)[:, term['var'][j]].reshape(-1, 1) * w
            out = _apply_pow(out, s, term['pow'][j])
        return coeff * out


class Derivative_autograd(DerivativeInt):
    def __init__(self, model: torch.nn.Module):
        self.model = model

    @staticmethod
    def _nn_autograd(model, points: torch.Tensor, var, axis: List[int] = [0]):
        points.requires_grad = True
        f = model(points)[:, var].sum(0)
        g = None
        for ax in axis:
            g, = torch.autograd.grad(f, points, create_graph=True)
            f = g[:, ax].sum()
        return g[:, axis[-1]].reshape(-1, 1)

    def take_derivative(self, term: dict, grid_points: torch.Tensor) -> torch.Tensor:
        key = list(term.keys())[1]
        coeff = term['coeff']
        if callable(coeff) and not isinstance(coeff, torch.Tensor):
            coeff = coeff(grid_points).reshape(-1, 1)
        out = 1.
        for j, axes in enumerate(term[key]):
            if axes == [None]:
                d = self.model(grid_points)[:, term['var'][j]].reshape(-1, 1)
            else:
                d = self._nn_autograd(self.model, grid_points, term['var'][j], axes)
            out = _apply_pow(out, d, term['pow'][j])
        return coeff * out


class Derivative_mat(DerivativeInt):
    def __init__(self, model: torch.Tensor, derivative_points: int):
        self.model = model
        self.derivative_points = int(derivative_points)
        ab, af, self.backward, self.farward = mat_alphas(self.derivative_points)
        self.alpha_backward, self.alpha_farward = ab, af

    @staticmethod
    def _steps(grid: torch.Tensor) -> List[torch.Tensor]:
        """Step of every axis of an ``ij`` meshgrid ``[d, n1, ...]``: the
        difference of its first two distinct values."""
        pts = torch.vstack([grid[i].reshape(-1) for i in range(grid.shape[0])]).T.float()
        out = []
        for i in range(pts.shape[-1]):
            v = torch.unique(pts[:, i])
            out.append(abs(v[1] - v[0]))
        return out

    def _along(self, u: torch.Tensor, h, axis: int) -> torch.Tensor:
        """One derivative application along ``axis`` of the field ``u``."""
        if u.dim() == 1 or u.shape[0] == 1:
            u1 = u.reshape(-1)
            return self._line(u1, h).reshape(u.shape)
        if u.dim() > 3:
            raise NotImplementedError("'mat' fields of more than three axes")
        v = u.movedim(axis, -1)
        return self._line(v, h).movedim(-1, axis)

    def _line(self, v: torch.Tensor, h) -> torch.Tensor:
        """Stencil along the last dimension: one-sided at the ends."""
        n = v.shape[-1]
        dp = self.derivative_points
        pad = dp - 1
        back = torch.zeros_like(v)
        forw = torch.zeros_like(v)
        for a, off in zip(self.alpha_backward, self.backward):
            back = back + torch.roll(v, -off, dims=-1) * float(a)
        for a, off in zip(self.alpha_farward, self.farward):
            forw = forw + torch.roll(v, -off, dims=-1) * float(a)
        du = (back + forw) / (2 * h)
        if pad:
            du[..., n - pad:] = back[..., n - pad:] / h
            du[..., :pad] = forw[..., :pad] / h
        return du

    def take_derivative(self, term: dict, grid_points: torch.Tensor) -> torch.Tensor:
        key = list(term.keys())[1]
        out = torch.zeros_like(self.model) + 1
        steps = None
        for j, axes in enumerate(term[key]):
            prod = self.model[term['var'][j]]
            if axes != [None]:
                for ax in axes:
                    if ax is None:
                        continue
                    if steps is None:
                        steps = self._steps(grid_points)
                    prod = self._along(prod, steps[ax], ax)
            out = _apply_pow(out, prod, term['pow'][j])
        coeff = term['coeff']
        if callable(coeff) and not isinstance(coeff, torch.Tensor):
            return coeff(grid_points) * out
        return coeff * out


class Derivative:
    """Strategy dispatch (tedeous/derivative.py:326-363)."""

    def __init__(self, model: Union[torch.nn.Module, torch.Tensor], derivative_points: int):
        self.model = model
        self.derivative_points = derivative_points

    def set_strategy(self, strategy: str) -> Union[Derivative_NN, Derivative_autograd,
                                                   Derivative_mat]:
        if strategy == 'NN':
            return Derivative_NN(self.model)
        if strategy == 'autograd':
            return Derivative_autograd(self.model)
        if strategy == 'mat':
            return Derivative_mat(self.model, self.derivative_points)
        raise NotImplementedError(f"mode {strategy!r} is not served by this engine")
