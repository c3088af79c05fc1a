"""Equation preparation objects of the reference's plugin interface
(tedeous/input_preprocessing.py:173-644).

``Operator_bcond_preproc(grid, operator, bconds, ...).set_strategy(mode)``
returns one of ``Equation_NN``, ``Equation_mat``, ``Equation_autograd``.
Each keeps the compile inputs (``grid``, ``operator``, ``bconds``, ``h``,
``inner_order``, ``boundary_order``, ``mode``) that ``Solution`` lowers to
kernel tables, and offers the reference's ``operator_prepare()`` and
``bnd_prepare()``, which return the operator and the conditions in the
reference's prepared form for code written against that interface (a
custom derivative strategy, the reference's ``Operator``):

* 'NN': per term, ``coeff`` (scalar, ``(fn, points)`` for a callable, the
  gathered ``[n, 1]`` values for a tensor) and ``[shifted grids per factor,
  weights per factor]`` on the 'central' points; a condition operator per
  point type of its points (a list of such lists for a periodic one);
* 'mat': the unified operator; condition points as per-point index tuples
  into the meshgrid;
* 'autograd': the unified operator with checked coefficients; conditions as
  given.

The prepared structures are built from this package's own stencil tables
(``finite_diffs.nn_scheme``) and point typing (``points_type``); the
structures are copies, so preparing never changes the operator the solver
lowers.  None of this is on the training path: ``Solution.evaluate`` runs
the native engines.
"""

import copy
from typing import List, Optional, Union

import numpy as np
import torch

from .device import check_device
from .equations import equation_unify
from .finite_diffs import nn_scheme
from .points_type import grid_sort

MODES = ('NN', 'mat', 'autograd')


def _copy_operator(op: dict) -> dict:
    """Two-level copy: new term dicts, coefficient objects shared."""
    return {k: dict(v) for k, v in op.items()}


def _operator_list(operator) -> List[dict]:
    if isinstance(operator, list) and operator and isinstance(operator[0], dict):
        return list(operator)
    return [operator]


def _row_positions(grid: torch.Tensor, pts: torch.Tensor) -> torch.Tensor:
    """Grid row of every point (exact coordinate match)."""
    g = grid.detach().cpu().numpy()
    lut = {tuple(r): i for i, r in enumerate(g.tolist())}
    pos = [lut.get(tuple(p)) for p in pts.detach().cpu().numpy().tolist()]
    if any(p is None for p in pos):
        raise ValueError("a coefficient point is not a grid point")
    return torch.as_tensor(pos, dtype=torch.long, device=grid.device)


class _Prepared:
    """Compile inputs of one problem in one mode."""

    mode: Optional[str] = None

    def __init__(self, grid: torch.Tensor, operator, bconds, h: float = 0.001,
                 inner_order: str = '1', boundary_order: str = '2'):
        self.grid = grid
        self.operator = operator
        self.bconds = bconds
        self.h = h
        self.inner_order = inner_order
        self.boundary_order = boundary_order

    def set_strategy(self, strategy: str):
        """Re-tag the same inputs with another mode (the holder's own
        interface, kept for code that calls it on a prepared object)."""
        return Operator_bcond_preproc(self.grid, self.operator, self.bconds, self.h,
                                      self.inner_order, self.boundary_order
                                      ).set_strategy(strategy)


class Equation_NN(_Prepared):
    """'NN' preparation (tedeous/input_preprocessing.py:173-408)."""

    mode = 'NN'

    def _types(self):
        g = self.grid.detach().cpu().numpy()
        return {k: self.grid[torch.as_tensor(v, device=self.grid.device)]
                for k, v in grid_sort(g).items()}

    def _coeff(self, coeff, points: torch.Tensor):
        # order of the checks as the reference's _checking_coeff
        # (input_preprocessing.py:266-295): a Parameter is a tensor first
        if isinstance(coeff, (int, float)):
            return coeff
        if callable(coeff) and not isinstance(coeff, torch.Tensor):
            return (coeff, points)
        if isinstance(coeff, torch.Tensor):
            c = check_device(coeff)
            return c.reshape(-1)[_row_positions(self.grid, points).to(c.device)].reshape(-1, 1)
        raise NameError('"coeff" should be: torch.Tensor or callable or int or float!')

    def _shifted(self, points: torch.Tensor, axes, point_type: str):
        """[grids], [weights] of one factor on the points of one type."""
        if axes is None or axes == [None]:
            return [points], [1]
        order = self.inner_order if point_type == 'central' else self.boundary_order
        nvars = points.shape[-1]
        grids, weights = [], []
        for shift, w in nn_scheme(list(axes), nvars, point_type, order, self.h):
            s = points.clone()
            for j, k in enumerate(shift):
                if k:
                    s[:, j] = points[:, j] + k * self.h
            grids.append(s)
            weights.append(w)
        return grids, weights

    def _one(self, operator: dict, points: torch.Tensor, point_type: str) -> dict:
        op = equation_unify(_copy_operator(operator))
        for label in op:
            term = op[label]
            key = list(term.keys())[1]
            term['coeff'] = self._coeff(term['coeff'], points)
            grids, weights = [], []
            for axes in term[key]:
                g, w = self._shifted(points, axes, point_type)
                grids.append(g)
                weights.append(w)
            term[key] = [grids, weights]
        return op

    def operator_prepare(self) -> List[dict]:
        central = self._types()['central']
        return [self._one(op, central, 'central') for op in _operator_list(self.operator)]

    def _by_type(self, types: dict, pts: torch.Tensor) -> dict:
        """Points of ``pts`` grouped by the type of the equal grid point,
        in ``pts`` order within each type."""
        out = {}
        for t, v in types.items():
            vn = v.detach().cpu().numpy()
            keys = {tuple(r) for r in vn.tolist()}
            sel = [i for i, p in enumerate(pts.detach().cpu().numpy().tolist())
                   if tuple(p) in keys]
            if sel:
                out[t] = pts[torch.as_tensor(sel, device=pts.device)]
        return out

    def bnd_prepare(self) -> List[dict]:
        types = self._types()
        out = []
        for bc in self.bconds:
            bc = dict(bc)
            if bc['bop'] is not None:
                if bc['type'] == 'periodic':
                    bc['bop'] = [[self._one(bc['bop'], p, t)
                                  for t, p in self._by_type(types, b).items()]
                                 for b in bc['bnd']]
                else:
                    bc['bop'] = [self._one(bc['bop'], p, t)
                                 for t, p in self._by_type(types, bc['bnd']).items()]
            out.append(bc)
        return out


class Equation_autograd(_Prepared):
    """'autograd' preparation (tedeous/input_preprocessing.py:411-505)."""

    mode = 'autograd'

    @staticmethod
    def _coeff(coeff):
        if isinstance(coeff, (int, float)):
            return coeff
        if callable(coeff) and not isinstance(coeff, torch.Tensor):
            return coeff
        if isinstance(coeff, torch.Tensor):
            return check_device(coeff).reshape(-1, 1)
        raise NameError('"coeff" should be: torch.Tensor or callable or int or float!')

    def operator_prepare(self) -> List[dict]:
        out = []
        for op in _operator_list(self.operator):
            op = equation_unify(_copy_operator(op))
            for label in op:
                op[label]['coeff'] = self._coeff(op[label]['coeff'])
            out.append(op)
        return out

    def bnd_prepare(self) -> Optional[List[dict]]:
        return None if self.bconds is None else [dict(bc) for bc in self.bconds]


class Equation_mat(_Prepared):
    """'mat' preparation (tedeous/input_preprocessing.py:511-594)."""

    mode = 'mat'

    def operator_prepare(self) -> List[dict]:
        return [equation_unify(_copy_operator(op)) for op in _operator_list(self.operator)]

    def _positions(self, pts: torch.Tensor) -> list:
        """Per point, the index tuple of the grid positions whose
        coordinates are all ``isclose`` to it; in 1-D the ``isclose`` match
        (else the nearest point) as a one-element tuple."""
        g = self.grid
        out = []
        if g.shape[0] == 1:
            gd = g.reshape(-1)
            for p in pts:
                m = torch.nonzero(torch.isclose(gd.double(), p.double().reshape(-1)[0])).reshape(-1)
                i = int(m[0]) if m.numel() == 1 else \
                    int(torch.argmin((gd.double() - p.double().reshape(-1)[0]).abs()))
                out.append((torch.tensor(i),))
            return out
        for p in pts:
            hit = torch.ones_like(g[0], dtype=torch.bool)
            for a in range(g.shape[0]):
                hit = hit & torch.isclose(p[a].float(), g[a].float())
            out.append(torch.where(hit))
        return out

    def bnd_prepare(self) -> List[dict]:
        out = []
        for bc in self.bconds:
            bc = dict(bc)
            if bc['type'] == 'periodic':
                bc['bnd'] = [self._positions(b) for b in bc['bnd']]
            else:
                bc['bnd'] = self._positions(bc['bnd'])
            if bc['bop'] is not None:
                bc['bop'] = equation_unify(_copy_operator(bc['bop']))
            out.append(bc)
        return out


class Operator_bcond_preproc:
    """Mode dispatch of the preparation (tedeous/input_preprocessing.py:597-644)."""

    def __init__(self, grid: torch.Tensor, operator: Union[dict, list], bconds: list,
                 h: float = 0.001, inner_order: str = '1', boundary_order: str = '2'):
        self.grid = check_device(grid)
        self.operator = operator
        self.bconds = bconds
        self.h = h
        self.inner_order = inner_order
        self.boundary_order = boundary_order
        self.mode = None

    def set_strategy(self, strategy: str) -> Union[Equation_NN, Equation_mat, Equation_autograd]:
        cls = {'NN': Equation_NN, 'mat': Equation_mat, 'autograd': Equation_autograd}.get(strategy)
        if cls is None:
            raise NotImplementedError(f"mode {strategy!r} is not served by this engine")
        return cls(self.grid, self.operator, self.bconds, self.h, self.inner_order,
                   self.boundary_order)
