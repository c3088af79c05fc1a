"""Problem definition API: ``Domain``, ``Conditions``, ``Equation``.

Mirrors the public surface of tedeous/data.py (Domain 33-107,
Conditions 110-323, Equation 326-339) so that user scripts written for the
reference run unchanged.  Grid values are produced with the same torch
routines (``linspace``, ``cartesian_prod``, an ``ij`` meshgrid) so that the
coordinates are bit-identical to the reference's.
"""

from copy import deepcopy
from typing import Callable, Dict, List, Union

import numpy as np
import torch

from .device import check_device
from .equations import equation_unify


def tensor_dtype(dtype: Union[str, torch.dtype]) -> torch.dtype:
    """Map the reference's dtype strings to torch dtypes (tedeous/data.py:14-30)."""
    table = {'float32': torch.float32, 'float64': torch.float64,
             'float16': torch.float16}
    return table.get(dtype, dtype)


class Domain:
    """Axes of a tensor-product grid (tedeous/data.py:33-107)."""

    def __init__(self, type: str = 'uniform'):
        self.type = type
        self.variable_dict: Dict[str, torch.Tensor] = {}

    def variable(self, variable_name: str,
                 variable_set: Union[List, torch.Tensor],
                 n_points: Union[None, int],
                 dtype: str = 'float32') -> None:
        """Declare one axis.

        A ``[start, end]`` pair gives ``n_points + 1`` uniformly spaced values
        (tedeous/data.py:62-67: the reference adds one point); a tensor is taken
        as the axis values unchanged (tedeous/data.py:58-61).
        """
        dt = tensor_dtype(dtype)
        if isinstance(variable_set, torch.Tensor):
            self.variable_dict[variable_name] = variable_set.to(dt)
        elif self.type == 'uniform':
            start, end = variable_set
            self.variable_dict[variable_name] = torch.linspace(
                start, end, n_points + 1, dtype=dt)
        else:
            raise ValueError(f"unknown domain type {self.type!r}")

    def axes(self) -> List[torch.Tensor]:
        return [v.detach().cpu() for v in self.variable_dict.values()]

    def build(self, mode: str, removed_domains: list = None) -> torch.Tensor:
        """Grid for ``mode`` (tedeous/data.py:69-107).

        'NN' / 'autograd': points ``[N, d]`` from ``cartesian_prod`` (first
        variable slowest).  'mat': an ``ij`` meshgrid stacked as ``[d, n1, ...]``.
        """
        if removed_domains is not None:
            raise NotImplementedError(
                "removed_domains (CSG holes) is not supported by this engine")
        axes = self.axes()
        if mode in ('NN', 'autograd'):
            if len(axes) == 1:
                grid = axes[0].reshape(-1, 1)
            else:
                grid = torch.cartesian_prod(*axes)
        else:
            mesh = np.meshgrid(*[a.numpy() for a in axes], indexing='ij')
            grid = torch.from_numpy(np.stack(mesh))
        return check_device(grid)


class Conditions:
    """Boundary, initial and data conditions (tedeous/data.py:110-323).

    Each condition is stored as the reference stores it: a dict with keys
    ``bnd``, ``bop``, ``bval``, ``var`` and ``type``.
    """

    def __init__(self):
        self.conditions_lst: List[dict] = []

    def dirichlet(self, bnd, value, var: int = 0):
        self.conditions_lst.append({'bnd': bnd, 'bop': None, 'bval': value,
                                    'var': var, 'type': 'dirichlet'})

    def operator(self, bnd, operator: dict, value):
        # The reference looks the variable up with ``operator.keys()[0]``,
        # which always raises and so always falls back to variable 0
        # (tedeous/data.py:151-154).  That behaviour is kept.
        self.conditions_lst.append({'bnd': bnd,
                                    'bop': equation_unify(operator),
                                    'bval': value, 'var': 0,
                                    'type': 'operator'})

    def periodic(self, bnd, operator: dict = None, var: int = 0):
        # Periodic conditions compare values between point sets and have a
        # zero target (tedeous/data.py:178).
        value = torch.tensor([0.])
        if operator is not None:
            operator = equation_unify(operator)
            var = 0
        self.conditions_lst.append({'bnd': bnd, 'bop': operator,
                                    'bval': value, 'var': var,
                                    'type': 'periodic'})

    def robin(self, bnd, value, operator: dict = None, var: int = 0):
        self.conditions_lst.append({'bnd': bnd,
                                    'bop': equation_unify(operator),
                                    'bval': value, 'var': var,
                                    'type': 'robin'})

    def data(self, bnd, operator, value, var: int = 0):
        if operator is not None:
            operator = equation_unify(operator)
        self.conditions_lst.append({'bnd': bnd, 'bop': operator,
                                    'bval': value, 'var': var,
                                    'type': 'data'})

    @staticmethod
    def _bnd_grid(bnd, variable_dict: dict) -> torch.Tensor:
        """Point set of one condition (tedeous/data.py:241-286).

        A dict pins scalar coordinates, filters an axis to ``[lo, hi]`` for a
        list, or takes a tensor of values; the product of those per-axis sets
        is the point set.
        """
        dtype = next(iter(variable_dict.values())).dtype
        if isinstance(bnd, torch.Tensor):
            pts = bnd.to(dtype)
        else:
            if 'circle' in bnd:
                raise NotImplementedError("CSG circle boundaries are not supported")
            per_axis = []
            for name, axis in variable_dict.items():
                spec = bnd[name]
                if isinstance(spec, torch.Tensor):
                    per_axis.append(check_device(spec).to(dtype))
                elif isinstance(spec, (int, float)):
                    per_axis.append(check_device(torch.tensor([spec])).to(dtype))
                elif isinstance(spec, list):
                    lo, hi = spec
                    sel = axis[(axis >= lo) & (axis <= hi)]
                    per_axis.append(check_device(sel).to(dtype))
                else:
                    raise TypeError(f"bad boundary spec for {name!r}: {spec!r}")
            pts = torch.cartesian_prod(*per_axis).to(dtype)
        if pts.dim() == 1:
            pts = pts.reshape(-1, 1)
        return pts

    def build(self, variable_dict: dict) -> Union[None, List[dict]]:
        """Resolve point sets and values (tedeous/data.py:288-323).

        Returns new dicts; the stored conditions are left as given so that
        ``build`` can be called more than once.
        """
        if not self.conditions_lst:
            return None
        dtype = next(iter(variable_dict.values())).dtype
        built = []
        for cond in self.conditions_lst:
            c = dict(cond)
            c['bop'] = deepcopy(cond['bop'])
            if c['type'] == 'periodic':
                c['bnd'] = [self._bnd_grid(b, variable_dict) for b in cond['bnd']]
            else:
                c['bnd'] = self._bnd_grid(cond['bnd'], variable_dict)
            val = cond['bval']
            if isinstance(val, torch.Tensor):
                c['bval'] = check_device(val).to(dtype)
            elif isinstance(val, (int, float)):
                c['bval'] = check_device(torch.ones_like(c['bnd'][:, 0]) * val).to(dtype)
            elif callable(val):
                c['bval'] = check_device(val(c['bnd'])).to(dtype)
            built.append(c)
        return built


class Equation:
    """List of operator dicts, one per equation (tedeous/data.py:326-339)."""

    def __init__(self):
        self.equation_lst: List[dict] = []

    def add(self, eq: dict) -> None:
        self.equation_lst.append(eq)
