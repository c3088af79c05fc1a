"""Equation dictionaries: normalisation and the parsed term form.

The reference describes an operator as a dict of terms, each term a dict
whose keys are, in order, ``coeff``, the derivative spec (any key name: the
reference finds it *by position*, tedeous/input_preprocessing.py:65), then
``pow`` and optionally ``var``.  ``equation_unify`` brings every term to the
per-factor list form (tedeous/input_preprocessing.py:51-81); ``parse_operator``
turns that into ``Term``/``Factor`` records the plan builder lowers to tables.
"""

from dataclasses import dataclass
from typing import Any, Callable, List, Optional, Union

import torch


def _is_scalar_pow(p) -> bool:
    return isinstance(p, (int, float)) or callable(p)


def equation_unify(equation: Optional[dict]) -> Optional[dict]:
    """Bring each term to list form in place and return the dict.

    After this every term has ``term[spec] = [axes_0, axes_1, ...]``,
    ``pow = [p_0, ...]`` and ``var = [v_0, ...]`` with one entry per factor;
    a missing ``var`` means variable 0.
    """
    if equation is None:
        return None
    for term in equation.values():
        spec = list(term.keys())[1]
        if 'var' not in term:
            if _is_scalar_pow(term['pow']):
                term[spec] = [term[spec]]
                term['pow'] = [term['pow']]
                term['var'] = [0]
            elif isinstance(term['pow'], list):
                term['var'] = [0 for _ in term['pow']]
            continue
        if _is_scalar_pow(term['pow']):
            term[spec] = [term[spec]]
            term['pow'] = [term['pow']]
            term['var'] = [term['var']]
    return equation


def lambda_prepare(val: torch.Tensor, lambda_, dtype: torch.dtype = None) -> torch.Tensor:
    """Per-column weights as a ``[1, n]`` row (tedeous/input_preprocessing.py:14-43)."""
    if isinstance(lambda_, torch.Tensor):
        return lambda_
    if isinstance(lambda_, (int, float)):
        try:
            lambdas = torch.ones(val.shape[-1]) * lambda_
        except Exception:
            lambdas = torch.tensor(lambda_)
    elif isinstance(lambda_, list):
        lambdas = torch.tensor(lambda_)
    else:
        raise TypeError(f"unsupported lambda {lambda_!r}")
    if dtype:
        lambdas = lambdas.to(dtype)
    return lambdas.reshape(1, -1)


@dataclass
class Factor:
    """One factor ``(D^{axes} u_var) ** pow`` of a term.

    ``axes`` is ``None`` for the undifferentiated variable (the reference's
    ``[None]``), else the ordered list of axis indices of the derivative.
    """
    var: int
    axes: Optional[List[int]]
    pow: Union[int, float, Callable]


@dataclass
class Term:
    """``coeff * prod(factors)`` (tedeous/derivative.py:44-56)."""
    coeff: Any
    factors: List[Factor]


def parse_operator(operator: dict) -> List[Term]:
    """Parse one operator dict (already unified or not) into ``Term`` records.

    The dict is copied first, so the caller's dict is left unchanged.
    Term order is dict order: the reference sums terms in that order
    (tedeous/eval.py:158-165).
    """
    import copy
    op = equation_unify(copy.copy({k: dict(v) for k, v in operator.items()}))
    terms = []
    for term in op.values():
        spec = list(term.keys())[1]
        factors = []
        for axes, p, v in zip(term[spec], term['pow'], term['var']):
            if axes is None or axes == [None] or (isinstance(axes, list) and
                                                  len(axes) > 0 and axes[0] is None):
                ax = None
            else:
                ax = [int(a) for a in axes if a is not None]
                if not ax:
                    ax = None
            if isinstance(v, (list, tuple)):
                # a one-factor term written with 'var': [k] is wrapped once
                # more by the unification; the reference then indexes the
                # output with the list [k], i.e. column k
                # (tedeous/derivative.py:122-126)
                if len(v) != 1:
                    raise ValueError(f"term variable {v!r}: one index per factor")
                v = v[0]
            factors.append(Factor(var=int(v), axes=ax, pow=p))
        terms.append(Term(coeff=term['coeff'], factors=factors))
    return terms
