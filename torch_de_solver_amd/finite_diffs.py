"""Finite-difference stencil tables for both solution modes.

'NN' mode (tedeous/finite_diffs.py:9-268): a derivative along axes
``[a_1, ..., a_k]`` expands recursively, one axis at a time, into integer
shift vectors (multiples of ``h``) with weights:

* first-order scheme ('1'): central ``{+e_a: +c, -e_a: -c}`` with
  ``c = 1/(2h)``; forward ``{+e_a: +1/h, 0: -1/h}``; backward
  ``{0: +1/h, -e_a: -1/h}``.  Duplicate shifts are kept, so a second
  central derivative is ``(u(x+2h) - 2u(x) + u(x-2h)) / (4h^2)`` with step
  ``2h`` (tedeous/finite_diffs.py:59-117).
* second-order boundary scheme ('2', boundary point types only): forward
  ``{+2e: 3c, +e: -4c, 0: c}``, backward ``{-2e: -3c, -e: 4c, 0: -c}``
  (tedeous/finite_diffs.py:145-223).

Weights are formed in double precision with the same sequence of products
and quotients as the reference, so that after rounding to float32 they are
the same numbers the reference multiplies by.

'mat' mode (tedeous/derivative.py:157-227): one application of ``D`` along
an axis uses ``dp`` points each side; the coefficient vectors come from a
Vandermonde solve.  Interior points use ``(back + forw) / (2h)``; the first
``dp-1`` points use ``forw / h`` and the last ``dp-1`` use ``back / h``.
"""

from typing import List, Optional, Sequence, Tuple

import numpy as np


Shift = Tuple[int, ...]


def nn_scheme(axes: Optional[Sequence[int]], nvars: int, point_type: str,
              order_label: str, h: float) -> List[Tuple[Shift, float]]:
    """Shift/weight list of one factor, in the reference's evaluation order.

    ``axes`` is ``None`` for an undifferentiated factor, which evaluates the
    network once at the point with weight 1 (tedeous/finite_diffs.py:257-258).
    ``point_type`` is ``'central'`` or a per-axis string of ``'f'``/``'b'``.
    """
    if axes is None:
        return [(tuple([0] * nvars), 1.0)]
    axes = list(axes)
    if order_label == '1':
        if point_type == 'central':
            dirs = ['central'] * len(axes)
        else:
            dirs = [point_type[a] for a in axes]
        all_central = sorted(set(dirs))[0] == 'central'
        entries: List[Tuple[list, float]] = [([0] * nvars, 1)]
        for a, d in zip(axes, dirs):
            nxt = []
            for shift, w in entries:
                plus, minus = list(shift), list(shift)
                if d == 'central':
                    plus[a] += 1
                    minus[a] -= 1
                elif d == 'f':
                    plus[a] += 1
                elif d == 'b':
                    minus[a] -= 1
                if all_central:
                    c = 1 / (2 * h)
                    nxt.append((plus, w * c))
                    nxt.append((minus, -w * c))
                else:
                    nxt.append((plus, w / h))
                    nxt.append((minus, -w / h))
            entries = nxt
    elif order_label == '2':
        if point_type == 'central':
            # The reference indexes the characters of the string 'central'
            # here and then fails on an empty weight list
            # (tedeous/finite_diffs.py:138-143, 166-167, 210-222).
            raise ValueError("inner_order='2' is not a usable scheme: the "
                             "second-order scheme exists for boundary point "
                             "types only")
        dirs = [point_type[a] for a in axes]
        c = 1 / (2 * h)
        entries = [([0] * nvars, 1)]
        for a, d in zip(axes, dirs):
            nxt = []
            for shift, w in entries:
                s2, s1, s0 = list(shift), list(shift), list(shift)
                if d == 'f':
                    s2[a] += 2
                    s1[a] += 1
                    nxt += [(s2, 3 * c * w), (s1, -4 * c * w), (s0, c * w)]
                elif d == 'b':
                    s2[a] -= 2
                    s1[a] -= 1
                    nxt += [(s2, -3 * c * w), (s1, 4 * c * w), (s0, -c * w)]
                else:
                    raise ValueError(f"bad direction {d!r}")
            entries = nxt
    else:
        raise ValueError(f"unknown scheme order {order_label!r}")
    return [(tuple(s), float(w)) for s, w in entries]


def mat_alphas(derivative_points: int) -> Tuple[np.ndarray, np.ndarray,
                                                 List[int], List[int]]:
    """Backward/forward coefficient vectors and their offsets.

    Offsets are ``[-(dp-1), ..., 0]`` (backward) and ``[0, ..., dp-1]``
    (forward).  The coefficients solve ``sum_j alpha_j * l_j**i = [i == 1]``
    for ``i < dp`` (tedeous/derivative.py:157-197): the one-sided first
    derivative exact on polynomials of degree ``< dp``.
    """
    dp = int(derivative_points)
    if dp < 2:
        raise ValueError("derivative_points must be >= 2")
    back = list(range(-dp + 1, 1))
    forw = list(range(dp))

    def solve(labels):
        lab = np.array(labels, dtype=np.float64)
        A = np.vstack([lab ** i for i in range(len(labels))])
        b = np.zeros(len(labels))
        b[1] = 1.0
        return np.linalg.solve(A, b)

    return solve(back), solve(forw), back, forw
