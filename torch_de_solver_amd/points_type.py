"""Point typing for 'NN' mode (tedeous/points_type.py:9-157).

The reference shifts every point by +-1e-4 along each axis and asks whether
the shifted point stays in the convex hull of the grid (a qhull Delaunay
test in d>1, a bounds test in d=1).  A point whose 2d shifts all stay in is
``'central'``; otherwise its type has one character per axis, ``'f'`` when
the + shift stays in and ``'b'`` when it leaves.  In 1-D every point is
``'central'`` (tedeous/points_type.py:102-103).

Grids here are tensor products of axis values, whose hull is the bounding
box, so the hull test is a per-axis bounds test on the float32 shifted
coordinate.  ``types_by_hull`` runs the reference's qhull formulation for
arbitrary point clouds and is what the tests compare the box form against.
"""

from typing import Dict, List

import numpy as np

SHIFT = np.float32(1e-4)


def _inside_box(p: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    return np.all((p >= lo) & (p <= hi), axis=1)


def point_types(grid: np.ndarray) -> np.ndarray:
    """Type string of every grid point (numpy array of str)."""
    g = np.asarray(grid, dtype=np.float32)
    n, d = g.shape
    if d == 1:
        return np.array(['central'] * n, dtype=object)
    lo, hi = g.min(axis=0), g.max(axis=0)
    plus_in = np.zeros((n, d), dtype=bool)
    minus_in = np.zeros((n, d), dtype=bool)
    for a in range(d):
        p = g.copy()
        p[:, a] = g[:, a] + SHIFT
        plus_in[:, a] = _inside_box(p, lo, hi)
        p[:, a] = g[:, a] - SHIFT
        minus_in[:, a] = _inside_box(p, lo, hi)
    central = np.all(plus_in & minus_in, axis=1)
    out = np.full(n, 'central', dtype=object)
    chars = np.where(plus_in, 'f', 'b')
    for i in np.nonzero(~central)[0]:
        out[i] = ''.join(chars[i])
    return out


def types_by_hull(grid: np.ndarray) -> np.ndarray:
    """Same classification through a qhull Delaunay hull (any point cloud)."""
    from scipy.spatial import Delaunay
    g = np.asarray(grid, dtype=np.float32)
    n, d = g.shape
    if d == 1:
        return np.array(['central'] * n, dtype=object)
    tri = Delaunay(g.astype(np.float64))
    plus_in = np.zeros((n, d), dtype=bool)
    minus_in = np.zeros((n, d), dtype=bool)
    for a in range(d):
        p = g.copy()
        p[:, a] = g[:, a] + SHIFT
        plus_in[:, a] = tri.find_simplex(p.astype(np.float64)) >= 0
        p[:, a] = g[:, a] - SHIFT
        minus_in[:, a] = tri.find_simplex(p.astype(np.float64)) >= 0
    central = np.all(plus_in & minus_in, axis=1)
    chars = np.where(plus_in, 'f', 'b')
    out = np.empty(n, dtype=object)
    for i in range(n):
        out[i] = 'central' if central[i] else ''.join(chars[i])
    return out


def grid_sort(grid: np.ndarray) -> Dict[str, np.ndarray]:
    """Indices of the points of each type, in grid order
    (tedeous/points_type.py:108-125)."""
    types = point_types(grid)
    out: Dict[str, List[int]] = {}
    for i, t in enumerate(types):
        out.setdefault(t, []).append(i)
    return {k: np.asarray(v, dtype=np.int64) for k, v in out.items()}
