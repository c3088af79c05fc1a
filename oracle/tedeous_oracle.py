"""CPU oracle: an op-for-op restatement of the reference's residual and loss.

TEST INFRASTRUCTURE ONLY.  Nothing in the product imports this module; it is
used by tests/, by ``__graft_entry__.smoke()`` and by the CPU-baseline leg of
bench.py, always as the checker (never as the thing measured or shipped).

It restates, in torch CPU ops (the arithmetic library the reference itself
computes with), the reference's 'NN', 'mat' and 'autograd' evaluation path:

* grids and conditions        tedeous/data.py:41-107, 241-323
* point typing (qhull)        tedeous/points_type.py:21-125
* stencil schemes             tedeous/finite_diffs.py:9-268
* operator/bc preparation     tedeous/input_preprocessing.py:51-81, 173-408,
                              511-594
* derivatives                 tedeous/derivative.py:18-58, 61-132, 135-323
* residual / bcs              tedeous/eval.py:55-87, 143-193, 283-461
* loss                        tedeous/losses.py:84-135
* training loop               tedeous/model.py:154-188,
                              tedeous/optimizers/closure.py:49-64

The structure follows the reference's data flow (materialised shifted grids,
flattened ``torch.roll`` stencils, padded boundary matrices) so that, in
float32, the residual fields come out with the same rounding as the
reference's.  Parity checks of this module against the reference itself
live in tests/test_oracle.py (golden vectors written by
tests/golden/make_golden.py, closed forms, brute force).
"""

from copy import copy, deepcopy
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch


# ----------------------------------------------------------------------------
# problem definition
# ----------------------------------------------------------------------------

def axis_linspace(a: float, b: float, n_points: int, dtype=torch.float32) -> torch.Tensor:
    """``Domain.variable(name, [a, b], n)``: n + 1 points (tedeous/data.py:62-67)."""
    return torch.linspace(a, b, n_points + 1, dtype=dtype)


def grid_points(axes: List[torch.Tensor]) -> torch.Tensor:
    """'NN' grid ``[N, d]`` (tedeous/data.py:83-88)."""
    if len(axes) == 1:
        return axes[0].reshape(-1, 1)
    return torch.cartesian_prod(*axes)


def grid_mesh(axes: List[torch.Tensor]) -> torch.Tensor:
    """'mat' grid ``[d, n1, ...]`` (tedeous/data.py:102-104)."""
    m = np.meshgrid(*[a.numpy() for a in axes], indexing='ij')
    return torch.tensor(np.array(m))


def condition_points(bnd, axes: Dict[str, torch.Tensor]) -> torch.Tensor:
    """Point set of one condition (tedeous/data.py:241-286)."""
    dtype = next(iter(axes.values())).dtype
    if isinstance(bnd, torch.Tensor):
        out = bnd.to(dtype)
    else:
        parts = []
        for name, ax in axes.items():
            spec = bnd[name]
            if isinstance(spec, torch.Tensor):
                parts.append(spec.to(dtype))
            elif isinstance(spec, (int, float)):
                parts.append(torch.tensor([spec]).to(dtype))
            else:
                lo, hi = spec
                parts.append(ax[(ax >= lo) & (ax <= hi)].to(dtype))
        out = torch.cartesian_prod(*parts).to(dtype)
    return out.reshape(-1, 1) if out.dim() == 1 else out


def unify(op: Optional[dict]) -> Optional[dict]:
    """Per-factor list form of every term (tedeous/input_preprocessing.py:51-81)."""
    if op is None:
        return None
    op = deepcopy(op)
    for term in op.values():
        key = list(term.keys())[1]
        scalar = isinstance(term['pow'], (int, float)) or callable(term['pow'])
        if 'var' not in term:
            if scalar:
                term[key], term['pow'], term['var'] = [term[key]], [term['pow']], [0]
            else:
                term['var'] = [0] * len(term['pow'])
        elif scalar:
            term[key], term['pow'], term['var'] = [term[key]], [term['pow']], [term['var']]
    return op


def build_conditions(conds: List[dict], axes: Dict[str, torch.Tensor]) -> List[dict]:
    """Resolve point sets and values (tedeous/data.py:288-323)."""
    dtype = next(iter(axes.values())).dtype
    out = []
    for c in conds:
        c = dict(c)
        c['bop'] = unify(c.get('bop'))
        if c['type'] == 'periodic':
            c['bnd'] = [condition_points(b, axes) for b in c['bnd']]
        else:
            c['bnd'] = condition_points(c['bnd'], axes)
        v = c['bval']
        if isinstance(v, torch.Tensor):
            c['bval'] = v.to(dtype)
        elif isinstance(v, (int, float)):
            c['bval'] = (torch.ones_like(c['bnd'][:, 0]) * v).to(dtype)
        else:
            c['bval'] = v(c['bnd']).to(dtype)
        out.append(c)
    return out


# ----------------------------------------------------------------------------
# 'NN' mode
# ----------------------------------------------------------------------------

def point_types(grid: torch.Tensor) -> List[str]:
    """Per-point type via the qhull hull test (tedeous/points_type.py:39-106)."""
    from scipy.spatial import Delaunay
    n, d = grid.shape
    tests = []
    tri = Delaunay(grid.numpy()) if d > 1 else None
    for ax in range(d):
        for sign in (1, -1):
            p = grid.clone()
            p[:, ax] = grid[:, ax] + sign * 0.0001
            if d > 1:
                tests.append(tri.find_simplex(p.numpy()) >= 0)
            else:
                tests.append(((p <= grid.max()) & (p >= grid.min())).numpy().reshape(-1))
    tests = np.array(tests).T
    out = []
    for i in range(n):
        if d == 1 or tests[i].all():
            out.append('central')
        else:
            out.append(''.join('f' if tests[i, 2 * a] else 'b' for a in range(d)))
    return out


def scheme(term_axes, nvars: int, ptype: str, label: str, h: float):
    """(shift vectors, weights) of one factor (tedeous/finite_diffs.py:9-268)."""
    if term_axes == [None]:
        return [None], [1]
    if label == '1':
        dirs = ['central'] * len(term_axes) if ptype == 'central' else [ptype[a] for a in term_axes]
        shifts = [[0] * nvars]
        for a, dr in zip(term_axes, dirs):
            nxt = []
            for s in shifts:
                sp, sm = copy(s), copy(s)
                if dr == 'central':
                    sp[a] += 1
                    sm[a] -= 1
                elif dr == 'f':
                    sp[a] += 1
                else:
                    sm[a] -= 1
                nxt += [sp, sm]
            shifts = nxt
        signs = [1]
        central = np.unique(dirs)[0] == 'central'
        for _ in term_axes:
            nxt = []
            for sg in signs:
                nxt += ([sg * (1 / (2 * h)), -sg * (1 / (2 * h))] if central
                        else [sg / h, -sg / h])
            signs = nxt
        return shifts, signs
    dirs = [ptype[a] for a in term_axes]
    shifts = [[0] * nvars]
    for a, dr in zip(term_axes, dirs):
        nxt = []
        for s in shifts:
            s3, s2, s1 = copy(s), copy(s), copy(s)
            step = 1 if dr == 'f' else -1
            s3[a] += 2 * step
            s2[a] += step
            nxt += [s3, s2, s1]
        shifts = nxt
    signs = [1]
    for dr in dirs:
        nxt = []
        for sg in signs:
            if dr == 'f':
                nxt += [3 * (1 / (2 * h)) * sg, -4 * (1 / (2 * h)) * sg, (1 / (2 * h)) * sg]
            else:
                nxt += [-3 * (1 / (2 * h)) * sg, 4 * (1 / (2 * h)) * sg, -(1 / (2 * h)) * sg]
        signs = nxt
    return shifts, signs


def _shift(grid: torch.Tensor, axis: int, amount: float) -> torch.Tensor:
    out = grid.clone()
    out[:, axis] = grid[:, axis] + amount
    return out


def _positions(grid: torch.Tensor, pts: torch.Tensor) -> List[int]:
    """Grid positions of points (tedeous/input_preprocessing.py:84-170)."""
    g = grid.double()
    out = []
    for p in pts.double():
        m = torch.where(torch.all(torch.isclose(g, p), dim=1))[0]
        if m.numel() == 1:
            out.append(int(m))
        else:
            out.append(int(torch.argmin(torch.linalg.norm(g - p, dim=1))))
    return out


def nn_prepare_operator(op: dict, grid: torch.Tensor, pts: torch.Tensor, ptype: str,
                        h: float, inner: str, boundary: str) -> dict:
    """Materialise the shifted grids of every factor
    (tedeous/input_preprocessing.py:208-345)."""
    op = unify(op)
    label = inner if ptype == 'central' else boundary
    nvars = grid.shape[-1]
    for term in op.values():
        key = list(term.keys())[1]
        c = term['coeff']
        if isinstance(c, (int, float)):
            pass
        elif callable(c):
            term['coeff'] = (c, pts)
        elif isinstance(c, torch.Tensor):
            term['coeff'] = c[_positions(grid, pts)].reshape(-1, 1)
        schemes, signs = [], []
        for axes in term[key]:
            sh, sg = scheme(axes, nvars, ptype, label, h)
            grids = []
            for s in sh:
                if s is None:
                    grids.append(pts)
                    continue
                x = pts
                for ax, k in enumerate(s):
                    x = _shift(x, ax, k * h)
                grids.append(x)
            schemes.append(grids)
            signs.append(sg)
        term[key] = [schemes, signs]
    return op


def nn_term_value(net, term: dict) -> torch.Tensor:
    """One term (tedeous/derivative.py:30-58)."""
    key = list(term.keys())[1]
    c = term['coeff']
    if isinstance(c, tuple):
        c = c[0](c[1]).reshape(-1, 1)
    der = 1.
    for j, grids in enumerate(term[key][0]):
        s = 0.
        for k, x in enumerate(grids):
            s += net(x)[:, term['var'][j]].reshape(-1, 1) * term[key][1][j][k]
        p = term['pow'][j]
        der = der * s ** p if isinstance(p, (int, float)) else p(der * s)
    return c * der


def apply_operator(value_fn, op: dict) -> torch.Tensor:
    """In-order sum of term values (tedeous/eval.py:143-165)."""
    total = None
    for term in op.values():
        v = value_fn(term)
        total = v if total is None else total + v
    return total


def pad_matrix(bval: Dict[str, torch.Tensor], true: Dict[str, torch.Tensor]):
    """Columns per type, zero-padded to the longest (tedeous/eval.py:55-87,
    tedeous/utils.py:276-290)."""
    keys = list(bval.keys())
    max_len = max(len(v) for v in bval.values())

    def pad(x):
        return torch.nn.functional.pad(x, (0, max_len - x.size(-1))) if x.size(-1) < max_len else x

    b = torch.hstack([pad(bval[k]).reshape(-1, 1) for k in keys])
    t = torch.hstack([pad(true[k]).reshape(-1, 1) for k in keys])
    return b, t, keys, [len(bval[k]) for k in keys]


def lam_row(val: torch.Tensor, lam) -> torch.Tensor:
    """(tedeous/input_preprocessing.py:14-43)"""
    if isinstance(lam, torch.Tensor):
        return lam
    if isinstance(lam, (int, float)):
        return (torch.ones(val.shape[-1]) * lam).reshape(1, -1)
    return torch.tensor(lam).reshape(1, -1)


def default_loss(op, bval, true, lam_op, lam_bnd):
    """(tedeous/losses.py:84-135)"""
    mse_op = torch.mean((op - torch.zeros(op.shape, dtype=op.dtype)) ** 2, 0)
    mse_b = torch.mean((bval - true) ** 2, 0)
    loss = mse_op @ lam_op.T + mse_b @ lam_bnd.T
    with torch.no_grad():
        norm = mse_op @ torch.ones_like(lam_op).T + mse_b @ torch.ones_like(lam_bnd).T
    return loss, norm


class NNOracle:
    """Reference-semantics evaluation of an 'NN' problem on the CPU."""

    def __init__(self, axes: Dict[str, torch.Tensor], operators: List[dict], conds: List[dict],
                 h: float = 0.001, inner_order: str = '1', boundary_order: str = '2'):
        self.grid = grid_points(list(axes.values()))
        self.types = point_types(self.grid)
        cidx = [i for i, t in enumerate(self.types) if t == 'central']
        self.central = self.grid[cidx]
        self.ops = [nn_prepare_operator(op, self.grid, self.central, 'central', h,
                                        inner_order, boundary_order) for op in operators]
        self.conds = build_conditions(conds, axes)
        lut = {tuple(r): i for i, r in enumerate(self.grid.tolist())}

        def by_type(bop, pts):
            # the operator prepared per point type of the points
            # (tedeous/input_preprocessing.py:371-408)
            typ = [self.types[lut[tuple(p)]] for p in pts.tolist()]
            prepared = []
            for t in dict.fromkeys(typ):
                sel = [i for i, tt in enumerate(typ) if tt == t]
                prepared.append((sel, nn_prepare_operator(bop, self.grid, pts[sel], t,
                                                          h, inner_order, boundary_order)))
            return prepared

        for c in self.conds:
            if c['bop'] is None:
                continue
            if c['type'] in ('operator', 'data'):
                c['prepared'] = by_type(c['bop'], c['bnd'])
            elif c['type'] == 'periodic':
                # one prepared operator per point set (tedeous/input_preprocessing.py:401-403)
                c['prepared'] = [by_type(c['bop'], b) for b in c['bnd']]

    def residual(self, net) -> torch.Tensor:
        cols = [apply_operator(lambda t: nn_term_value(net, t), op).reshape(-1, 1) for op in self.ops]
        return cols[0] if len(cols) == 1 else torch.cat(cols, 1)

    @staticmethod
    def _operator_values(net, prepared):
        """Operator values at a condition's points, per point type, put back
        in the points' order: each point keeps its own target (the
        reference concatenates the types in the iteration order of a Python
        set of strings, which changes from run to run)."""
        parts = [(sel, apply_operator(lambda t: nn_term_value(net, t), op)) for sel, op in prepared]
        v = torch.cat([p for _, p in parts]).reshape(-1, 1)
        order = [i for sel, _ in parts for i in sel]
        inv = torch.empty(len(order), dtype=torch.long)
        inv[torch.tensor(order)] = torch.arange(len(order))
        return v[inv]

    def boundary(self, net):
        bval, true = {}, {}
        for c in self.conds:
            if c['type'] in ('dirichlet', 'data') and c['bop'] is None:
                v = net(c['bnd'])[:, c['var']].reshape(-1, 1)
            elif c['type'] == 'periodic' and c['bop'] is None:
                v = net(c['bnd'][0])[:, c['var']].reshape(-1, 1)
                for b in c['bnd'][1:]:
                    v = v - net(b)[:, c['var']].reshape(-1, 1)
            elif c['type'] == 'periodic':
                # operator at set 0 minus the operator at the other sets
                # (tedeous/eval.py:346-350)
                v = self._operator_values(net, c['prepared'][0])
                for prep in c['prepared'][1:]:
                    v = v - self._operator_values(net, prep)
            elif c['type'] in ('operator', 'data'):
                v = self._operator_values(net, c['prepared'])
            else:
                raise NotImplementedError(c['type'])
            k = c['type']
            if k in bval:
                bval[k] = torch.cat((bval[k], v.reshape(-1)))
                true[k] = torch.cat((true[k], c['bval'].reshape(-1)))
            else:
                bval[k] = v.reshape(-1)
                true[k] = c['bval'].reshape(-1)
        return pad_matrix(bval, true)

    def evaluate(self, net, lam_op=1, lam_bnd=1, tol=0, weak_form=None):
        op = self.residual(net)
        b, t, keys, lens = self.boundary(net)
        if weak_form:
            # the residual on the 'central' points, weighted and integrated
            # (tedeous/eval.py:195-221, tedeous/losses.py:184-228)
            ow = weak_operator(op, self.central, weak_form)
            loss, norm = weak_loss(ow, b, t, lam_row(op, lam_op).to(op.dtype),
                                   lam_row(b, lam_bnd).to(op.dtype))
            return loss, norm, op, b, t, keys, lens
        if tol:
            # slices: distinct first coordinates of the 'central' points
            # (tedeous/solution.py:51-53)
            n_t = len(self.central[:, 0].unique())
            loss, norm = causal_loss(op, b, t, lam_row(b, lam_bnd).to(op.dtype), tol, n_t)
        else:
            loss, norm = default_loss(op, b, t, lam_row(op, lam_op).to(op.dtype),
                                      lam_row(b, lam_bnd).to(op.dtype))
        return loss, norm, op, b, t, keys, lens


def integration(func: torch.Tensor, grid: torch.Tensor, power: int = 2):
    """Trapezoid over the last column of a sorted grid (tedeous/eval.py:13-52):
    per run of equal second-to-last coordinates on a 2-D grid (pairs across
    runs are skipped), over all points on a 1-D grid; the integrand is
    func ** power and the steps are Python floats."""
    column = -1 if grid.shape[-1] == 1 else -2
    marker = grid[0][column]
    index = [0]
    result = []
    u = 0.
    for i in range(1, len(grid)):
        if grid[i][column] == marker or column == -1:
            u += (grid[i][-1] - grid[i - 1][-1]).item() * (func[i] ** power + func[i - 1] ** power) / 2
        else:
            result.append(u)
            marker = grid[i][column]
            index.append(i)
            u = 0.
    if column == -1:
        return u, 0.
    result.append(u)
    return result, grid[index, :-1]


def weak_operator(op: torch.Tensor, grid: torch.Tensor, weak_form) -> torch.Tensor:
    """Operator._weak_pde_compute (tedeous/eval.py:195-221): each equation's
    residual times every basis function, integrated over every axis."""
    sols = []
    for i in range(op.shape[-1]):
        sol = op[:, i]
        for func in weak_form:
            sol = sol * func(grid).reshape(-1)
        g = torch.clone(grid)
        for _ in range(grid.shape[-1]):
            sol, g = integration(sol, g)
            if isinstance(sol, list):
                sol = torch.stack([v if isinstance(v, torch.Tensor) else torch.tensor(v) for v in sol])
        sols.append(sol.reshape(-1, 1))
    return sols[0] if len(sols) == 1 else torch.cat(sols).reshape(1, -1)


def weak_loss(op_w, bval, true, lam_op, lam_bnd):
    """Losses._weak_loss (tedeous/losses.py:184-228): op @ lambda + padded
    boundary mean @ lambda; the normalized loss with lambdas 1."""
    op_w = op_w.reshape(1, -1)
    mse_b = torch.mean((bval - true) ** 2, 0)
    loss = op_w @ lam_op.T + mse_b @ lam_bnd.T
    with torch.no_grad():
        norm = op_w @ torch.ones_like(lam_op).T + mse_b @ torch.ones_like(lam_bnd).T
    return loss, norm


def causal_loss(op, bval, true, lam_bnd, tol, n_t):
    """(tedeous/losses.py:137-182)"""
    try:
        res = torch.sum(op ** 2, dim=1).reshape(n_t, -1)
    except RuntimeError:
        n_t = op.size()[0]
        res = torch.sum(op ** 2, dim=1).reshape(n_t, -1)
    m = torch.triu(torch.ones((n_t, n_t), dtype=res.dtype), diagonal=1).T
    with torch.no_grad():
        w = torch.exp(- tol * (m @ res))
    loss_oper = torch.mean(w * res)
    bval_diff = torch.mean((bval - true) ** 2, 0)
    loss_bnd = bval_diff @ lam_bnd.T
    loss = loss_oper + loss_bnd
    lam_norm = torch.ones_like(lam_bnd)
    with torch.no_grad():
        norm = loss_oper + lam_norm @ bval_diff
    return loss, norm


# ----------------------------------------------------------------------------
# 'mat' mode
# ----------------------------------------------------------------------------

def mat_alphas(dp: int):
    """Vandermonde solve for the one-sided coefficients (tedeous/derivative.py:157-197)."""
    from scipy import linalg
    back = list(range(-dp + 1, 1))
    forw = list(range(dp))

    def solve(labels):
        lab = np.array(labels)
        A = np.array([lab ** i for i in range(len(labels))])
        b = np.zeros_like(lab)
        b[1] = 1
        return linalg.solve(A, b)

    return back, forw, solve(back), solve(forw)


class MatDerivative:
    """Per-axis stencil through flattened rolls (tedeous/derivative.py:199-323)."""

    def __init__(self, dp: int):
        self.back, self.forw, self.ab, self.af = mat_alphas(dp)
        n = len(self.back) - 1
        self.back_idx = [-i for i in range(1, n + 1)]
        self.forw_idx = list(range(n))

    def _combine(self, u):
        b, f = 0, 0
        for i, (sb, sf) in enumerate(zip(self.back, self.forw)):
            b += torch.roll(u, -sb) * self.ab[i]
            f += torch.roll(u, -sf) * self.af[i]
        return b, f

    @staticmethod
    def step(grid: torch.Tensor, axis: int):
        flat = torch.vstack([grid[i].reshape(-1) for i in range(grid.shape[0])]).T.float()
        v = torch.unique(flat[:, axis])
        return abs(v[1] - v[0])

    def d(self, u: torch.Tensor, h, axis: int) -> torch.Tensor:
        if u.dim() == 1 or u.shape[0] == 1:
            shape = u.shape
            x = u.reshape(-1)
            b, f = self._combine(x)
            du = (b + f) / (2 * h)
            du[self.back_idx] = b[self.back_idx] / h
            du[self.forw_idx] = f[self.forw_idx] / h
            return du.reshape(shape)
        pos = u.dim() - 1
        x = torch.transpose(u, pos, axis)
        b, f = self._combine(x)
        du = (b + f) / (2 * h)
        if pos == 1:
            du[:, self.back_idx] = b[:, self.back_idx] / h
            du[:, self.forw_idx] = f[:, self.forw_idx] / h
        elif pos == 2:
            du[:, :, self.back_idx] = b[:, :, self.back_idx] / h
            du[:, :, self.forw_idx] = f[:, :, self.forw_idx] / h
        return torch.transpose(du, pos, axis)

    def term(self, model: torch.Tensor, term: dict, grid: torch.Tensor) -> torch.Tensor:
        key = list(term.keys())[1]
        der = torch.zeros_like(model) + 1
        for j, axes in enumerate(term[key]):
            prod = model[term['var'][j]]
            if axes != [None]:
                for ax in axes:
                    if ax is None:
                        continue
                    prod = self.d(prod, self.step(grid, ax), ax)
            p = term['pow'][j]
            der = der * prod ** p if isinstance(p, (int, float)) else p(der * prod)
        c = term['coeff']
        return c(grid) * der if callable(c) else c * der


class MatOracle:
    """Reference-semantics evaluation of a 'mat' problem on the CPU."""

    def __init__(self, axes: Dict[str, torch.Tensor], operators: List[dict], conds: List[dict],
                 derivative_points: int = 2):
        self.grid = grid_mesh(list(axes.values()))
        self.D = MatDerivative(derivative_points)
        self.ops = [unify(op) for op in operators]
        self.conds = build_conditions(conds, axes)
        for c in self.conds:
            sets = c['bnd'] if c['type'] == 'periodic' else [c['bnd']]
            c['pos'] = [self._positions(s) for s in sets]

    def _positions(self, pts):
        """(tedeous/input_preprocessing.py:553-575)"""
        out = []
        for p in pts:
            if self.grid.shape[0] == 1:
                out.append((torch.tensor(_positions(self.grid.reshape(-1, 1), p.reshape(1, -1))),))
                continue
            prod = torch.ones_like(self.grid[0], dtype=torch.bool)
            for ax in range(self.grid.shape[0]):
                prod &= torch.isclose(p[ax].float(), self.grid[ax].float())
            out.append(torch.where(prod))
        return out

    def field(self, model, op):
        return apply_operator(lambda t: self.D.term(model, t, self.grid), op)

    def residual(self, model):
        cols = [self.field(model, op).reshape(-1, 1) for op in self.ops]
        return cols[0] if len(cols) == 1 else torch.cat(cols, 1)

    def boundary(self, model):
        bval, true = {}, {}
        for c in self.conds:
            if c['type'] == 'robin':
                # alpha u[var] + sum_beta beta * (operator at the points), alpha
                # and the betas being the terms' coefficients in order
                # (tedeous/eval.py:357-388)
                keys = list(c['bop'].keys())
                alpha, *betas = [c['bop'][k]['coeff'] for k in keys]
                pos = c['pos'][0]
                value_term = alpha * torch.cat([model[c['var']][p] for p in pos]).reshape(-1, 1)
                var = c['bop'][keys[0]]['var'][0]
                f = self.field(model, c['bop'])
                op_val = torch.cat([f[var][p] for p in pos]).reshape(-1, 1)
                derivative_term = 0
                for beta in betas:
                    derivative_term += beta * op_val
                vals = [value_term + derivative_term]
            elif c['bop'] is None:
                vals = [torch.cat([model[c['var']][p] for p in pos]).reshape(-1, 1) for pos in c['pos']]
            else:
                var = c['bop'][list(c['bop'].keys())[0]]['var'][0]
                f = self.field(model, c['bop'])
                vals = [torch.cat([f[var][p] for p in pos]).reshape(-1, 1) for pos in c['pos']]
            v = vals[0]
            for w in vals[1:]:
                v = v - w
            k = c['type']
            if k in bval:
                bval[k] = torch.cat((bval[k], v.reshape(-1)))
                true[k] = torch.cat((true[k], c['bval'].reshape(-1)))
            else:
                bval[k] = v.reshape(-1)
                true[k] = c['bval'].reshape(-1)
        return pad_matrix(bval, true)

    def evaluate(self, model, lam_op=1, lam_bnd=1, tol=0):
        op = self.residual(model)
        b, t, keys, lens = self.boundary(model)
        if tol:
            # the operator matrix stacks each equation's field over n_eq
            # broadcast copies ([n_eq N, n_eq]) and the slices are its
            # columns: n_t = op.shape[1] (tedeous/solution.py:62-64,
            # tedeous/losses.py:160-161)
            loss, norm = causal_loss(op, b, t, lam_row(b, lam_bnd).to(op.dtype), tol, op.shape[1])
        else:
            loss, norm = default_loss(op, b, t, lam_row(op, lam_op).to(op.dtype),
                                      lam_row(b, lam_bnd).to(op.dtype))
        return loss, norm, op, b, t, keys, lens


# ----------------------------------------------------------------------------
# 'autograd' mode
# ----------------------------------------------------------------------------

def ad_derivative(net, points: torch.Tensor, var: int, axis: List[int]) -> torch.Tensor:
    """Derivative_autograd._nn_autograd (tedeous/derivative.py:73-98): the
    network's output ``var`` summed over the points, differentiated along
    each axis in turn with ``create_graph=True``; the last gradient's
    column is the mixed partial at every point."""
    points = points.detach().clone().requires_grad_(True)
    fi = net(points)[:, var].sum(0)
    grads = None
    for ax in axis:
        grads, = torch.autograd.grad(fi, points, create_graph=True)
        fi = grads[:, ax].sum()
    return grads[:, axis[-1]].reshape(-1, 1)


def ad_term_value(net, term: dict, pts: torch.Tensor) -> torch.Tensor:
    """Derivative_autograd.take_derivative (tedeous/derivative.py:100-132)
    with the coefficient checks of tedeous/input_preprocessing.py:433-458."""
    key = list(term.keys())[1]
    c = term['coeff']
    if callable(c) and not isinstance(c, torch.Tensor):
        c = c(pts).reshape(-1, 1)
    elif isinstance(c, torch.Tensor):
        c = c.reshape(-1, 1)
    der = 1.
    for j, derivative in enumerate(term[key]):
        if derivative == [None]:
            d = net(pts)[:, term['var'][j]].reshape(-1, 1)
        else:
            d = ad_derivative(net, pts, term['var'][j], derivative)
        p = term['pow'][j]
        der = der * d ** p if isinstance(p, (int, float)) else p(der * d)
    return c * der


class ADOracle:
    """Reference-semantics evaluation of an 'autograd' problem on the CPU:
    the operator on every grid point (tedeous/eval.py:117-123), conditions
    at their points (tedeous/eval.py:283-461, mode 'autograd')."""

    def __init__(self, axes: Dict[str, torch.Tensor], operators: List[dict], conds: List[dict]):
        self.grid = grid_points(list(axes.values()))
        self.ops = [unify(op) for op in operators]
        self.conds = build_conditions(conds, axes)

    def _op(self, net, op, pts):
        return apply_operator(lambda t: ad_term_value(net, t, pts), op).reshape(-1, 1)

    def residual(self, net) -> torch.Tensor:
        cols = [self._op(net, op, self.grid) for op in self.ops]
        return cols[0] if len(cols) == 1 else torch.cat(cols, 1)

    def boundary(self, net):
        bval, true = {}, {}
        for c in self.conds:
            typ, bop = c['type'], c['bop']
            if typ in ('dirichlet', 'data') and bop is None:
                v = net(c['bnd'])[:, c['var']].reshape(-1, 1)
            elif typ == 'periodic' and bop is None:
                v = net(c['bnd'][0])[:, c['var']].reshape(-1, 1)
                for b in c['bnd'][1:]:
                    v = v - net(b)[:, c['var']].reshape(-1, 1)
            elif typ == 'periodic':
                # tedeous/eval.py:351-354
                v = self._op(net, bop, c['bnd'][0])
                for b in c['bnd'][1:]:
                    v = v - self._op(net, bop, b)
            elif typ in ('operator', 'data'):
                v = self._op(net, bop, c['bnd'])
            elif typ == 'robin':
                # tedeous/eval.py:357-388: alpha * u + beta * (whole operator)
                alpha, *betas = [bop[k]['coeff'] for k in bop.keys()]
                v = alpha * net(c['bnd'])[:, c['var']].reshape(-1, 1)
                dt = 0
                for beta in betas:
                    dt += beta * self._op(net, bop, c['bnd'])
                v = v + dt
            else:
                raise NotImplementedError(typ)
            k = typ
            if k in bval:
                bval[k] = torch.cat((bval[k], v.reshape(-1)))
                true[k] = torch.cat((true[k], c['bval'].reshape(-1)))
            else:
                bval[k] = v.reshape(-1)
                true[k] = c['bval'].reshape(-1)
        return pad_matrix(bval, true)

    def evaluate(self, net, lam_op=1, lam_bnd=1, tol=0, weak_form=None):
        op = self.residual(net)
        b, t, keys, lens = self.boundary(net)
        if weak_form:
            # the residual on the grid, weighted and integrated
            # (tedeous/eval.py:195-221, tedeous/losses.py:184-228)
            ow = weak_operator(op, self.grid, weak_form)
            loss, norm = weak_loss(ow, b, t, lam_row(op, lam_op).to(op.dtype),
                                   lam_row(b, lam_bnd).to(op.dtype))
            return loss, norm, op, b, t, keys, lens
        if tol:
            # slices: distinct first coordinates of the grid
            # (tedeous/solution.py:55-57)
            n_t = len(self.grid[:, 0].unique())
            loss, norm = causal_loss(op, b, t, lam_row(b, lam_bnd).to(op.dtype), tol, n_t)
            return loss, norm, op, b, t, keys, lens
        loss, norm = default_loss(op, b, t, lam_row(op, lam_op).to(op.dtype),
                                  lam_row(b, lam_bnd).to(op.dtype))
        return loss, norm, op, b, t, keys, lens


# ----------------------------------------------------------------------------
# training loop
# ----------------------------------------------------------------------------

def train_adam(params: List[torch.Tensor], loss_fn: Callable[[], torch.Tensor], lr: float,
               epochs: int, betas=(0.9, 0.999), eps=1e-8) -> List[float]:
    """``Model.train(Optimizer('Adam', {'lr': lr}), epochs)``: one initial
    evaluation, then ``epochs - 1`` closure steps (tedeous/model.py:154-188,
    tedeous/optimizers/closure.py:49-64).  Returns the pre-step losses."""
    opt = torch.optim.Adam(params, lr=lr, betas=betas, eps=eps)
    history = []

    def closure():
        opt.zero_grad()
        loss = loss_fn()
        loss.backward()
        history.append(float(loss))
        return loss

    for _ in range(1, epochs):
        opt.zero_grad()
        opt.step(closure)
    return history


# ----------------------------------------------------------------------------
# natural gradient (NGD)
# ----------------------------------------------------------------------------

def ngd_gram(residuals: torch.Tensor, params: List[torch.Tensor]) -> torch.Tensor:
    """``NGD.gram_factory`` (tedeous/optimizers/ngd.py:57-78): one
    ``torch.autograd.grad`` per residual entry gives a row of J; the Gram
    matrix is ``1/len(residuals) * J.T @ J``."""
    rows = []
    for r in residuals:
        g = torch.autograd.grad(r, params, retain_graph=True, allow_unused=True)
        g = [torch.zeros_like(p) if x is None else x for p, x in zip(params, g)]
        rows.append(torch.cat([x.reshape(-1) for x in g]).reshape(1, -1))
    J = torch.cat(rows)
    return 1.0 / len(residuals) * J.T @ J


def ngd_lstsq(A: torch.Tensor, B: torch.Tensor, tol: Optional[float] = None) -> torch.Tensor:
    """``NGD.torch_cuda_lstsq`` (tedeous/optimizers/ngd.py:104-126): the
    pseudo-inverse solution by SVD, singular values above ``tol`` (machine
    epsilon of the dtype, absolute) inverted."""
    tol = torch.finfo(A.dtype).eps if tol is None else tol
    U, S, Vh = torch.linalg.svd(A, full_matrices=False)
    Spinv = torch.zeros_like(S)
    Spinv[S > tol] = 1 / S[S > tol]
    UhB = U.adjoint() @ B
    if Spinv.ndim != UhB.ndim:
        Spinv = Spinv.unsqueeze(-1)
    return Vh.adjoint() @ (Spinv * UhB)


def ngd_steps(grid_steps_number: int) -> torch.Tensor:
    """Candidate step sizes ``0.5 ** linspace(0, n, n + 1)`` (float32,
    tedeous/optimizers/ngd.py:24-25)."""
    return 0.5 ** torch.linspace(0, grid_steps_number, grid_steps_number + 1)


def ngd_line_search(params: List[torch.Tensor], f_nat_grad: torch.Tensor,
                    loss_fn: Callable[[], torch.Tensor], steps: torch.Tensor):
    """``NGD.grid_line_search_update`` (tedeous/optimizers/ngd.py:29-55): the
    loss at ``params - step * f`` for every candidate step, the first
    minimum taken; returns (losses, chosen index)."""
    from torch.nn.utils import parameters_to_vector, vector_to_parameters
    base = parameters_to_vector(params).detach().clone()
    losses = []
    # autograd stays on: the 'autograd' operator differentiates the network
    for st in steps:
        vector_to_parameters(base - st * f_nat_grad, params)
        losses.append(loss_fn().detach().reshape(1))
    losses = torch.cat(losses)
    k = int(torch.argmin(losses))
    vector_to_parameters(base - steps[k] * f_nat_grad, params)
    return losses, k


def ngd_step(net, orc, lam_op, lam_bnd, grid_steps_number: int = 30):
    """One ``NGD.step`` with the closure of tedeous/optimizers/closure.py:
    98-117 on an oracle problem: loss and gradient, the Gram matrices of the
    operator residuals (flattened [N, n_eq]) and of the padded boundary
    residuals ``bval - true_bval`` (flattened [max_len, n_types]), the
    Marquardt term ``min(loss, 0) * I`` (zero for a non-negative loss), the
    SVD least-squares natural gradient and the grid line search.  Returns a
    dict of the intermediate values."""
    params = list(net.parameters())
    loss, _, op, b, t, _, _ = orc.evaluate(net, lam_op, lam_bnd)
    grads = torch.autograd.grad(loss, params, retain_graph=True, allow_unused=True)
    f_grads = torch.cat([g.reshape(-1) for g in grads])
    G_int = ngd_gram(op.reshape(-1), params)
    G_b = ngd_gram((b - t).reshape(-1), params)
    G = G_int + G_b
    damp = torch.minimum(loss.detach().reshape(()), torch.zeros((), dtype=loss.dtype))
    G = damp * torch.eye(len(G), dtype=G.dtype) + G
    f = ngd_lstsq(G.detach(), f_grads.detach())
    steps = ngd_steps(grid_steps_number)
    losses, k = ngd_line_search(params, f, lambda: orc.evaluate(net, lam_op, lam_bnd)[0], steps)
    return dict(loss=float(loss.detach()), grad=f_grads.detach(), G_int=G_int.detach(), G_b=G_b.detach(),
                G=G.detach(), f=f, losses=losses, k=k, step=float(steps[k]))
