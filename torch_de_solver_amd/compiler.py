"""Lowering of a problem (grid, operator dicts, conditions) into flat tables.

This is the compile step of the reference (``Model.compile`` ->
``Solution.__init__`` -> ``Equation_NN/Equation_mat.operator_prepare`` and
``bnd_prepare``, tedeous/model.py:96-113, tedeous/solution.py:23-88,
tedeous/input_preprocessing.py:319-408, 534-594), restated as a table
builder.  Nothing here runs per training step.

'NN' mode lowers to *residual groups*.  A group has ``n_entries`` entries
(collocation points, or boundary points) and ``n_slots`` evaluation slots;
slot ``s`` of entry ``e`` is the network input
``base_s[e] + offset_s`` (the reference's shifted grid, computed in the
working precision as ``x + fl(shift * h)`` per axis,
tedeous/points_type.py:35-37, tedeous/input_preprocessing.py:255-264).
Each group carries one or more equations; an equation is the in-order sum
of terms ``coeff * prod_j (sum_k w_jk u_{slot_jk}[var_j]) ** pow_j``
(tedeous/derivative.py:30-58, tedeous/eval.py:158-165).  A periodic
condition ``u(set0) - u(set1) - ...`` (tedeous/eval.py:342-345) is one
factor whose stencil has weights ``+1, -1, ...`` over per-set slots.

The interior group's entries are the ``'central'`` points only
(tedeous/input_preprocessing.py:356).  Identical shift vectors are
evaluated once: the reference evaluates duplicates separately but gets the
same values, so the residual is unchanged.

'mat' mode lowers to a term table over the whole grid tensor plus boundary
entries given by flat grid positions (tedeous/input_preprocessing.py:553-594).
"""

import itertools
from dataclasses import dataclass, field
from typing import Any, Dict, List, Optional, Sequence, Tuple

import numpy as np
import torch

from .equations import Factor, Term, parse_operator
from .finite_diffs import mat_alphas, nn_scheme
from .points_type import point_types

BC_TYPES = ('dirichlet', 'operator', 'periodic', 'robin', 'data')


# ----------------------------------------------------------------------------
# shared helpers
# ----------------------------------------------------------------------------

def _lambda_list(lam, n: int, what: str) -> List[float]:
    if isinstance(lam, torch.Tensor):
        lam = lam.detach().cpu().reshape(-1).tolist()
    if isinstance(lam, (int, float)):
        return [float(lam)] * n
    lam = [float(x) for x in lam]
    if len(lam) != n:
        raise ValueError(f"{what}: expected {n} weights, got {len(lam)}")
    return lam


def _pow_code(p) -> Tuple[int, float]:
    """(kind, value): kind 0 = integer power, 1 = real power."""
    if callable(p):
        raise NotImplementedError(
            "a callable 'pow' is served in 'NN' and 'autograd' modes only")
    if float(p) == int(p):
        return 0, float(int(p))
    return 1, float(p)


# one-argument operations of a callable power (the TDX_AD_OP_* codes)
AD_OP = dict(ADD=1, SUB=2, RSUB=3, MUL=4, DIV=5, RDIV=6, POWI=7, POWR=8, NEG=9, SIN=10,
             COS=11, EXP=12, LOG=13, TANH=14, SQRT=15)
AD_MAX_CHAIN = 16
_UNARY = {'sin': 'SIN', 'cos': 'COS', 'exp': 'EXP', 'log': 'LOG', 'tanh': 'TANH',
          'sqrt': 'SQRT', 'neg': 'NEG'}


def callable_chain(fn) -> List[Tuple[int, float]]:
    """A callable power (applied by the reference to the term's running
    product, tedeous/derivative.py:124-129) as the chain of one-argument
    operations the 'autograd' kernels evaluate, [(op, arg)]: arg the real
    constant (ADD..RDIV, POWR), the integer power (POWI) or 0.  The callable
    is traced with torch.fx; it must apply, in sequence, +, -, *, / or **
    with a number, unary minus, or torch.sin/cos/exp/log/tanh/sqrt to its
    argument (e.g. lambda u: torch.sin(k * u ** 2)); anything else raises
    NotImplementedError."""
    import operator
    import torch.fx as fx
    try:
        graph = fx.symbolic_trace(fn).graph
    except Exception as exc:  # control flow, non-traceable callables
        raise NotImplementedError(f"a callable 'pow' that torch.fx cannot trace: {exc}")
    bad = NotImplementedError("a callable 'pow' must be a chain of one-argument operations "
                              "(+, -, *, / or ** with a number, unary minus, "
                              "torch.sin/cos/exp/log/tanh/sqrt)")
    chain: List[Tuple[int, float]] = []
    cur = None

    def num(v):
        if isinstance(v, bool) or not isinstance(v, (int, float)):
            raise bad
        return float(v)

    for node in graph.nodes:
        if node.op == 'placeholder':
            if cur is not None:
                raise bad
            cur = node
            continue
        if node.op == 'output':
            if node.args[0] is not cur:
                raise bad
            break
        if node.op == 'call_method':
            name = node.target
        elif node.op == 'call_function':
            name = getattr(node.target, '__name__', '')
        else:
            raise bad
        args = list(node.args)
        if node.kwargs:
            raise bad
        nodes = [a for a in args if isinstance(a, fx.Node)]
        if any(a is not cur for a in nodes) or not nodes:
            raise bad
        tgt = node.target
        if name in _UNARY or tgt in (operator.neg,):
            if len(args) != 1:
                raise bad
            chain.append((AD_OP[_UNARY.get(name, 'NEG')], 0.0))
        elif tgt in (operator.mul, torch.mul) or name == 'mul':
            if len(nodes) == 2:
                chain.append((AD_OP['POWI'], 2.0))      # u * u
            else:
                c = num(args[1] if args[0] is cur else args[0])
                chain.append((AD_OP['MUL'], c))
        elif tgt in (operator.add, torch.add) or name == 'add':
            if len(nodes) == 2:
                chain.append((AD_OP['MUL'], 2.0))       # u + u
            else:
                c = num(args[1] if args[0] is cur else args[0])
                chain.append((AD_OP['ADD'], c))
        elif tgt in (operator.sub, torch.sub) or name == 'sub':
            if len(nodes) == 2:
                raise bad
            if args[0] is cur:
                chain.append((AD_OP['SUB'], num(args[1])))
            else:
                chain.append((AD_OP['RSUB'], num(args[0])))
        elif tgt in (operator.truediv, torch.div) or name in ('div', 'truediv'):
            if len(nodes) == 2:
                raise bad
            if args[0] is cur:
                chain.append((AD_OP['DIV'], num(args[1])))
            else:
                chain.append((AD_OP['RDIV'], num(args[0])))
        elif tgt in (operator.pow, torch.pow) or name == 'pow':
            if len(nodes) == 2 or args[0] is not cur:
                raise bad
            e = num(args[1])
            if e == int(e) and abs(e) <= 64:
                chain.append((AD_OP['POWI'], float(int(e))))
            else:
                chain.append((AD_OP['POWR'], e))
        else:
            raise bad
        cur = node
    if len(chain) > AD_MAX_CHAIN:
        raise NotImplementedError(f"a callable 'pow' of more than {AD_MAX_CHAIN} operations")
    return chain


def _ad_pow_code(p):
    """(kind, value) of an 'NN' or 'autograd' factor's power: 0 integer,
    1 real, 2 callable (value: its operation chain)."""
    if callable(p):
        return 2, callable_chain(p)
    return _pow_code(p)


# ----------------------------------------------------------------------------
# 'NN' mode
# ----------------------------------------------------------------------------

@dataclass
class NNFactor:
    var: int
    pow_kind: int
    pow_val: float
    stencil: List[Tuple[int, float]]          # (slot, weight as double)


@dataclass
class NNTerm:
    coeff_scalar: Optional[float]
    coeff_array: Optional[np.ndarray]          # per entry, or None
    factors: List[NNFactor]


@dataclass
class NNEquation:
    terms: List[NNTerm]                        # summed in order
    lam: float                                 # lambda of its loss column
    denom: float                               # N (operator) or max_len (bc)
    target: Optional[np.ndarray]               # per entry, None -> zeros
    column: int = -1                           # bc column (-1: operator eq)
    col_pos: Optional[np.ndarray] = None       # entry positions in column


@dataclass
class NNGroup:
    kind: str                                  # 'operator' or a bc type
    n_entries: int
    slot_base: List[int]                       # index of base array per slot
    slot_offset: List[List[float]]             # per slot, per axis (double)
    bases: List[np.ndarray]                    # base coordinate arrays [n, d]
    equations: List[NNEquation]


@dataclass
class NNProblem:
    d_in: int
    n_out: int
    groups: List[NNGroup]
    n_central: int
    central_index: np.ndarray                  # grid indices of entries
    bc_keys: List[str]                         # column types, first-seen order
    bc_lengths: List[int]
    bc_max_len: int
    dtype: torch.dtype
    h: float
    # causal loss (tedeous/losses.py:137-182): tolerance and the number of
    # slices the operator points are read in (0: default loss)
    causal_tol: float = 0.0
    causal_nt: int = 0
    # weak form (tedeous/eval.py:13-52, 195-221): 0 = none, 1 = one
    # trapezoid level (1-D grids), 2 = two levels (2-D grids); per 'central'
    # point its inner trapezoid weight times the squared basis product, per
    # run of equal first coordinates its outer trapezoid weight, and each
    # point's run (W = sum_g cx_g (sum_{i in g} ctf2_i r_i^2)^2, or
    # sum_i ctf2_i r_i^2 on one level)
    weak_levels: int = 0
    weak_ctf2: Optional[np.ndarray] = None
    weak_cx: Optional[np.ndarray] = None
    weak_gstart: Optional[np.ndarray] = None


def _coeff_for_entries(coeff, grid: torch.Tensor, entry_idx: np.ndarray,
                       entry_pts: torch.Tensor):
    """Coefficient at the entries (tedeous/input_preprocessing.py:266-295).

    Scalars stay scalars.  A tensor is indexed by grid position (the
    reference looks positions up with ``bndpos``).  A callable is evaluated
    at the entry points once, here; the reference re-evaluates it every
    step on the same points, which gives the same values for a function of
    the coordinates.
    """
    if isinstance(coeff, torch.nn.Parameter):
        raise NotImplementedError("trainable coefficients are not supported")
    if isinstance(coeff, (int, float)):
        return float(coeff), None
    if isinstance(coeff, torch.Tensor):
        flat = coeff.detach().cpu().reshape(-1)
        if flat.numel() == 1:
            return float(flat[0]), None
        if flat.numel() != grid.shape[0]:
            raise ValueError("a tensor coefficient must have one value per "
                             "grid point")
        return None, flat.numpy()[entry_idx]
    if callable(coeff):
        vals = coeff(entry_pts)
        if isinstance(vals, torch.Tensor):
            vals = vals.detach().cpu().numpy()
        vals = np.asarray(vals).reshape(-1)
        if vals.size == 1:
            vals = np.full(len(entry_idx), float(vals[0]))
        return None, vals
    raise NameError('"coeff" should be: torch.Tensor or callable or int or float!')


def _nn_terms(terms: List[Term], nvars: int, point_type: str, order: str,
              h: float, grid: torch.Tensor, entry_idx: np.ndarray,
              entry_pts: torch.Tensor, slot_index: Dict, shifts: List):
    """Lower one operator on one point set; extends the shared slot table."""
    out_terms = []
    for t in terms:
        cs, ca = _coeff_for_entries(t.coeff, grid, entry_idx, entry_pts)
        factors = []
        for f in t.factors:
            sch = nn_scheme(f.axes, nvars, point_type, order, h)
            st = []
            for shift, w in sch:
                if shift not in slot_index:
                    slot_index[shift] = len(shifts)
                    shifts.append(shift)
                st.append((slot_index[shift], w))
            kind, pv = _ad_pow_code(f.pow)
            factors.append(NNFactor(var=f.var, pow_kind=kind, pow_val=pv,
                                    stencil=st))
        out_terms.append(NNTerm(coeff_scalar=cs, coeff_array=ca,
                                factors=factors))
    return out_terms


def _match_grid_rows(grid_np: np.ndarray, pts: np.ndarray) -> np.ndarray:
    """Index of the grid row equal to each point, -1 if none."""
    lut = {tuple(r): i for i, r in enumerate(grid_np.tolist())}
    return np.array([lut.get(tuple(p), -1) for p in pts.tolist()], dtype=np.int64)


def _trapezoid_weights(steps: np.ndarray) -> np.ndarray:
    """Weights c_i with sum_i c_i f_i = sum_{i>=1} s_i (f_i + f_{i-1}) / 2,
    s_i = x_i - x_{i-1} (len(steps) + 1 points)."""
    c = np.zeros(len(steps) + 1, dtype=np.float64)
    c[1:] += steps / 2
    c[:-1] += steps / 2
    return c


def weak_tables(pts: torch.Tensor, weak_form) -> Tuple[int, np.ndarray, np.ndarray, np.ndarray]:
    """(levels, ctf2, cx, gstart) of the weak form on the sorted points
    ``pts`` (tedeous/eval.py:13-52, 195-221).  Steps are taken in the grid's
    precision and then as Python floats, as the reference's ``.item()``."""
    n, d = pts.shape
    if d > 2:
        raise NotImplementedError("the weak form is served on 1-D and 2-D grids")
    f = torch.ones(n, dtype=torch.float64)
    for func in weak_form:
        f = f * torch.as_tensor(func(pts)).reshape(-1).to(torch.float64)
    f2 = (f * f).numpy()
    last = pts[:, -1]
    if d == 1:
        steps = (last[1:] - last[:-1]).to(torch.float64).numpy()
        return 1, _trapezoid_weights(steps) * f2, np.ones(1), np.array([0, n], dtype=np.int64)
    first = pts[:, 0]
    starts = [0] + [i for i in range(1, n) if first[i] != first[i - 1]]
    gstart = np.array(starts + [n], dtype=np.int64)
    ct = np.zeros(n, dtype=np.float64)
    for a, b in zip(gstart[:-1], gstart[1:]):
        steps = (last[a + 1:b] - last[a:b - 1]).to(torch.float64).numpy()
        ct[a:b] = _trapezoid_weights(steps)
    gx = first[torch.from_numpy(gstart[:-1])]
    steps = (gx[1:] - gx[:-1]).to(torch.float64).numpy()
    cx = _trapezoid_weights(steps)
    return 2, ct * f2, cx, gstart


def compile_nn(grid: torch.Tensor, operators: List[dict], bconds: Optional[List[dict]],
               n_out: int, h: float = 0.001, inner_order: str = '1',
               boundary_order: str = '2', lambda_operator=1,
               lambda_bound=1, tol: float = 0.0, weak_form=None) -> NNProblem:
    """Lower an 'NN' problem (tedeous/input_preprocessing.py:173-408,
    tedeous/eval.py:90-461, tedeous/losses.py:84-135)."""
    if inner_order == '2':
        nn_scheme([0], 1, 'central', '2', h)  # raises with the reason
    grid_cpu = grid.detach().cpu()
    dtype = grid_cpu.dtype
    grid_np = grid_cpu.numpy()
    n, d = grid_np.shape
    types = point_types(grid_np)
    central_idx = np.nonzero(types == 'central')[0].astype(np.int64)
    central_pts = grid_cpu[torch.from_numpy(central_idx)]

    if isinstance(operators, dict):
        operators = [operators]
    n_eq = len(operators)
    lam_op = _lambda_list(lambda_operator, n_eq, 'lambda_operator')

    slot_index: Dict = {}
    shifts: List = []
    eqs = []
    for i, op in enumerate(operators):
        terms = _nn_terms(parse_operator(op), d, 'central', inner_order, h, grid_cpu,
                          central_idx, central_pts, slot_index, shifts)
        eqs.append(NNEquation(terms=terms, lam=lam_op[i],
                              denom=float(len(central_idx)), target=None))
    interior = NNGroup(kind='operator', n_entries=len(central_idx),
                       slot_base=[0] * len(shifts),
                       slot_offset=[[s * h for s in sh] for sh in shifts],
                       bases=[central_pts.numpy()], equations=eqs)
    groups = [interior]

    # ---- boundary conditions: columns by type, first-seen order
    # (tedeous/eval.py:448-459, 55-87)
    keys: List[str] = []
    col_vals: Dict[str, List[np.ndarray]] = {}
    col_len: Dict[str, int] = {}
    pending = []   # (group, eq, column key, start position, count)
    for bc in (bconds or []):
        typ = bc['type']
        if typ not in keys:
            keys.append(typ)
            col_vals[typ] = []
            col_len[typ] = 0
        true = bc['bval'].detach().cpu().reshape(-1).numpy()
        col_vals[typ].append(true)
        if typ == 'dirichlet' or (typ in ('periodic', 'data') and bc['bop'] is None):
            # 'data' without an operator: the network's values at the data
            # points (tedeous/eval.py:402-403)
            sets = list(bc['bnd']) if typ == 'periodic' else [bc['bnd']]
            pts = [s.detach().cpu().to(dtype).numpy() for s in sets]
            cnt = pts[0].shape[0]
            if any(p.shape[0] != cnt for p in pts):
                raise ValueError("periodic point sets must have equal sizes")
            # value u(set0) - u(set1) - ... as one +-1 stencil: the same
            # float operations as tedeous/eval.py:342-345
            stencil = [(k, 1.0 if k == 0 else -1.0) for k in range(len(pts))]
            fac = NNFactor(var=int(bc['var']), pow_kind=0, pow_val=1.0,
                           stencil=stencil)
            eq = NNEquation(terms=[NNTerm(1.0, None, [fac])], lam=0.0, denom=1.0, target=None)
            g = NNGroup(kind=typ, n_entries=cnt,
                        slot_base=list(range(len(pts))),
                        slot_offset=[[0.0] * d for _ in pts], bases=pts,
                        equations=[eq])
            groups.append(g)
            pending.append((eq, typ, col_len[typ], np.arange(cnt)))
            col_len[typ] += cnt
        elif typ == 'periodic':
            # u-operator at set 0 minus the same operator at sets 1, 2, ...
            # (tedeous/eval.py:346-350); each set's points take the scheme of
            # their own point type, so entries are grouped by the tuple of
            # their sets' types, and a set's entry e pairs with set 0's e
            sets = [s.detach().cpu().to(dtype) for s in bc['bnd']]
            cnt = sets[0].shape[0]
            if any(s.shape[0] != cnt for s in sets):
                raise ValueError("periodic point sets must have equal sizes")
            rows_k = [_match_grid_rows(grid_np, s.numpy()) for s in sets]
            if any(np.any(r < 0) for r in rows_k):
                raise ValueError("periodic-condition points must be grid points in 'NN' mode")
            tuples = list(zip(*[types[r].tolist() for r in rows_k]))
            terms = parse_operator(bc['bop'])
            start = col_len[typ]
            for tt in dict.fromkeys(tuples):
                sel = np.array([e for e in range(cnt) if tuples[e] == tt], dtype=np.int64)
                slot_base: List[int] = []
                slot_offset: List[List[float]] = []
                all_terms: List[NNTerm] = []
                for k, pt in enumerate(tt):
                    order = inner_order if pt == 'central' else boundary_order
                    s_index: Dict = {}
                    s_list: List = []
                    tl = _nn_terms(terms, d, pt, order, h, grid_cpu, rows_k[k][sel],
                                   sets[k][torch.from_numpy(sel)], s_index, s_list)
                    first = len(slot_base)
                    slot_base += [k] * len(s_list)
                    slot_offset += [[s * h for s in sh] for sh in s_list]
                    for t in tl:
                        for f in t.factors:
                            f.stencil = [(slot + first, w) for slot, w in f.stencil]
                        if k > 0:
                            t.coeff_scalar = None if t.coeff_scalar is None else -t.coeff_scalar
                            t.coeff_array = None if t.coeff_array is None else -t.coeff_array
                        all_terms.append(t)
                eq = NNEquation(terms=all_terms, lam=0.0, denom=1.0, target=None)
                g = NNGroup(kind=typ, n_entries=len(sel), slot_base=slot_base,
                            slot_offset=slot_offset,
                            bases=[sk[torch.from_numpy(sel)].numpy() for sk in sets],
                            equations=[eq])
                groups.append(g)
                pending.append((eq, typ, start, sel))
            col_len[typ] += cnt
        elif typ in ('operator', 'data'):
            # 'data' with an operator: the operator's values at the data
            # points in their own column (tedeous/eval.py:390-406)
            bnd = bc['bnd'].detach().cpu().to(dtype)
            rows = _match_grid_rows(grid_np, bnd.numpy())
            if np.any(rows < 0):
                raise ValueError("operator-condition points must be grid points "
                                 "in 'NN' mode")
            btypes = types[rows]
            terms = parse_operator(bc['bop'])
            start = col_len[typ]
            # one group per point type; each entry keeps its own target
            for pt in dict.fromkeys(btypes.tolist()):
                sel = np.nonzero(btypes == pt)[0]
                order = inner_order if pt == 'central' else boundary_order
                s_index: Dict = {}
                s_list: List = []
                pts_t = bnd[torch.from_numpy(sel)]
                tl = _nn_terms(terms, d, pt, order, h, grid_cpu, rows[sel],
                               pts_t, s_index, s_list)
                eq = NNEquation(terms=tl, lam=0.0, denom=1.0, target=None)
                g = NNGroup(kind=typ, n_entries=len(sel),
                            slot_base=[0] * len(s_list),
                            slot_offset=[[s * h for s in sh] for sh in s_list],
                            bases=[pts_t.numpy()], equations=[eq])
                groups.append(g)
                pending.append((eq, typ, start, sel))
            col_len[typ] += bnd.shape[0]
        else:
            raise NotImplementedError(
                f"boundary type {typ!r}{' with an operator' if bc['bop'] is not None else ''}"
                " is not supported by the native 'NN' engine")

    max_len = max(col_len.values()) if keys else 0
    lam_b = _lambda_list(lambda_bound, len(keys), 'lambda_bound') if keys else []
    # padded target columns (PadTransform, tedeous/utils.py:257-290)
    padded = {}
    for k in keys:
        t = np.concatenate(col_vals[k]) if col_vals[k] else np.zeros(0)
        col = np.zeros(max_len, dtype=np.float64)
        m = min(len(t), max_len)
        col[:m] = t[:m]
        padded[k] = col
    for eq, typ, start, sel in pending:
        ci = keys.index(typ)
        eq.column = ci
        eq.lam = lam_b[ci]
        eq.denom = float(max_len)
        pos = start + np.asarray(sel, dtype=np.int64)
        eq.col_pos = pos
        eq.target = padded[typ][pos]

    # causal slices: the number of distinct first coordinates among the
    # 'central' points (tedeous/solution.py:51-53); when it does not divide
    # the points, the reference's reshape fails and it takes one slice per
    # point (tedeous/losses.py:162-166)
    causal_nt = 0
    if tol:
        n_c = len(central_idx)
        causal_nt = int(torch.unique(central_pts[:, 0]).numel())
        if causal_nt == 0 or n_c % causal_nt != 0:
            causal_nt = n_c
    prob = NNProblem(d_in=d, n_out=n_out, groups=groups,
                     n_central=len(central_idx), central_index=central_idx,
                     bc_keys=keys, bc_lengths=[col_len[k] for k in keys],
                     bc_max_len=max_len, dtype=dtype, h=h,
                     causal_tol=float(tol), causal_nt=causal_nt)
    if weak_form:
        if tol:
            raise NotImplementedError("a weak form with a causal loss is not served")
        lv, ctf2, cx, gstart = weak_tables(central_pts, weak_form)
        prob.weak_levels, prob.weak_ctf2, prob.weak_cx, prob.weak_gstart = lv, ctf2, cx, gstart
    return prob


# ----------------------------------------------------------------------------
# 'mat' mode
# ----------------------------------------------------------------------------

@dataclass
class MatFactor:
    var: int
    axes: List[int]                            # [] for the undifferentiated field
    pow_kind: int
    pow_val: float


@dataclass
class MatTerm:
    coeff_scalar: Optional[float]
    coeff_field: Optional[np.ndarray]          # full-grid field or None
    factors: List[MatFactor]


@dataclass
class MatBC:
    typ: str
    var: int
    positions: List[np.ndarray]                # per set: flat grid indices
    terms: Optional[List[MatTerm]]             # None: plain value u[var]
    column: int = -1
    col_start: int = 0
    # Robin: value alpha * u[var] + sum_beta beta * (operator), with alpha
    # the first term's coefficient and the betas the later terms'
    robin: Optional[Tuple[float, List[float]]] = None


@dataclass
class MatProblem:
    shape: Tuple[int, ...]                     # grid shape (n1, ..., nd)
    n_eq: int
    h: List[float]                             # per-axis step (working precision)
    dp: int
    alpha_back: np.ndarray
    alpha_forw: np.ndarray
    equations: List[List[MatTerm]]
    lam_op: List[float]
    bcs: List[MatBC]
    bc_keys: List[str]
    bc_lengths: List[int]
    bc_max_len: int
    lam_bnd: List[float]
    targets: np.ndarray                        # [max_len, n_cols] padded
    dtype: torch.dtype


def _mat_terms(operator: dict, grid: torch.Tensor, shape, n_eq: int = 1) -> List[MatTerm]:
    out = []
    for t in parse_operator(operator):
        c = t.coeff
        if isinstance(c, torch.nn.Parameter):
            raise NotImplementedError("trainable coefficients are not supported")
        if isinstance(c, (int, float)):
            cs, cf = float(c), None
        else:
            # tedeous/derivative.py:318-321: ``coeff(grid) * der_term`` or
            # ``coeff * der_term`` broadcast against the [n_eq, *grid] field
            val = c(grid) if callable(c) else c
            if not isinstance(val, torch.Tensor):
                val = torch.as_tensor(val)
            val = val.detach().cpu()
            if val.numel() == 1:
                cs, cf = float(val.reshape(-1)[0]), None
            else:
                full = torch.broadcast_to(val, (n_eq,) + tuple(shape))
                if n_eq > 1 and not bool(torch.all(full == full[0:1])):
                    raise NotImplementedError("a coefficient field that differs between "
                                              "equation planes is not supported")
                cs, cf = None, full[0].numpy().copy()
        facs = []
        for f in t.factors:
            kind, pv = _pow_code(f.pow)
            facs.append(MatFactor(var=f.var, axes=list(f.axes or []),
                                  pow_kind=kind, pow_val=pv))
        out.append(MatTerm(coeff_scalar=cs, coeff_field=cf, factors=facs))
    return out


def _mat_positions(grid: torch.Tensor, pts: torch.Tensor) -> np.ndarray:
    """Flat positions of boundary points (tedeous/input_preprocessing.py:553-575).

    In d>1 every position whose coordinates are all ``isclose`` to the point
    is taken (zero or several are possible, as in the reference), in
    row-major order.  The grid is an ``ij`` meshgrid, so the test factors
    into one ``isclose`` per axis on the axis values.  In 1-D a single
    ``isclose`` match is taken, else the nearest point
    (tedeous/input_preprocessing.py:84-150).
    """
    g = grid.detach().cpu()
    d = g.shape[0]
    shape = tuple(g.shape[1:])
    out = []
    if d == 1:
        gd = g.reshape(-1).double()
        for p in pts.detach().cpu():
            v = p.double().reshape(-1)[0]
            m = torch.nonzero(torch.isclose(gd, v)).reshape(-1)
            out.append(int(m[0]) if m.numel() == 1 else int(torch.argmin((gd - v).abs())))
        return np.asarray(out, dtype=np.int64)
    # axis values of the meshgrid: grid[a] varies along axis a only
    axis_vals = []
    for a in range(d):
        idx = [0] * d
        idx[a] = slice(None)
        axis_vals.append(g[a][tuple(idx)].reshape(-1).to(torch.float32))
    strides = np.cumprod((1,) + shape[::-1])[::-1][1:]
    P = pts.detach().cpu().to(torch.float32)
    per_axis = []
    for a in range(d):
        # [n_pts, n_a] closeness of each point's coordinate to the axis values
        per_axis.append(torch.isclose(P[:, a:a + 1], axis_vals[a].reshape(1, -1)).numpy())
    import itertools
    for i in range(P.shape[0]):
        lists = [np.nonzero(per_axis[a][i])[0] for a in range(d)]
        for combo in itertools.product(*lists):
            out.append(int(sum(int(c) * int(s) for c, s in zip(combo, strides))))
    return np.asarray(out, dtype=np.int64)


def compile_mat(grid: torch.Tensor, operators: List[dict], bconds: Optional[List[dict]],
                n_eq: int, derivative_points: int = 2, lambda_operator=1,
                lambda_bound=1) -> MatProblem:
    """Lower a 'mat' problem (tedeous/derivative.py:135-323,
    tedeous/eval.py:283-461, tedeous/losses.py:84-135)."""
    g = grid.detach().cpu()
    d = g.shape[0]
    if d > 3:
        raise NotImplementedError("'mat' mode supports grids of up to 3 axes")
    shape = tuple(g.shape[1:])
    # step per axis: |v1 - v0| of the sorted unique axis values, in the grid
    # dtype (tedeous/derivative.py:229-247)
    hs = []
    for a in range(d):
        # the meshgrid holds each axis value along its own axis only, so the
        # sorted unique values of the field are those of the axis
        idx = [0] * d
        idx[a] = slice(None)
        u = torch.unique(g[a][tuple(idx)].reshape(-1).float())
        hs.append(float(torch.abs(u[1] - u[0])) if u.numel() > 1 else 1.0)
    ab, af, _, _ = mat_alphas(derivative_points)
    if isinstance(operators, dict):
        operators = [operators]
    eqs = [_mat_terms(op, g, shape, n_eq) for op in operators]
    lam_op = _lambda_list(lambda_operator, len(eqs), 'lambda_operator')

    bcs: List[MatBC] = []
    keys: List[str] = []
    col_len: Dict[str, int] = {}
    col_true: Dict[str, List[np.ndarray]] = {}
    for bc in (bconds or []):
        typ = bc['type']
        if typ not in keys:
            keys.append(typ)
            col_len[typ] = 0
            col_true[typ] = []
        col_true[typ].append(bc['bval'].detach().cpu().reshape(-1).numpy())
        robin = None
        if typ == 'dirichlet' or (typ == 'data' and bc['bop'] is None):
            # 'data' without an operator: grid values at the data points
            # (tedeous/eval.py:402-403)
            pos = [_mat_positions(g, bc['bnd'])]
            terms = None
            var = int(bc['var'])
        elif typ == 'robin':
            # alpha * u[var] + sum over the later terms' coefficients beta of
            # beta * (the whole operator at the points, read from its first
            # term's variable) (tedeous/eval.py:357-388)
            pos = [_mat_positions(g, bc['bnd'])]
            parsed = parse_operator(bc['bop'])
            coeffs = []
            for t in parsed:
                c = t.coeff
                if isinstance(c, torch.Tensor) and c.numel() == 1:
                    c = float(c.reshape(-1)[0])
                if not isinstance(c, (int, float)):
                    raise NotImplementedError("Robin conditions take scalar coefficients "
                                              "in the native 'mat' engine")
                coeffs.append(float(c))
            terms = _mat_terms(bc['bop'], g, shape, n_eq)
            robin = (coeffs[0], coeffs[1:])
            var = int(bc['var'])
        elif typ in ('operator', 'data'):
            pos = [_mat_positions(g, bc['bnd'])]
            terms = _mat_terms(bc['bop'], g, shape, n_eq)
            # value read from the bop's first term variable
            # (tedeous/eval.py:320-326)
            var = terms[0].factors[0].var
        elif typ == 'periodic':
            pos = [_mat_positions(g, b) for b in bc['bnd']]
            terms = _mat_terms(bc['bop'], g, shape, n_eq) if bc['bop'] is not None else None
            var = int(bc['var']) if terms is None else terms[0].factors[0].var
            if any(len(p) != len(pos[0]) for p in pos):
                raise ValueError("periodic point sets must have equal sizes")
        else:
            raise NotImplementedError(
                f"boundary type {typ!r} is not supported by the native 'mat' engine")
        m = MatBC(typ=typ, var=var, positions=pos, terms=terms,
                  column=keys.index(typ), col_start=col_len[typ], robin=robin)
        col_len[typ] += len(pos[0])
        bcs.append(m)
    max_len = max(col_len.values()) if keys else 0
    targets = np.zeros((max_len, len(keys)), dtype=np.float64)
    for ci, k in enumerate(keys):
        t = np.concatenate(col_true[k])
        m = min(len(t), max_len)
        targets[:m, ci] = t[:m]
    lam_b = _lambda_list(lambda_bound, len(keys), 'lambda_bound') if keys else []
    return MatProblem(shape=shape, n_eq=n_eq, h=hs, dp=int(derivative_points),
                      alpha_back=ab, alpha_forw=af, equations=eqs, lam_op=lam_op,
                      bcs=bcs, bc_keys=keys, bc_lengths=[col_len[k] for k in keys],
                      bc_max_len=max_len, lam_bnd=lam_b, targets=targets,
                      dtype=g.dtype)


# ----------------------------------------------------------------------------
# 'autograd' mode
# ----------------------------------------------------------------------------

AD_MAX_ORDER = 4


@dataclass
class ADFactor:
    var: int
    pow_kind: int
    pow_val: Any          # the power, or (kind 2) the operation chain of a callable
    stencil: List[Tuple[int, int, float]]      # (point set, jet, weight)


@dataclass
class ADTerm:
    coeff_scalar: Optional[float]
    coeff_array: Optional[np.ndarray]          # per entry, or None
    factors: List[ADFactor]


@dataclass
class ADGroup:
    kind: str                                  # 'operator' or a bc type
    sets: List[np.ndarray]                     # point coordinates per set [n, d]
    equations: List[NNEquation]                # terms are ADTerm


@dataclass
class ADProblem:
    d_in: int
    n_out: int
    jets: List[Tuple[int, ...]]                # jet 0 = () (the value)
    groups: List[ADGroup]
    n_grid: int
    bc_keys: List[str]
    bc_lengths: List[int]
    bc_max_len: int
    dtype: torch.dtype
    # causal loss (tedeous/losses.py:137-182): tolerance and the number of
    # slices the grid points are read in (0: default loss)
    causal_tol: float = 0.0
    causal_nt: int = 0
    # weak form on the grid points (as NNProblem's, any number of equations)
    weak_levels: int = 0
    weak_ctf2: Optional[np.ndarray] = None
    weak_cx: Optional[np.ndarray] = None
    weak_gstart: Optional[np.ndarray] = None

    def partitions(self) -> List[List[List[int]]]:
        """Per jet, its set partitions as lists of block jets (Faa di Bruno:
        the k-th derivative of tanh(z) along the axes of a multiset is the
        sum over the set partitions of its positions of
        tanh^(#blocks)(z) * prod over blocks of the block's derivative)."""
        index = {j: i for i, j in enumerate(self.jets)}
        out = []
        for jet in self.jets:
            parts = []
            for blocks in _set_partitions(list(range(len(jet)))):
                parts.append([index[tuple(sorted(jet[q] for q in b))] for b in blocks])
            out.append(parts if jet else [])
        return out


def _set_partitions(items: List[int]):
    """All set partitions of ``items`` (each a list of blocks)."""
    if not items:
        yield []
        return
    first, rest = items[0], items[1:]
    for part in _set_partitions(rest):
        # put ``first`` in its own block, or into one of the blocks
        yield [[first]] + part
        for k in range(len(part)):
            yield part[:k] + [[first] + part[k]] + part[k + 1:]


def _jet_of(f: Factor) -> Tuple[int, ...]:
    return () if f.axes is None else tuple(sorted(int(a) for a in f.axes))


def _ad_coeff(coeff, pts: torch.Tensor):
    """Coefficient at the entries of a point set: a scalar, or one value per
    entry (a tensor of that length, or a callable evaluated at the points
    once, here; the reference calls it every step on the same points,
    tedeous/derivative.py:116-118)."""
    if isinstance(coeff, torch.nn.Parameter):
        raise NotImplementedError("trainable coefficients are not supported")
    if isinstance(coeff, (int, float)):
        return float(coeff), None
    n = pts.shape[0]
    if isinstance(coeff, torch.Tensor):
        flat = coeff.detach().cpu().reshape(-1)
        if flat.numel() == 1:
            return float(flat[0]), None
        if flat.numel() != n:
            raise ValueError(f"a tensor coefficient must have one value per point ({n})")
        return None, flat.numpy().astype(np.float64)
    if callable(coeff):
        vals = coeff(pts)
        if isinstance(vals, torch.Tensor):
            vals = vals.detach().cpu().numpy()
        vals = np.asarray(vals, dtype=np.float64).reshape(-1)
        if vals.size == 1:
            vals = np.full(n, float(vals[0]))
        return None, vals
    raise NameError('"coeff" should be: torch.Tensor or callable or int or float!')


def _ad_terms(terms: List[Term], pts: torch.Tensor, set_k: int, sign: float = 1.0,
              scale: float = 1.0) -> List[ADTerm]:
    out = []
    for t in terms:
        cs, ca = _ad_coeff(t.coeff, pts)
        k = sign * scale
        cs = None if cs is None else (cs if k == 1.0 else k * cs)
        ca = None if ca is None else (ca if k == 1.0 else k * ca)
        factors = []
        for f in t.factors:
            if f.axes is not None and len(f.axes) > AD_MAX_ORDER:
                raise NotImplementedError(f"derivative order above {AD_MAX_ORDER} in 'autograd' mode")
            kind, pv = _ad_pow_code(f.pow)
            factors.append(ADFactor(var=f.var, pow_kind=kind, pow_val=pv,
                                    stencil=[(set_k, _jet_of(f), 1.0)]))
        if len(factors) > 4:
            raise NotImplementedError("more than 4 factors in a term")
        out.append(ADTerm(coeff_scalar=cs, coeff_array=ca, factors=factors))
    return out


def compile_ad(grid: torch.Tensor, operators: List[dict], bconds: Optional[List[dict]],
               n_out: int, lambda_operator=1, lambda_bound=1, tol: float = 0.0,
               weak_form=None) -> ADProblem:
    """Lower an 'autograd' problem (tedeous/input_preprocessing.py:411-505,
    tedeous/eval.py:90-461 with mode 'autograd', tedeous/losses.py:84-135).

    The operator is evaluated on every grid point (the reference's sorted
    grid is the grid itself in this mode, tedeous/eval.py:121-122); boundary
    operators on the condition's points (tedeous/eval.py:318-319)."""
    grid_cpu = grid.detach().cpu()
    dtype = grid_cpu.dtype
    n, d = grid_cpu.shape
    if isinstance(operators, dict):
        operators = [operators]
    n_eq = len(operators)
    lam_op = _lambda_list(lambda_operator, n_eq, 'lambda_operator')

    eqs = []
    for i, op in enumerate(operators):
        eqs.append(NNEquation(terms=_ad_terms(parse_operator(op), grid_cpu, 0), lam=lam_op[i],
                              denom=float(n), target=None))
    groups = [ADGroup(kind='operator', sets=[grid_cpu.numpy()], equations=eqs)]

    keys: List[str] = []
    col_vals: Dict[str, List[np.ndarray]] = {}
    col_len: Dict[str, int] = {}
    pending = []
    for bc in (bconds or []):
        typ = bc['type']
        if typ not in keys:
            keys.append(typ)
            col_vals[typ] = []
            col_len[typ] = 0
        col_vals[typ].append(bc['bval'].detach().cpu().reshape(-1).numpy())
        var = int(bc['var']) if bc.get('var') is not None else 0
        bop = bc['bop']
        if typ == 'periodic':
            sets = [s.detach().cpu().to(dtype) for s in bc['bnd']]
        else:
            sets = [bc['bnd'].detach().cpu().to(dtype)]
        cnt = sets[0].shape[0]
        if any(s.shape[0] != cnt for s in sets):
            raise ValueError("periodic point sets must have equal sizes")
        if typ in ('dirichlet', 'periodic', 'data') and bop is None:
            # u(set0) - u(set1) - ... (tedeous/eval.py:283-303, 342-345)
            st = [(k, (), 1.0 if k == 0 else -1.0) for k in range(len(sets))]
            terms = [ADTerm(1.0, None, [ADFactor(var=var, pow_kind=0, pow_val=1.0, stencil=st)])]
        elif typ in ('operator', 'data'):
            terms = _ad_terms(parse_operator(bop), sets[0], 0)
        elif typ == 'periodic':
            # operator at set 0 minus the operator at sets 1, 2, ...
            # (tedeous/eval.py:351-354)
            parsed = parse_operator(bop)
            terms = []
            for k, s in enumerate(sets):
                terms += _ad_terms(parsed, s, k, sign=1.0 if k == 0 else -1.0)
        elif typ == 'robin':
            # alpha * u[var] + sum over beta of beta * (the whole operator)
            # (tedeous/eval.py:357-388)
            parsed = parse_operator(bop)
            coeffs = []
            for t in parsed:
                c = t.coeff
                if isinstance(c, torch.Tensor) and c.numel() == 1:
                    c = float(c.reshape(-1)[0])
                if not isinstance(c, (int, float)):
                    raise NotImplementedError("Robin conditions take scalar coefficients")
                coeffs.append(float(c))
            terms = [ADTerm(coeffs[0], None,
                            [ADFactor(var=var, pow_kind=0, pow_val=1.0, stencil=[(0, (), 1.0)])])]
            for beta in coeffs[1:]:
                terms += _ad_terms(parsed, sets[0], 0, scale=beta)
        else:
            raise NotImplementedError(f"boundary type {typ!r} is not supported")
        eq = NNEquation(terms=terms, lam=0.0, denom=1.0, target=None)
        groups.append(ADGroup(kind=typ, sets=[s.numpy() for s in sets], equations=[eq]))
        pending.append((eq, typ, col_len[typ], np.arange(cnt)))
        col_len[typ] += cnt

    max_len = max(col_len.values()) if keys else 0
    lam_b = _lambda_list(lambda_bound, len(keys), 'lambda_bound') if keys else []
    padded = {}
    for k in keys:
        t = np.concatenate(col_vals[k]) if col_vals[k] else np.zeros(0)
        col = np.zeros(max_len, dtype=np.float64)
        m = min(len(t), max_len)
        col[:m] = t[:m]
        padded[k] = col
    for eq, typ, start, sel in pending:
        ci = keys.index(typ)
        eq.column = ci
        eq.lam = lam_b[ci]
        eq.denom = float(max_len)
        pos = start + np.asarray(sel, dtype=np.int64)
        eq.col_pos = pos
        eq.target = padded[typ][pos]

    # jets: every derivative read, closed under sub-multisets, by order
    jets = {()}
    for g in groups:
        for eq in g.equations:
            for t in eq.terms:
                for f in t.factors:
                    for _, jet, _ in f.stencil:
                        for r in range(len(jet) + 1):
                            for sub in itertools.combinations(jet, r):
                                jets.add(tuple(sorted(sub)))
    for jet in jets:
        if any(a < 0 or a >= d for a in jet):
            raise ValueError(f"derivative axis out of range in {jet}")
    jet_list = sorted(jets, key=lambda j: (len(j), j))
    # causal slices: the distinct first coordinates of the grid
    # (tedeous/solution.py:55-57); one slice per point when they do not
    # divide the points (the reference's reshape fails and it takes one,
    # tedeous/losses.py:162-166)
    causal_nt = 0
    if tol:
        causal_nt = int(torch.unique(grid_cpu[:, 0]).numel())
        if causal_nt == 0 or n % causal_nt != 0:
            causal_nt = n
    prob = ADProblem(d_in=d, n_out=n_out, jets=jet_list, groups=groups, n_grid=n,
                     bc_keys=keys, bc_lengths=[col_len[k] for k in keys],
                     bc_max_len=max_len, dtype=dtype, causal_tol=float(tol),
                     causal_nt=causal_nt)
    if weak_form:
        if tol:
            raise NotImplementedError("a weak form with a causal loss is not served")
        # the operator points of this mode are the grid (tedeous/eval.py:206-207)
        lv, ctf2, cx, gstart = weak_tables(grid_cpu, weak_form)
        prob.weak_levels, prob.weak_ctf2, prob.weak_cx, prob.weak_gstart = lv, ctf2, cx, gstart
    return prob
