"""'mat' mode engine: table packing, the native plan and its autograd Function."""

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import native
from .compiler import MatProblem, MatTerm


def _chain_functional(g0: int, g1: int, axes: List[int], shape: Tuple[int, int], dp: int,
                      h: List[float], ab, af) -> Dict[int, float]:
    """Coefficients of ``D_{axes[-1]} ... D_{axes[0]} u`` at (g0, g1) over the
    flat grid points, in double precision (the transposed stencil of the
    boundary terms)."""
    n0, n1 = shape
    if not axes:
        return {g0 * n1 + g1: 1.0}
    ax = axes[-1]
    rest = axes[:-1]
    g = g0 if ax == 0 else g1
    n = n0 if ax == 0 else n1

    def at(off):
        return _chain_functional(g0 + off if ax == 0 else g0, g1 if ax == 0 else g1 + off,
                                 rest, shape, dp, h, ab, af)

    out: Dict[int, float] = {}

    def add(fn, w):
        for k, v in fn.items():
            out[k] = out.get(k, 0.0) + w * v

    if g < dp - 1:
        for i in range(dp):
            add(at(i), af[i] / h[ax])
    elif g >= n - dp + 1:
        for i in range(dp):
            add(at(i - (dp - 1)), ab[i] / h[ax])
    else:
        for i in range(dp):
            add(at(i - (dp - 1)), ab[i] / (2 * h[ax]))
        for i in range(dp):
            add(at(i), af[i] / (2 * h[ax]))
    return out


class MatTables:
    def __init__(self, prob: MatProblem, np_dtype):
        ints: List[int] = []
        reals: List[float] = []
        data_parts: List[np.ndarray] = []
        self.n_data = 0
        shape = prob.shape
        self.n2 = 1
        if len(shape) == 1:
            self.n0, self.n1 = 1, shape[0]
            axis_map = {0: 1}
            self.h = [1.0, prob.h[0], 0.0]
        elif len(shape) == 2:
            self.n0, self.n1 = shape
            axis_map = {0: 0, 1: 1}
            self.h = [prob.h[0], prob.h[1], 0.0]
        elif len(shape) == 3:
            self.n0, self.n1, self.n2 = shape
            axis_map = {0: 0, 1: 1, 2: 2}
            self.h = [prob.h[0], prob.h[1], prob.h[2]]
        else:
            raise NotImplementedError("the native 'mat' engine handles grids of 1 to 3 axes")
        N = self.n0 * self.n1 * self.n2
        chains = [len(f.axes) for terms in prob.equations for t in terms for f in t.factors]
        chains += [len(f.axes) for bc in prob.bcs for t in (bc.terms or []) for f in t.factors]
        # plans of 3-D grids or longer chains run full-grid passes, which take
        # the boundary seeds without the transposed-stencil table
        self.full_grid = len(shape) == 3 or max(chains, default=0) > 3
        self.N = N

        def add_data(arr) -> int:
            arr = np.ascontiguousarray(np.asarray(arr, dtype=np_dtype).reshape(-1))
            off = self.n_data
            data_parts.append(arr)
            self.n_data += arr.size
            return off

        def add_real(v) -> int:
            reals.append(float(v))
            return len(reals) - 1

        def add_terms(terms: List[MatTerm]) -> int:
            term_i = len(ints)
            ints.extend([0] * (len(terms) * native.TDX_MTERM_INTS))
            for ti, t in enumerate(terms):
                factor_i = len(ints)
                for f in t.factors:
                    ax = [axis_map[a] for a in f.axes]
                    if len(ax) > native.TDX_MAT_MAX_CHAIN:
                        raise NotImplementedError(
                            f"derivative order above {native.TDX_MAT_MAX_CHAIN} in 'mat' mode")
                    pow_r = add_real(f.pow_val)
                    ints.extend([f.var, f.pow_kind, int(f.pow_val) if f.pow_kind == 0 else 0,
                                 pow_r, len(ax)] + ax + [0] * (native.TDX_MAT_MAX_CHAIN - len(ax)))
                if t.coeff_field is not None:
                    ck, ci = 1, add_data(np.asarray(t.coeff_field).reshape(-1))
                else:
                    ck, ci = 0, add_real(np_dtype(t.coeff_scalar))
                ints[term_i + ti * native.TDX_MTERM_INTS:
                     term_i + (ti + 1) * native.TDX_MTERM_INTS] = [len(t.factors), factor_i, ck, ci]
            return term_i

        # equations
        self.eq_i = len(ints)
        ints.extend([0] * (len(prob.equations) * native.TDX_MEQ_INTS))
        for q, terms in enumerate(prob.equations):
            term_i = add_terms(terms)
            lam_r = add_real(prob.lam_op[q] / N)
            ints[self.eq_i + q * native.TDX_MEQ_INTS:
                 self.eq_i + (q + 1) * native.TDX_MEQ_INTS] = [len(terms), term_i, lam_r]

        # boundary entries
        entries = []  # (record fields), built after term tables
        seed = 0
        csr: Dict[int, List[Tuple[int, float]]] = {}
        ab = [float(np_dtype(x)) for x in prob.alpha_back]
        af = [float(np_dtype(x)) for x in prob.alpha_forw]
        self.bc_entries: List[Tuple[int, int, int]] = []   # (bc index, column, column pos)
        max_len = prob.bc_max_len
        for bi, bc in enumerate(prob.bcs):
            term_i = add_terms(bc.terms) if bc.terms is not None else 0
            n_terms = len(bc.terms) if bc.terms is not None else 0
            lam_r = add_real(prob.lam_bnd[bc.column] / max_len)
            inv_r = add_real(1.0 / max_len)
            n_ent = len(bc.positions[0])
            for e in range(n_ent):
                pos_i = len(ints)
                sets = [int(p[e]) for p in bc.positions]
                ints.extend(sets)
                tgt = prob.targets[bc.col_start + e, bc.column] if bc.col_start + e < max_len else 0.0
                target_r = add_data([tgt])
                seed_i = seed
                robin_rec = 0
                if bc.robin is not None:
                    # ints {number of betas, reals index of alpha}, then the
                    # betas follow alpha in the reals; first seed: alpha u[var]
                    alpha, betas = bc.robin
                    robin_rec = len(ints) + 1
                    ints.extend([len(betas), add_real(np_dtype(alpha))])
                    for b in betas:
                        add_real(np_dtype(b))
                    if not self.full_grid:
                        csr.setdefault(bc.var * N + sets[0], []).append((seed, 1.0))
                    seed += 1
                for s, pos in enumerate(sets):
                    g0, g1 = divmod(pos, self.n1)
                    if n_terms == 0:
                        if not self.full_grid:
                            csr.setdefault(bc.var * N + pos, []).append((seed, 1.0))
                        seed += 1
                        continue
                    for t in bc.terms:
                        for f in t.factors:
                            if not self.full_grid:
                                ax = [axis_map[a] for a in f.axes]
                                fn = _chain_functional(g0, g1, ax, (self.n0, self.n1), prob.dp,
                                                       self.h, ab, af)
                                for j, w in fn.items():
                                    if w != 0.0:
                                        csr.setdefault(f.var * N + j, []).append((seed, w))
                            seed += 1
                entries.append([len(sets), pos_i, n_terms, term_i, bc.var, target_r, lam_r,
                                inv_r, seed_i, robin_rec])
                self.bc_entries.append((bi, bc.column, bc.col_start + e))
        self.bc_i = len(ints)
        for rec in entries:
            ints.extend(rec)
        self.n_bc = len(entries)
        self.n_bc_seeds = seed
        keys = sorted(csr.keys())
        self.n_touch = len(keys)
        ptr, pos_l, seed_l, w_l = [0], [], [], []
        for k in keys:
            for s, w in csr[k]:
                seed_l.append(s)
                w_l.append(add_real(np_dtype(w)))
            ptr.append(len(seed_l))
            pos_l.append(k)
        self.bct_ptr_i = len(ints)
        ints.extend(ptr)
        self.bct_pos_i = len(ints)
        ints.extend(pos_l)
        self.bct_seed_i = len(ints)
        ints.extend(seed_l)
        self.bct_w_i = len(ints)
        ints.extend(w_l)
        self.ints = np.asarray(ints if ints else [0], dtype=np.int32)
        self.reals = np.asarray(reals if reals else [0.0], dtype=np_dtype)
        self.data = (np.concatenate(data_parts) if data_parts
                     else np.zeros(1, dtype=np_dtype)).astype(np_dtype)
        self.ab = np.asarray(prob.alpha_back, dtype=np.float64)
        self.af = np.asarray(prob.alpha_forw, dtype=np.float64)
        self.dtype_code = 0 if np_dtype == np.float32 else 1
        self.n_eq = prob.n_eq
        self.dp = prob.dp

    def desc(self, exact_div=1, tile_rows: int = 0, specialize: int = 0,
             band_rows: int = 0) -> 'native.MatDesc':
        """Plan descriptor over these tables (which must outlive it).
        exact_div: 0 fast product, 1 correctly rounded, 2 IEEE division."""
        d = native.MatDesc()
        d.dtype = self.dtype_code
        d.n_eq = self.n_eq
        d.n0, d.n1, d.n2 = self.n0, self.n1, self.n2
        d.dp = self.dp
        d.h[0], d.h[1], d.h[2] = self.h[0], self.h[1], self.h[2]
        d.alpha_back = self.ab.ctypes.data_as(native.c_dp)
        d.alpha_forw = self.af.ctypes.data_as(native.c_dp)
        d.eq_i = self.eq_i
        d.n_bc, d.bc_i, d.n_bc_seeds = self.n_bc, self.bc_i, self.n_bc_seeds
        d.n_touch = self.n_touch
        d.bct_ptr_i, d.bct_pos_i, d.bct_seed_i, d.bct_w_i = (self.bct_ptr_i, self.bct_pos_i,
                                                              self.bct_seed_i, self.bct_w_i)
        d.n_ints, d.ints = self.ints.size, self.ints.ctypes.data_as(native.c_i32p)
        d.n_reals, d.reals = self.reals.size, self.reals.ctypes.data
        d.n_data, d.data = self.data.size, self.data.ctypes.data
        d.exact_div = int(exact_div)
        d.tile_rows = int(tile_rows)
        d.specialize = int(specialize)
        d.band_rows = int(band_rows)
        return d

    def specialized_source(self, exact_div=1, compile: bool = False) -> str:
        """Source of the interior kernel a plan over these tables generates
        ('' when it has none); with compile, also compiled for gfx950."""
        d = self.desc(exact_div)
        lib = native.lib()
        n = int(lib.tdx_mat_specialized_source(ctypes.byref(d), None, 0, 0))
        if n < 0:
            native.check(int(n))
        if n == 0:
            return ''
        buf = ctypes.create_string_buffer(n + 1)
        rc = int(lib.tdx_mat_specialized_source(ctypes.byref(d), buf, n + 1, 1 if compile else 0))
        if rc < 0:
            native.check(rc)
        return buf.value.decode()


class MatEngine:
    """Native 'mat' residual/loss engine for one grid tensor."""

    def __init__(self, prob: MatProblem, device: torch.device, exact_div: bool = True,
                 tile_rows: int = 0, specialize: int = 0, band_rows: int = 0):
        if prob.dtype not in (torch.float32, torch.float64):
            raise NotImplementedError(f"dtype {prob.dtype} is not supported")
        self.prob = prob
        self.dtype = prob.dtype
        np_dtype = np.float32 if prob.dtype == torch.float32 else np.float64
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise native.NativeUnavailable("the native engine runs on a HIP device only")
        t = MatTables(prob, np_dtype)
        self.tables = t
        d = t.desc(exact_div, tile_rows, specialize, band_rows)
        handle = ctypes.c_void_p()
        lib = native.lib()
        with torch.cuda.device(self.device):
            native.check(lib.tdx_mat_plan_create(ctypes.byref(d), ctypes.byref(handle)))
        self._plan = handle
        nbytes = int(lib.tdx_mat_workspace_bytes(handle))
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        self.numel = prob.n_eq * t.N

    @property
    def specialized_points(self) -> int:
        """Grid points evaluated by the kernel generated for this program."""
        return int(native.lib().tdx_mat_specialized_points(self._plan))

    def __del__(self):
        try:
            if getattr(self, '_plan', None) is not None and self._plan.value:
                native.lib().tdx_mat_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass

    def _check(self, u: torch.Tensor) -> None:
        if u.device.type != 'cuda' or u.dtype != self.dtype or not u.is_contiguous() \
                or u.numel() != self.numel:
            raise ValueError("u must be a contiguous device tensor of the grid's shape and dtype")

    def forward(self, u, loss, op=None, bval=None) -> None:
        self._check(u)
        native.check(native.lib().tdx_mat_forward(
            self._plan, native.ptr(u), native.ptr(loss), native.ptr(op), native.ptr(bval),
            native.ptr(self.workspace), native.stream_ptr()))

    def backward(self, u, dloss, grad) -> None:
        self._check(u)
        native.check(native.lib().tdx_mat_backward(
            self._plan, native.ptr(u), native.ptr(dloss), native.ptr(grad),
            native.ptr(self.workspace), native.stream_ptr()))

    def train_step(self, u, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps, loss=None):
        self._check(u)
        native.check(native.lib().tdx_mat_train_step(
            self._plan, native.ptr(u), native.ptr(exp_avg), native.ptr(exp_avg_sq),
            native.ptr(step), float(lr), float(beta1), float(beta2), float(eps),
            native.ptr(loss), native.ptr(self.workspace), native.stream_ptr()))

    def train_step_to(self, u_in, u_out, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps,
                      loss=None):
        """One fused Adam step from u_in into u_out (u_in unchanged): the
        result of copying u_in to u_out and calling train_step on it."""
        self._check(u_in)
        self._check(u_out)
        native.check(native.lib().tdx_mat_train_step_to(
            self._plan, native.ptr(u_in), native.ptr(u_out), native.ptr(exp_avg),
            native.ptr(exp_avg_sq), native.ptr(step), float(lr), float(beta1), float(beta2),
            float(eps), native.ptr(loss), native.ptr(self.workspace), native.stream_ptr()))

    def loss(self, u: torch.Tensor) -> torch.Tensor:
        return MatResidualLoss.apply(u, self)

    def fields(self, u: torch.Tensor):
        """(loss[2], op field [n_eq, N], boundary values [n_bc])."""
        loss = torch.empty(2, dtype=self.dtype, device=self.device)
        op = torch.empty(self.prob.n_eq * self.tables.N, dtype=self.dtype, device=self.device)
        bval = torch.empty(max(self.tables.n_bc, 1), dtype=self.dtype, device=self.device)
        self.forward(u.detach().contiguous(), loss, op, bval)
        return loss, op.view(self.prob.n_eq, -1), bval[:self.tables.n_bc]


class MatResidualLoss(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u: torch.Tensor, engine: MatEngine):
        x = u.detach().contiguous()
        loss = torch.empty(2, dtype=engine.dtype, device=engine.device)
        engine.forward(x, loss)
        ctx.engine = engine
        ctx.save_for_backward(x)
        ctx.shape = u.shape
        return loss

    @staticmethod
    def backward(ctx, grad_out):
        (x,) = ctx.saved_tensors
        eng: MatEngine = ctx.engine
        dloss = grad_out[:1].contiguous().to(eng.dtype)
        grad = torch.empty_like(x)
        eng.backward(x, dloss, grad)
        return grad.view(ctx.shape), None
