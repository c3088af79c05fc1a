"""'NN' mode engine: table packing, the native plan and its autograd Function.

``NNEngine`` owns one native plan for one or more instances of the same
network architecture (a single solve is one instance; a parameter sweep is
many).  Its ``loss`` goes through ``NNResidualLoss``, a
``torch.autograd.Function`` whose forward is ``tdx_nn_forward`` and whose
backward is ``tdx_nn_backward``; ``train_step`` is the fused native step
(forward, adjoint and Adam) that a HIP graph can capture.
"""

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np
import torch

from . import native
from .compiler import AD_OP, NNProblem


# ----------------------------------------------------------------------------
# network description
# ----------------------------------------------------------------------------

@dataclass
class MLPArch:
    d_in: int
    width: int
    n_out: int
    n_layers: int

    @property
    def n_params(self) -> int:
        H, d, o, L = self.width, self.d_in, self.n_out, self.n_layers
        return H * d + H + (L - 2) * (H * H + H) + o * H + o


def mlp_arch(net: torch.nn.Module) -> MLPArch:
    """Check that ``net`` is Linear/Tanh/.../Linear with one hidden width."""
    layers = list(net) if isinstance(net, torch.nn.Sequential) else None
    if not layers:
        raise NotImplementedError("the native 'NN' engine needs an nn.Sequential "
                                  "of Linear and Tanh layers")
    lin = [m for m in layers if isinstance(m, torch.nn.Linear)]
    kinds = ['L' if isinstance(m, torch.nn.Linear) else
             'T' if isinstance(m, torch.nn.Tanh) else '?' for m in layers]
    expect = 'LT' * (len(lin) - 1) + 'L'
    if ''.join(kinds) != expect or len(lin) < 2:
        raise NotImplementedError(
            f"unsupported network layout {''.join(kinds)}: expected {expect} "
            "(Linear layers separated by Tanh)")
    width = lin[0].out_features
    for m in lin[:-1]:
        if m.out_features != width or m.bias is None:
            raise NotImplementedError("all hidden layers must have the same width and a bias")
    for m in lin[1:-1]:
        if m.in_features != width:
            raise NotImplementedError("hidden layers must be width x width")
    if lin[-1].bias is None:
        raise NotImplementedError("the output layer needs a bias")
    return MLPArch(d_in=lin[0].in_features, width=width,
                   n_out=lin[-1].out_features, n_layers=len(lin))


# ----------------------------------------------------------------------------
# table packing
# ----------------------------------------------------------------------------

@dataclass
class ResidSlot:
    inst: int
    group: int
    eq: int
    kind: str
    offset: int
    count: int
    column: int
    col_pos: Optional[np.ndarray]


class NNTables:
    """Flat tables of include/tdx.h for a list of instances."""

    def __init__(self, problems: Sequence[NNProblem], arch: MLPArch, np_dtype):
        self.np_dtype = np_dtype
        groups: List[List[int]] = []
        ints: List[int] = []
        reals: List[float] = []
        data_parts: List[np.ndarray] = []
        self.n_data = 0
        self.resid: List[ResidSlot] = []
        n_resid = 0
        row = 0

        def add_data(arr) -> int:
            arr = np.ascontiguousarray(np.asarray(arr, dtype=np_dtype).reshape(-1))
            off = self.n_data
            data_parts.append(arr)
            self.n_data += arr.size
            return off

        def add_real(v) -> int:
            reals.append(float(v))
            return len(reals) - 1

        for inst, prob in enumerate(problems):
            if prob.d_in != arch.d_in:
                raise ValueError("grid dimension does not match the network input")
            for gi, g in enumerate(prob.groups):
                base_off = [add_data(b) for b in g.bases]
                slot_base_i = len(ints)
                ints.extend(base_off[b] for b in g.slot_base)
                slot_off_r = len(reals)
                for off in g.slot_offset:
                    reals.extend(float(np_dtype(o)) for o in off)
                # equation records are reserved first, terms after
                eq_i = len(ints)
                ints.extend([0] * (len(g.equations) * native.TDX_EQ_INTS))
                for qi, eq in enumerate(g.equations):
                    term_i = len(ints)
                    ints.extend([0] * (len(eq.terms) * native.TDX_TERM_INTS))
                    for ti, t in enumerate(eq.terms):
                        factor_i = len(ints)
                        ints.extend([0] * (len(t.factors) * native.TDX_FACTOR_INTS))
                        for fi, f in enumerate(t.factors):
                            st_i = len(ints)
                            for slot, w in f.stencil:
                                ints.extend([slot, add_real(np_dtype(w))])
                            if f.pow_kind == 2:
                                # callable power: n, then (op, arg) pairs at
                                # the record's integer word
                                pow_i, pow_r = len(ints), 0
                                ints.append(len(f.pow_val))
                                for op, arg in f.pow_val:
                                    ints.extend([op, int(arg) if op == AD_OP['POWI'] else
                                                 (add_real(np_dtype(arg)) if op <= AD_OP['RDIV']
                                                  or op == AD_OP['POWR'] else 0)])
                            else:
                                pow_r = add_real(f.pow_val)
                                pow_i = int(f.pow_val) if f.pow_kind == 0 else 0
                            rec = [f.var, f.pow_kind, pow_i, len(f.stencil), st_i, pow_r]
                            ints[factor_i + fi * native.TDX_FACTOR_INTS:
                                 factor_i + (fi + 1) * native.TDX_FACTOR_INTS] = rec
                        if t.coeff_array is not None:
                            ck, ci = 1, add_data(t.coeff_array)
                        else:
                            ck, ci = 0, add_real(np_dtype(t.coeff_scalar))
                        ints[term_i + ti * native.TDX_TERM_INTS:
                             term_i + (ti + 1) * native.TDX_TERM_INTS] = [len(t.factors), factor_i, ck, ci]
                    target_off = add_data(eq.target) if eq.target is not None else -1
                    if prob.weak_levels and gi == 0:
                        # weak form: the operator term is sum_eq lambda W_eq,
                        # formed after the forward pass; its seeds carry
                        # lambda here and the trapezoid factors there
                        lam_r = add_real(0.0)
                        inv_r = add_real(0.0)
                        seed_r = add_real(eq.lam)
                    elif prob.causal_tol and gi == 0:
                        # causal loss: the operator term is mean(w res),
                        # formed after the forward pass; its seeds carry
                        # 1/N here and w there
                        lam_r = add_real(0.0)
                        inv_r = add_real(0.0)
                        seed_r = add_real(1.0 / eq.denom)
                    else:
                        lam_r = add_real(eq.lam / eq.denom)
                        inv_r = add_real(1.0 / eq.denom)
                        seed_r = lam_r
                    self.resid.append(ResidSlot(inst, gi, qi, g.kind, n_resid, g.n_entries,
                                                eq.column, eq.col_pos))
                    ints[eq_i + qi * native.TDX_EQ_INTS: eq_i + (qi + 1) * native.TDX_EQ_INTS] = [
                        len(eq.terms), term_i, target_off, n_resid, lam_r, inv_r, seed_r]
                    n_resid += g.n_entries
                groups.append([inst, g.n_entries, len(g.slot_base), slot_base_i, slot_off_r,
                               len(g.equations), eq_i, row])
                row += len(g.slot_base) * g.n_entries
        # weak-form tables of the first instance (one instance per plan)
        self.weak_levels = self.weak_runs = self.weak_i = 0
        self.weak_d = 0
        prob = problems[0]
        if prob.weak_levels:
            lams = [eq.lam for eq in prob.groups[0].equations]
            self.weak_levels = prob.weak_levels
            self.weak_runs = len(prob.weak_cx)
            self.weak_d = add_data(np.concatenate([prob.weak_ctf2, prob.weak_cx, lams]))
            self.weak_i = len(ints)
            ints.extend(int(v) for v in prob.weak_gstart)
        self.groups = np.asarray(groups, dtype=np.int32).reshape(-1)
        self.n_groups = len(groups)
        self.ints = np.asarray(ints if ints else [0], dtype=np.int32)
        self.reals = np.asarray(reals if reals else [0.0], dtype=np_dtype)
        self.data = (np.concatenate(data_parts) if data_parts
                     else np.zeros(1, dtype=np_dtype)).astype(np_dtype)
        self.n_resid = n_resid
        self.n_rows = row
        self.n_inst = len(problems)
        self.causal = (float(prob.causal_tol), int(prob.causal_nt)) if prob.causal_tol else None

    def fused_source(self, arch: MLPArch, compile: bool = False) -> str:
        """Source of the fused training-step kernel with the residual
        programs generated from these tables ('' when a plan of them has no
        fused step; the reason is in tdx_last_error); ``compile`` also
        compiles it for gfx950 (hipRTC, no device needed)."""
        desc = self.desc(arch)
        lib = native.lib()
        n = int(lib.tdx_nn_fused_source(ctypes.byref(desc), None, 0, 0))
        if n < 0:
            native.check(n)
        if n == 0:
            return ''
        buf = ctypes.create_string_buffer(n + 1)
        rc = int(lib.tdx_nn_fused_source(ctypes.byref(desc), buf, n + 1, 1 if compile else 0))
        if rc < 0:
            native.check(rc)
        return buf.value.decode()

    def desc(self, arch: MLPArch, tile_rows: int = 0, chunk_tiles: int = 0,
             persistent: int = 0) -> 'native.NNDesc':
        """The plan descriptor of include/tdx.h over these tables (which must
        outlive the plan-create call)."""
        desc = native.NNDesc()
        desc.dtype = 0 if self.np_dtype == np.float32 else 1
        desc.d_in, desc.n_out = arch.d_in, arch.n_out
        desc.width, desc.n_layers = arch.width, arch.n_layers
        desc.n_inst = self.n_inst
        desc.n_groups = self.n_groups
        desc.groups = self.groups.ctypes.data_as(native.c_i32p)
        desc.n_ints = self.ints.size
        desc.ints = self.ints.ctypes.data_as(native.c_i32p)
        desc.n_reals = self.reals.size
        desc.reals = self.reals.ctypes.data
        desc.n_data = self.data.size
        desc.data = self.data.ctypes.data
        desc.n_resid = self.n_resid
        desc.tile_rows = tile_rows
        desc.chunk_tiles = chunk_tiles
        desc.persistent = persistent
        if self.causal:
            if self.n_inst != 1:
                raise NotImplementedError("the causal loss serves one instance per plan")
            desc.causal_tol, desc.causal_nt = self.causal
        if self.weak_levels:
            if self.n_inst != 1:
                raise NotImplementedError("the weak form serves one instance per plan")
            desc.weak_levels, desc.weak_runs = self.weak_levels, self.weak_runs
            desc.weak_d, desc.weak_i = self.weak_d, self.weak_i
        return desc


# ----------------------------------------------------------------------------
# native plan
# ----------------------------------------------------------------------------

class NNEngine:
    """Native 'NN' residual/loss engine for ``len(problems)`` instances."""

    def __init__(self, problems: Sequence[NNProblem], arch: MLPArch,
                 device: torch.device, tile_rows: int = 0, chunk_tiles: int = 0,
                 persistent: int = 0):
        if not problems:
            raise ValueError("no problems")
        dtype = problems[0].dtype
        if dtype not in (torch.float32, torch.float64):
            raise NotImplementedError(f"dtype {dtype} is not supported")
        self.dtype = dtype
        self.np_dtype = np.float32 if dtype == torch.float32 else np.float64
        self.arch = arch
        self.problems = list(problems)
        self.n_inst = len(problems)
        self.device = torch.device(device)
        if self.device.type != 'cuda':
            raise native.NativeUnavailable("the native engine runs on a HIP device only")
        self.tables = NNTables(problems, arch, self.np_dtype)
        t = self.tables
        lib = native.lib()
        desc = t.desc(arch, tile_rows, chunk_tiles, persistent)
        handle = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            native.check(lib.tdx_nn_plan_create(ctypes.byref(desc), ctypes.byref(handle)))
        self._plan = handle
        self.n_params = int(lib.tdx_nn_param_count(handle))
        assert self.n_params == arch.n_params
        nbytes = int(lib.tdx_nn_workspace_bytes(handle))
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=self.device)
        self.loss_buf = torch.zeros(2 * self.n_inst, dtype=dtype, device=self.device)
        self.resid_buf = torch.zeros(max(t.n_resid, 1), dtype=dtype, device=self.device)
        self.n_rows = int(lib.tdx_nn_row_count(handle))
        # forward passes so far: the adjoint reads the seeds of the latest
        # one, so a backward checks that it belongs to its forward
        self.fwd_gen = 0

    def __del__(self):
        try:
            if getattr(self, '_plan', None) is not None and self._plan.value:
                native.lib().tdx_nn_plan_destroy(self._plan)
                self._plan = None
        except Exception:
            pass

    # -- raw entry points ----------------------------------------------------
    def forward(self, params: torch.Tensor, loss: torch.Tensor,
                resid: Optional[torch.Tensor] = None) -> None:
        self._check_params(params)
        native.check(native.lib().tdx_nn_forward(
            self._plan, native.ptr(params), native.ptr(loss),
            native.ptr(resid) if resid is not None else None,
            native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1

    def network_values(self, params: torch.Tensor) -> torch.Tensor:
        """Network output of every (group, slot, entry) row ``[n_rows, n_out]``
        (row ``row_base + slot * n_entries + entry``), from a forward pass."""
        self._check_params(params)
        out = torch.empty(self.n_rows, self.arch.n_out, dtype=self.dtype, device=self.device)
        native.check(native.lib().tdx_nn_network_values(
            self._plan, native.ptr(params), native.ptr(out), None,
            native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1
        return out

    def forward_values(self, values: torch.Tensor, loss: torch.Tensor,
                       resid: Optional[torch.Tensor] = None,
                       seeds: Optional[torch.Tensor] = None) -> None:
        """The residual program on given network values ``[n_rows, n_out]``:
        losses, residuals and the seeds dL/du of every row (to ``seeds`` when
        given, else to the workspace for a following backward)."""
        if values.device.type != 'cuda' or values.dtype != self.dtype or \
                not values.is_contiguous() or values.numel() != self.n_rows * self.arch.n_out:
            raise ValueError(f"values must be a contiguous [{self.n_rows}, {self.arch.n_out}] "
                             "device tensor of the problem dtype")
        if seeds is not None and (seeds.dtype != self.dtype or not seeds.is_contiguous() or
                                  seeds.numel() != values.numel()):
            raise ValueError("seeds must match values")
        native.check(native.lib().tdx_nn_forward_values(
            self._plan, native.ptr(values), native.ptr(loss),
            native.ptr(resid) if resid is not None else None,
            native.ptr(seeds) if seeds is not None else None,
            native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1

    def jacobian(self, params: torch.Tensor) -> torch.Tensor:
        """``[P, n_resid]``: d resid[r] / d params[p] of every residual entry
        (the transposed J of the natural gradient), one instance."""
        if self.n_inst != 1:
            raise NotImplementedError("the Jacobian serves one instance per plan")
        self._check_params(params)
        n = self.tables.n_resid
        if self.n_params * n * params.element_size() > (16 << 30):
            raise ValueError(f"the Jacobian would take {self.n_params} x {n} values")
        jt = torch.empty(self.n_params, n, dtype=self.dtype, device=self.device)
        native.check(native.lib().tdx_nn_jacobian(
            self._plan, native.ptr(params), native.ptr(jt), native.stream_ptr()))
        return jt

    def backward(self, params: torch.Tensor, dloss: torch.Tensor,
                 dparams: torch.Tensor) -> None:
        self._check_params(params)
        assert dparams.numel() == params.numel() and dparams.is_contiguous()
        native.check(native.lib().tdx_nn_backward(
            self._plan, native.ptr(params), native.ptr(dloss), native.ptr(dparams),
            native.ptr(self.workspace), native.stream_ptr()))

    def train_step(self, params, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps,
                   loss: Optional[torch.Tensor] = None) -> None:
        self._check_params(params)
        for t in (exp_avg, exp_avg_sq):
            assert t.shape == params.shape and t.dtype == params.dtype and t.is_contiguous()
        assert step.dtype == torch.int64 and step.device == params.device
        native.check(native.lib().tdx_nn_train_step(
            self._plan, native.ptr(params), native.ptr(exp_avg), native.ptr(exp_avg_sq),
            native.ptr(step), float(lr), float(beta1), float(beta2), float(eps),
            native.ptr(loss) if loss is not None else None,
            native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1

    @property
    def step_kernel(self) -> str:
        """The training step's kernels: 'separate' (forward, then adjoint),
        'fused' (one kernel interpreting the residual programs),
        'fused-generated' (one kernel compiled with the plan's generated
        residual programs), 'adjoint-seeds' (wide networks: the
        block-owning adjoint forms its own seeds, no forward kernel) or
        'adjoint-seeds-generated' (the same compiled with generated
        programs)."""
        k = int(native.lib().tdx_nn_step_kernel(self._plan))
        return {0: 'separate', 1: 'fused', 2: 'fused-generated', 3: 'adjoint-seeds',
                4: 'adjoint-seeds-generated'}[k]

    def profile(self, capacity: int) -> None:
        """Time up to ``capacity`` launches per kernel kind with HIP events
        (0 turns timing off)."""
        native.check(native.lib().tdx_nn_profile(self._plan, int(capacity)))

    def profile_read(self):
        """``{kind: (total_ms, launches)}`` for 'forward', 'backward', 'update'."""
        ms = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int * 3)()
        native.check(native.lib().tdx_nn_profile_read(self._plan, ms, cnt))
        return {k: (ms[i], cnt[i]) for i, k in enumerate(('forward', 'backward', 'update'))}

    def _check_params(self, params: torch.Tensor) -> None:
        if params.device.type != 'cuda' or params.dtype != self.dtype or not params.is_contiguous():
            raise ValueError("params must be a contiguous device tensor of the problem dtype")
        if params.numel() != self.n_inst * self.n_params:
            raise ValueError(f"params has {params.numel()} values, expected "
                             f"{self.n_inst * self.n_params}")

    # -- autograd ------------------------------------------------------------
    def loss(self, params: torch.Tensor) -> torch.Tensor:
        """Per-instance ``[loss, loss_normalized]`` pairs ``[n_inst, 2]``;
        differentiable in ``params`` through the native adjoint."""
        return NNResidualLoss.apply(params, self)

    def residuals(self, params: torch.Tensor) -> torch.Tensor:
        """Residual values of every (group, equation, entry) in table order."""
        loss = torch.empty_like(self.loss_buf)
        self.forward(params.detach().contiguous(), loss, self.resid_buf)
        return self.resid_buf


class NNResidualLoss(torch.autograd.Function):
    """Loss of ``Solution.evaluate`` (tedeous/solution.py:129-168) with the
    gradient of ``loss.backward()`` (tedeous/optimizers/closure.py:60)."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, engine: NNEngine):
        p = params.detach().contiguous()
        loss = torch.empty(2 * engine.n_inst, dtype=engine.dtype, device=engine.device)
        engine.forward(p, loss)
        ctx.engine = engine
        ctx.gen = engine.fwd_gen
        ctx.save_for_backward(p)
        return loss.view(engine.n_inst, 2)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (p,) = ctx.saved_tensors
        eng: NNEngine = ctx.engine
        if eng.fwd_gen != ctx.gen:
            # another forward pass ran since this node's: restore its seeds
            # before the adjoint
            eng.forward(p, torch.empty(2 * eng.n_inst, dtype=eng.dtype, device=eng.device))
        g = grad_out.reshape(eng.n_inst, 2)
        # the normalized loss carries no gradient (computed under no_grad in
        # tedeous/losses.py:124-126); instances share one adjoint call, which
        # needs one upstream scale
        scale = g[:, 0]
        if eng.n_inst > 1 and not bool(torch.all(scale == scale[0])):
            raise NotImplementedError("per-instance upstream gradients must be equal")
        dloss = scale[:1].contiguous().to(eng.dtype)
        dparams = torch.empty_like(p)
        eng.backward(p, dloss, dparams)
        return dparams, None


# ----------------------------------------------------------------------------
# flat parameter storage for nn.Module networks
# ----------------------------------------------------------------------------

def flatten_module_params(net: torch.nn.Module) -> torch.Tensor:
    """Move the parameters of ``net`` into one contiguous buffer in
    ``net.parameters()`` order and make every parameter a view of it, so
    that the native kernels and torch optimizers see the same storage."""
    params = list(net.parameters())
    flat = torch.cat([p.detach().reshape(-1) for p in params]).contiguous()
    off = 0
    for p in params:
        n = p.numel()
        p.data = flat[off:off + n].view_as(p)
        off += n
    return flat


def module_flat_view(net: torch.nn.Module) -> Optional[torch.Tensor]:
    """The shared flat buffer if ``flatten_module_params`` was applied."""
    params = list(net.parameters())
    if not params:
        return None
    base = params[0].data
    st = base.untyped_storage()
    off = 0
    for p in params:
        if p.data.untyped_storage().data_ptr() != st.data_ptr() or p.data.storage_offset() != off \
                or not p.data.is_contiguous():
            return None
        off += p.numel()
    return torch.empty(0, dtype=base.dtype, device=base.device).set_(
        st, 0, (off,), (1,))
