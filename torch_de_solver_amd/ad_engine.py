"""'autograd' mode engine: table packing, the native plan and its autograd
Function.

The reference's 'autograd' mode differentiates the network with nested
``torch.autograd.grad`` calls (tedeous/derivative.py:61-132).  ``ADEngine``
serves it with the native jet kernels (``tdx_ad_*``): one pass carries the
value and every partial derivative the program reads through the network,
the residual program runs on the output jets, and a hand-written adjoint
returns the parameter gradient.
"""

import ctypes
from typing import List, Optional

import numpy as np
import torch

from . import native
from .compiler import AD_OP, ADProblem
from .nn_engine import MLPArch


class ADTables:
    """Flat int32 / float64 / working-precision tables of one problem
    (layout documented with ``tdx_ad_desc`` in include/tdx.h)."""

    def __init__(self, prob: ADProblem, np_dtype):
        ints: List[int] = []
        reals: List[float] = []
        data: List[np.ndarray] = []
        n_data = [0]

        def add_real(v: float) -> int:
            reals.append(float(v))
            return len(reals) - 1

        def add_data(arr) -> int:
            a = np.asarray(arr, dtype=np_dtype).reshape(-1)
            off = n_data[0]
            data.append(a)
            n_data[0] += a.size
            return off

        jets = prob.jets
        J = len(jets)
        jet_index = {j: i for i, j in enumerate(jets)}
        # jets
        self.jet_i = len(ints)
        for j in jets:
            ints += [len(j), j[0] if len(j) == 1 else -1]
        parts = prob.partitions()
        self.pptr_i = len(ints)
        k = 0
        for p in parts:
            ints.append(k)
            k += len(p)
        ints.append(k)
        self.part_i = len(ints)
        for p in parts:
            for blocks in p:
                ints += [len(blocks)] + blocks + [-1] * (4 - len(blocks))
        # points: every set of every group, back to back
        pts = []
        n_pts = 0
        set_offsets = []
        for g in prob.groups:
            offs = []
            for s in g.sets:
                offs.append(n_pts)
                pts.append(np.asarray(s, dtype=np_dtype).reshape(-1, prob.d_in))
                n_pts += s.shape[0]
            set_offsets.append(offs)
        self.points = np.ascontiguousarray(np.concatenate(pts)) if pts else \
            np.zeros((0, prob.d_in), dtype=np_dtype)
        self.n_points = n_pts
        # residual offsets: per group, per equation
        self.resid = []   # (group, eq, offset, count, column, col_pos)
        n_resid = 0
        for gi, g in enumerate(prob.groups):
            ne = g.sets[0].shape[0]
            for qi, eq in enumerate(g.equations):
                self.resid.append((gi, qi, n_resid, ne, eq.column, eq.col_pos))
                n_resid += ne
        self.n_resid = n_resid
        # groups and their programs: headers first, records appended after
        self.groups_i = len(ints)
        ints += [0] * (5 * len(prob.groups))
        r_iter = iter(self.resid)
        for gi, g in enumerate(prob.groups):
            ne = g.sets[0].shape[0]
            set_i = len(ints)
            ints += set_offsets[gi]
            eq_i = len(ints)
            ints += [0] * (7 * len(g.equations))
            for qi, eq in enumerate(g.equations):
                _, _, roff, _, _, _ = next(r_iter)
                term_i = len(ints)
                ints += [0] * (4 * len(eq.terms))
                for ti, t in enumerate(eq.terms):
                    fac_i = len(ints)
                    ints += [0] * (5 * len(t.factors))
                    for fi, f in enumerate(t.factors):
                        st_i = len(ints)
                        for (s, jet, w) in f.stencil:
                            ints += [s, jet_index[jet], add_real(w)]
                        if f.pow_kind == 2:
                            # callable power: n, then (op, arg) pairs; arg
                            # the integer power (POWI) or a real's index
                            pref = len(ints)
                            ints.append(len(f.pow_val))
                            for op, arg in f.pow_val:
                                ints += [op, int(arg) if op == AD_OP['POWI'] else
                                         (add_real(arg) if op <= AD_OP['RDIV'] or
                                          op == AD_OP['POWR'] else 0)]
                        else:
                            pref = int(f.pow_val) if f.pow_kind == 0 else add_real(f.pow_val)
                        ints[fac_i + 5 * fi: fac_i + 5 * fi + 5] = [len(f.stencil), st_i, f.var,
                                                                     f.pow_kind, pref]
                    if t.coeff_array is not None:
                        ck, cr = 1, add_data(t.coeff_array)
                    else:
                        ck, cr = 0, add_real(t.coeff_scalar)
                    ints[term_i + 4 * ti: term_i + 4 * ti + 4] = [len(t.factors), fac_i, ck, cr]
                tgt = add_data(eq.target) if eq.target is not None else -1
                if prob.weak_levels and gi == 0:
                    # weak form: sum_eq lambda W_eq, formed after the forward
                    # pass; the seeds carry lambda here and dW/dr / (2 r) there
                    lw, nw, sw = 0.0, 0.0, eq.lam
                elif prob.causal_tol and gi == 0:
                    # causal loss: the operator term is mean(w res), formed
                    # after the forward pass; the seeds carry 1/N here and w there
                    lw, nw, sw = 0.0, 0.0, 1.0 / eq.denom
                else:
                    lw, nw, sw = eq.lam / eq.denom, 1.0 / eq.denom, eq.lam / eq.denom
                ints[eq_i + 7 * qi: eq_i + 7 * qi + 7] = [len(eq.terms), term_i, add_real(lw),
                                                          add_real(nw), tgt, roff, add_real(sw)]
            ints[self.groups_i + 5 * gi: self.groups_i + 5 * gi + 5] = [
                ne, len(g.sets), set_i, len(g.equations), eq_i]
        self.weak_d = self.weak_i = 0
        if prob.weak_levels:
            lams = [eq.lam for eq in prob.groups[0].equations]
            self.weak_d = add_data(np.concatenate([prob.weak_ctf2, prob.weak_cx, lams]))
            self.weak_i = len(ints)
            ints += [int(v) for v in prob.weak_gstart]
        self.n_jets = J
        self.ints = np.asarray(ints, dtype=np.int32)
        self.reals = np.asarray(reals, dtype=np.float64)
        self.data = np.concatenate(data).astype(np_dtype) if data else np.zeros(1, dtype=np_dtype)


    def desc(self, prob: ADProblem, arch: MLPArch) -> 'native.ADDesc':
        """The plan descriptor of include/tdx.h over these tables (which must
        outlive the plan-create call)."""
        desc = native.ADDesc()
        desc.dtype = 0 if prob.dtype == torch.float32 else 1
        desc.d_in = arch.d_in
        desc.n_out = arch.n_out
        desc.width = arch.width
        desc.n_layers = arch.n_layers
        desc.n_jets = self.n_jets
        desc.jet_i = self.jet_i
        desc.pptr_i = self.pptr_i
        desc.part_i = self.part_i
        desc.n_points = self.n_points
        desc.points = self.points.ctypes.data
        desc.n_groups = len(prob.groups)
        desc.groups_i = self.groups_i
        desc.n_ints = self.ints.size
        desc.ints = self.ints.ctypes.data_as(native.c_i32p)
        desc.n_reals = self.reals.size
        desc.reals = self.reals.ctypes.data_as(native.c_dp)
        desc.n_data = self.data.size
        desc.data = self.data.ctypes.data
        desc.n_resid = self.n_resid
        desc.causal_tol = float(prob.causal_tol)
        desc.causal_nt = int(prob.causal_nt)
        if prob.weak_levels:
            desc.weak_levels, desc.weak_runs = prob.weak_levels, len(prob.weak_cx)
            desc.weak_d, desc.weak_i = self.weak_d, self.weak_i
        return desc


def ad_step_source(prob: 'ADProblem', arch: 'MLPArch', compile: bool = False) -> str:
    """The generated source of the fused 'autograd' step kernel for a
    problem (no device needed; with ``compile`` also compiled for gfx950),
    '' when the fused step does not serve it."""
    t = ADTables(prob, np.float32 if prob.dtype == torch.float32 else np.float64)
    desc = t.desc(prob, arch)
    lib = native.lib()
    n = int(lib.tdx_ad_step_source(ctypes.byref(desc), None, 0, int(compile)))
    native.check(min(n, 0))
    if n == 0:
        return ''
    buf = ctypes.create_string_buffer(n + 1)
    lib.tdx_ad_step_source(ctypes.byref(desc), buf, n + 1, 0)
    return buf.value.decode()


class ADEngine:
    """One native 'autograd' plan (one network instance)."""

    def __init__(self, prob: ADProblem, arch: MLPArch, device: torch.device):
        if prob.dtype not in (torch.float32, torch.float64):
            raise NotImplementedError(f"dtype {prob.dtype} is not supported")
        if arch.d_in != prob.d_in or arch.n_out < prob.n_out:
            raise ValueError("network shape does not match the problem")
        self.problem = prob
        self.arch = arch
        self.device = device
        self.dtype = prob.dtype
        self.n_inst = 1
        np_dtype = np.float32 if prob.dtype == torch.float32 else np.float64
        self.tables = t = ADTables(prob, np_dtype)
        lib = native.lib()
        self._keep = (t.ints, t.reals, t.data, t.points)
        desc = t.desc(prob, arch)
        handle = ctypes.c_void_p()
        with torch.cuda.device(device):
            native.check(lib.tdx_ad_plan_create(ctypes.byref(desc), ctypes.byref(handle)))
        self.handle = handle
        self.n_params = int(lib.tdx_ad_param_count(handle))
        nbytes = int(lib.tdx_ad_workspace_bytes(handle))
        self.workspace = torch.empty(max(nbytes, 1), dtype=torch.uint8, device=device)
        # forward passes so far: the adjoint reads the jets and seeds of the
        # latest one, so a backward checks that it belongs to its forward
        self.fwd_gen = 0

    def __del__(self):
        h = getattr(self, 'handle', None)
        if h is not None and h.value:
            try:
                native.lib().tdx_ad_plan_destroy(h)
            except Exception:
                pass
            self.handle = None

    def _check(self, params: torch.Tensor) -> None:
        if params.device.type != 'cuda' or params.dtype != self.dtype or \
                not params.is_contiguous() or params.numel() != self.n_params:
            raise ValueError("params must be a contiguous device tensor of "
                             f"{self.n_params} {self.dtype} values")

    def forward(self, params: torch.Tensor, loss: torch.Tensor,
                resid: Optional[torch.Tensor] = None) -> None:
        self._check(params)
        native.check(native.lib().tdx_ad_forward(
            self.handle, native.ptr(params), native.ptr(loss), native.ptr(resid),
            native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1

    def backward(self, params: torch.Tensor, dloss: torch.Tensor, dparams: torch.Tensor) -> None:
        self._check(params)
        native.check(native.lib().tdx_ad_backward(
            self.handle, native.ptr(params), native.ptr(dloss), native.ptr(dparams),
            native.ptr(self.workspace), native.stream_ptr()))

    @property
    def step_kernel(self) -> str:
        """'fused' when train_step runs the one-kernel step interpreting the
        tables (k_ad_step), 'fused-generated' when it runs that kernel
        compiled with the plan's generated jet maps and residual programs
        (tdx_ad_step_jit), 'layers' for the layer kernels and the separate
        Adam pass."""
        k = native.lib().tdx_ad_step_kernel(self.handle)
        return {1: 'fused', 2: 'fused-generated'}.get(k, 'layers')

    def train_step(self, params, exp_avg, exp_avg_sq, step, lr, beta1, beta2, eps, loss) -> None:
        self._check(params)
        native.check(native.lib().tdx_ad_train_step(
            self.handle, native.ptr(params), native.ptr(exp_avg), native.ptr(exp_avg_sq),
            native.ptr(step), float(lr), float(beta1), float(beta2), float(eps),
            native.ptr(loss), native.ptr(self.workspace), native.stream_ptr()))
        self.fwd_gen += 1

    def jacobian(self, params: torch.Tensor) -> torch.Tensor:
        """``[P, n_resid]``: d resid[r] / d params[p] of every residual entry
        (the transposed J of the natural gradient)."""
        self._check(params)
        n = self.tables.n_resid
        if self.n_params * n * params.element_size() > (16 << 30):
            raise ValueError(f"the Jacobian would take {self.n_params} x {n} values")
        jt = torch.empty(self.n_params, n, dtype=self.dtype, device=self.device)
        native.check(native.lib().tdx_ad_jacobian(
            self.handle, native.ptr(params), native.ptr(jt), native.stream_ptr()))
        return jt

    def loss(self, params: torch.Tensor) -> torch.Tensor:
        """``[1, 2]``: (lambda-weighted loss, normalized loss), differentiable
        in ``params`` through the native adjoint."""
        return ADResidualLoss.apply(params, self)

    def residuals(self, params: torch.Tensor) -> torch.Tensor:
        p = params.detach().contiguous()
        loss = torch.empty(2, dtype=self.dtype, device=self.device)
        resid = torch.empty(max(self.tables.n_resid, 1), dtype=self.dtype, device=self.device)
        self.forward(p, loss, resid)
        return resid


class ADResidualLoss(torch.autograd.Function):
    """Loss of ``Solution.evaluate`` in 'autograd' mode
    (tedeous/solution.py:129-168) with the gradient of ``loss.backward()``
    (tedeous/optimizers/closure.py:60)."""

    @staticmethod
    def forward(ctx, params: torch.Tensor, engine: ADEngine):
        p = params.detach().contiguous()
        loss = torch.empty(2, dtype=engine.dtype, device=engine.device)
        engine.forward(p, loss)
        ctx.engine = engine
        ctx.gen = engine.fwd_gen
        ctx.save_for_backward(p)
        return loss.view(1, 2)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (p,) = ctx.saved_tensors
        eng: ADEngine = ctx.engine
        if eng.fwd_gen != ctx.gen:
            # another forward pass ran since this node's: restore its jets
            # and seeds before the adjoint
            eng.forward(p, torch.empty(2, dtype=eng.dtype, device=eng.device))
        # the normalized loss carries no gradient (tedeous/losses.py:124-126)
        dloss = grad_out.reshape(2)[:1].contiguous().to(eng.dtype)
        dparams = torch.empty_like(p)
        eng.backward(p, dloss, dparams)
        return dparams, None
