"""``Solution``: the per-step objective, served by the native engines.

Drop-in for tedeous/solution.py:19-168 in the 'NN', 'mat' and 'autograd'
modes.  ``evaluate()`` returns
``(loss, loss_normalized)`` with shape ``[1]`` each, like the reference
(tedeous/losses.py:119-135), and differentiates through the native adjoint.
The reference's side outputs ``op``, ``bval``, ``true_bval``, ``bval_keys``
and ``bval_length`` (read by callbacks, tedeous/callbacks/adaptive_lambda.py:
90-105) are computed on demand from the native residual output.
"""

from typing import List, Optional, Tuple, Union

import torch

from .ad_engine import ADEngine
from .compiler import ADProblem, MatProblem, NNProblem, compile_ad, compile_mat, compile_nn
from .input_preprocessing import Operator_bcond_preproc, _Prepared
from .mat_engine import MatEngine
from .nn_engine import MLPArch, NNEngine, mlp_arch


def _operators(op) -> List[dict]:
    if isinstance(op, dict):
        return [op]
    return list(op)


def _lambda_key(lam) -> Tuple:
    """Value identity of a lambda argument (number, list or tensor)."""
    if isinstance(lam, torch.Tensor):
        return tuple(float(x) for x in lam.detach().cpu().reshape(-1).tolist())
    if isinstance(lam, (int, float)):
        return (float(lam),)
    return tuple(float(x) for x in lam)


class Solution:
    """Objective of one problem (tedeous/solution.py:19-168)."""

    def __init__(self, grid: torch.Tensor, equal_cls: _Prepared,
                 model: Union[torch.nn.Module, torch.Tensor], mode: str,
                 weak_form=None, lambda_operator=1, lambda_bound=1, tol: float = 0,
                 derivative_points: int = 2, batch_size: int = None,
                 exact_div: bool = True):
        if weak_form and mode not in ('NN', 'autograd'):
            # the reference's weak residual takes its points from the 'NN' or
            # 'autograd' branch only and fails in 'mat' mode
            # (tedeous/eval.py:204-208: grid_central is never bound)
            raise NotImplementedError("weak-form losses exist in 'NN' and 'autograd' modes "
                                      "(the reference has no 'mat' weak form)")
        if tol and mode == 'mat' and len(_operators(equal_cls.operator)) > 1:
            raise NotImplementedError("the 'mat' causal loss is served for one equation")
        if batch_size is not None and mode != 'NN':
            raise NotImplementedError("mini-batching is served in 'NN' mode only")
        if batch_size is not None and tol:
            # the reference's causal loss re-derives n_t from the batch
            # residual with the 'NN' point-type lookup and fails
            # (tedeous/solution.py:54, tedeous/losses.py:160-161)
            raise NotImplementedError("mini-batching with a causal loss is not served")
        if mode not in ('NN', 'mat', 'autograd'):
            raise NotImplementedError(f"mode {mode!r} is not served by this engine")
        if not equal_cls.bconds:
            # the reference builds its boundary matrix with max() over the
            # condition columns and fails without conditions in every mode
            # (tedeous/eval.py:75); its np.inf return for 'mat'/'autograd'
            # (tedeous/losses.py:252-255) is never reached
            raise ValueError("a problem without boundary conditions has no boundary "
                             "matrix (max() arg is an empty sequence)")
        self.grid = grid
        self.mode = mode
        self.model = model
        self.lambda_operator = lambda_operator
        self.lambda_bound = lambda_bound
        self.derivative_points = derivative_points
        self._equal_cls = equal_cls
        self._weak_form = weak_form
        self._tol = tol
        self._exact_div = exact_div
        self._build()
        self.loss = None
        self.loss_normalized = None

    def _build(self) -> None:
        """Lower the problem with the current lambdas and create its plan."""
        eq = self._equal_cls
        ops = _operators(eq.operator)
        device = torch.device('cuda', torch.cuda.current_device()) if torch.cuda.is_available() \
            else torch.device('cpu')
        lam_op, lam_bnd = self.lambda_operator, self.lambda_bound
        if self.mode == 'NN':
            self.arch: MLPArch = mlp_arch(self.model)
            self.problem: NNProblem = compile_nn(
                self.grid, ops, eq.bconds, n_out=self.arch.n_out, h=eq.h,
                inner_order=eq.inner_order, boundary_order=eq.boundary_order,
                lambda_operator=lam_op, lambda_bound=lam_bnd, tol=self._tol,
                weak_form=self._weak_form)
            self.engine = NNEngine([self.problem], self.arch, device)
        elif self.mode == 'autograd':
            self.arch = mlp_arch(self.model)
            self.problem: ADProblem = compile_ad(
                self.grid, ops, eq.bconds, n_out=self.arch.n_out,
                lambda_operator=lam_op, lambda_bound=lam_bnd, tol=self._tol,
                weak_form=self._weak_form)
            self.engine = ADEngine(self.problem, self.arch, device)
        else:
            if self._tol:
                # the reference's 'mat' causal loss slices the operator matrix
                # by its columns (n_t = op.shape[1], tedeous/solution.py:62-64,
                # tedeous/losses.py:160-161): one equation is one slice, so
                # w = 1 and the operator term is mean(op^2) with no
                # lambda_operator (tedeous/losses.py:162-170)
                lam_op = 1
            self.problem: MatProblem = compile_mat(
                self.grid, ops, eq.bconds, n_eq=len(ops),
                derivative_points=self.derivative_points,
                lambda_operator=lam_op, lambda_bound=lam_bnd)
            self.engine = MatEngine(self.problem, device, exact_div=self._exact_div)
        self._lam_key = (_lambda_key(self.lambda_operator), _lambda_key(self.lambda_bound))

    def _sync_lambdas(self) -> None:
        """The reference reads ``lambda_operator``/``lambda_bound`` on every
        evaluate (tedeous/solution.py:149-150), and callbacks may replace
        them between steps (tedeous/callbacks/adaptive_lambda.py:114): a
        change rebuilds the plan's tables with the new weights."""
        key = (_lambda_key(self.lambda_operator), _lambda_key(self.lambda_bound))
        if key != self._lam_key:
            self._build()

    def _model_change(self, new_model: Union[torch.nn.Module, torch.Tensor]) -> None:
        """Rebind the solution to another model (tedeous/solution.py:109-127,
        used by the reference's cache and retraining paths).  The plan's
        tables do not depend on the weights: the same plan serves a model
        of the same architecture (or a grid of the same shape); another one
        is lowered anew."""
        if self.mode in ('NN', 'autograd'):
            same = mlp_arch(new_model) == self.arch
        else:
            same = tuple(new_model.shape) == tuple(self.model.shape)
        self.model = new_model
        if not same:
            self._build()

    # -- per step ------------------------------------------------------------
    def _flat_params(self) -> torch.Tensor:
        params = list(self.model.parameters())
        return torch.cat([p.reshape(-1) for p in params])

    def evaluate(self, save_graph: bool = True) -> Tuple[torch.Tensor, torch.Tensor]:
        """``(loss, loss_normalized)``, each of shape ``[1]``."""
        self._sync_lambdas()
        if self.mode in ('NN', 'autograd'):
            pair = self.engine.loss(self._flat_params())[0]
        else:
            u = self.model
            if not u.is_contiguous():
                u = u.contiguous()
            pair = self.engine.loss(u)
        loss = pair[0:1]
        if not save_graph:
            loss = loss.detach()
        self.loss = loss
        self.loss_normalized = pair[1:2].detach()
        return self.loss, self.loss_normalized

    # -- side outputs, on demand ----------------------------------------------
    def _side(self):
        if self.mode == 'NN':
            return self._side_nn()
        if self.mode == 'autograd':
            return self._side_ad()
        return self._side_mat()

    def _side_ad(self):
        eng: ADEngine = self.engine
        with torch.no_grad():
            resid = eng.residuals(self._flat_params().detach())
        prob = self.problem
        cols = [resid[off:off + cnt] for (_, _, off, cnt, col, _) in eng.tables.resid if col < 0]
        op = torch.stack(cols, 1) if len(cols) > 1 else cols[0].reshape(-1, 1)
        ncol = len(prob.bc_keys)
        bval = torch.zeros(prob.bc_max_len, max(ncol, 1), dtype=resid.dtype, device=resid.device)
        true = torch.zeros_like(bval)
        for (gi, qi, off, cnt, col, col_pos) in eng.tables.resid:
            if col < 0:
                continue
            pos = torch.as_tensor(col_pos, device=resid.device)
            bval[pos, col] = resid[off:off + cnt]
            t = prob.groups[gi].equations[qi].target
            true[pos, col] = torch.as_tensor(t, dtype=resid.dtype, device=resid.device)
        return op, bval[:, :ncol], true[:, :ncol], list(prob.bc_keys), list(prob.bc_lengths)

    def _side_nn(self):
        eng: NNEngine = self.engine
        with torch.no_grad():
            resid = eng.residuals(self._flat_params().detach())
        prob = self.problem
        # interior equations carry no boundary column
        cols = [resid[r.offset:r.offset + r.count] for r in eng.tables.resid if r.column < 0]
        n_eq = len(cols)
        op = torch.stack(cols, 1) if n_eq > 1 else cols[0].reshape(-1, 1)
        ncol = len(prob.bc_keys)
        bval = torch.zeros(prob.bc_max_len, max(ncol, 1), dtype=resid.dtype, device=resid.device)
        true = torch.zeros_like(bval)
        for r in eng.tables.resid:
            if r.column < 0:
                continue
            pos = torch.as_tensor(r.col_pos, device=resid.device)
            bval[pos, r.column] = resid[r.offset:r.offset + r.count]
            g = prob.groups[r.group]
            t = g.equations[r.eq].target
            true[pos, r.column] = torch.as_tensor(t, dtype=resid.dtype, device=resid.device)
        return op, bval[:, :ncol], true[:, :ncol], list(prob.bc_keys), list(prob.bc_lengths)

    def _side_mat(self):
        eng: MatEngine = self.engine
        _, op, bv = eng.fields(self.model)
        prob = self.problem
        N = op.shape[1]
        # the reference stacks per-equation residuals broadcast over n_eq
        # copies (tedeous/derivative.py:306, tedeous/eval.py:183-193)
        cols = [op[q].repeat(prob.n_eq).reshape(-1, 1) for q in range(prob.n_eq)]
        opm = cols[0] if len(cols) == 1 else torch.cat(cols, 1)
        ncol = len(prob.bc_keys)
        bval = torch.zeros(prob.bc_max_len, max(ncol, 1), dtype=op.dtype, device=op.device)
        for k, (bi, col, pos) in enumerate(eng.tables.bc_entries):
            bval[pos, col] = bv[k]
        true = torch.as_tensor(prob.targets, dtype=op.dtype, device=op.device)
        return opm, bval[:, :ncol], true, list(prob.bc_keys), list(prob.bc_lengths)

    @property
    def op(self):
        return self._side()[0]

    @property
    def bval(self):
        return self._side()[1]

    @property
    def true_bval(self):
        return self._side()[2]

    @property
    def bval_keys(self):
        return list(self.problem.bc_keys)

    @property
    def bval_length(self):
        return list(self.problem.bc_lengths)
