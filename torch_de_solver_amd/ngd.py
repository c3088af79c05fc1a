"""Natural-gradient descent (NGD) over the native 'NN' engine.

The reference's optimizer (tedeous/optimizers/ngd.py:8-178, with the closure
of tedeous/optimizers/closure.py:98-117) takes one step as:

1. loss and gradient of the lambda-weighted loss;
2. the Gram matrices ``1/N J^T J`` of the operator residuals (flattened
   ``[N, n_eq]``) and of the padded boundary residuals ``bval - true_bval``
   (flattened ``[max_len, n_types]``), J built with one
   ``torch.autograd.grad`` per residual entry (ngd.py:57-78);
3. the Marquardt term ``min(loss, 0) * I`` (zero for a non-negative loss);
4. the natural gradient ``f`` as the pseudo-inverse solution of
   ``G f = grad`` (singular values above the dtype's machine epsilon
   inverted, ngd.py:104-126);
5. a grid line search: the loss at ``params - 0.5**k * f`` for
   ``k = 0..grid_steps_number``, the first minimum taken (ngd.py:29-55).

Here J comes from one kernel (``tdx_nn_jacobian`` in 'NN' mode: every
residual's seed-weighted network adjoint, one thread per entry;
``tdx_ad_jacobian`` in 'autograd' mode: the same over the network's jets),
the Gram products are two library GEMMs on the device, the pseudo-inverse
comes from the device's symmetric eigensolver (G is symmetric; see
``lstsq``), and the line search evaluates all candidate parameter vectors
in one launch of a grouped plan that holds the problem once per candidate
('NN'), or one forward per candidate ('autograd').  Every value stays on
the device; the eigensolver's convergence check is the step's one wait for
the device.
"""

from typing import Optional

import torch

from .nn_engine import NNEngine


class NGD(torch.optim.Optimizer):
    """Drop-in for ``tedeous.optimizers.ngd.NGD`` on 'NN' and 'autograd'
    solutions."""

    def __init__(self, params, grid_steps_number: int = 30):
        defaults = {'grid_steps_number': grid_steps_number}
        super().__init__(params, defaults)
        self.params = self.param_groups[0]['params']
        self.grid_steps_number = grid_steps_number
        self.grid_steps = torch.linspace(0, grid_steps_number, grid_steps_number + 1)
        self.steps = 0.5 ** self.grid_steps
        self._cache = None      # (engine, grouped line-search engine, index sets)
        self.last: dict = {}    # intermediate values of the last step

    # -- pieces ----------------------------------------------------------------
    @staticmethod
    def lstsq(A: torch.Tensor, B: torch.Tensor, tol: Optional[float] = None) -> torch.Tensor:
        """The reference's pseudo-inverse solution (singular values above
        ``tol``, default the dtype's machine epsilon, inverted;
        tedeous/optimizers/ngd.py:104-126) for the symmetric Gram matrix A.
        For a symmetric A the singular values are the eigenvalues' magnitudes
        and the pseudo-inverse is sum over |lambda| > tol of v v^T / lambda,
        so the device's symmetric eigensolver gives the same solution; it
        is used because the device SVD loses accuracy on ill-conditioned
        Gram matrices (measured: 0.37 relative error against LAPACK at
        condition 1e9, against 8e-8 for the eigensolver)."""
        tol = torch.finfo(A.dtype).eps if tol is None else tol
        w, V = torch.linalg.eigh(A)
        winv = torch.where(w.abs() > tol, 1 / w, torch.zeros_like(w))
        VhB = V.adjoint() @ B
        if winv.ndim != VhB.ndim:
            winv = winv.unsqueeze(-1)
        return V @ (winv * VhB)

    def _prepare(self, sol):
        """Engine, line-search engine (None: the engine itself, once per
        candidate), operator and boundary residual indices, boundary
        padding length."""
        eng = sol.engine
        if self._cache is not None and self._cache[0] is eng:
            return self._cache
        if eng.n_inst != 1:
            raise NotImplementedError("NGD serves one instance per solution")
        op_idx, b_idx = [], []
        if sol.mode == 'NN':
            spans = [(r.offset, r.count, r.column) for r in eng.tables.resid]
        else:
            spans = [(off, cnt, col) for (_, _, off, cnt, col, _) in eng.tables.resid]
        for off, cnt, col in spans:
            (op_idx if col < 0 else b_idx).extend(range(off, off + cnt))
        dev = eng.device
        op_idx = torch.tensor(op_idx, dtype=torch.long, device=dev)
        b_idx = torch.tensor(b_idx, dtype=torch.long, device=dev)
        prob = sol.problem
        n_b = prob.bc_max_len * len(prob.bc_keys)
        search = NNEngine([prob] * len(self.steps), sol.arch, dev) if sol.mode == 'NN' else None
        self._cache = (eng, search, op_idx, b_idx, max(n_b, 1))
        return self._cache

    def _line_search(self, eng, search, cand: torch.Tensor, losses: torch.Tensor) -> None:
        """Loss pairs of every candidate parameter vector (rows of cand)."""
        if search is not None:
            search.forward(cand.view(-1), losses)
            return
        for k in range(cand.shape[0]):
            eng.forward(cand[k], losses[2 * k:2 * k + 2])

    def gram(self, jt: torch.Tensor, op_idx: torch.Tensor, b_idx: torch.Tensor,
             n_b: int) -> torch.Tensor:
        """``1/N_op J_op^T J_op + 1/N_b J_b^T J_b`` from the transposed J."""
        j_op = jt.index_select(1, op_idx)
        g = (1.0 / op_idx.numel() * j_op) @ j_op.T
        if b_idx.numel():
            j_b = jt.index_select(1, b_idx)
            g = g + (1.0 / n_b * j_b) @ j_b.T
        return g

    # -- one step --------------------------------------------------------------
    def step(self, closure=None):
        if closure is None:
            raise ValueError("NGD needs the closure of Closure.get_closure('NGD')")
        int_res, bval, true_bval, loss, loss_function = closure()
        sol = getattr(loss_function, '__self__', None)
        if sol is None or getattr(sol, 'mode', None) not in ('NN', 'autograd'):
            raise NotImplementedError("NGD is served for 'NN' and 'autograd' solutions")
        eng, search, op_idx, b_idx, n_b = self._prepare(sol)
        params = self.params
        flat = torch.cat([p.detach().reshape(-1) for p in params]).contiguous()
        grads = [p.grad for p in params]
        if any(g is None for g in grads):
            grads = torch.autograd.grad(loss, params, retain_graph=True, allow_unused=True)
            grads = [torch.zeros_like(p) if g is None else g for p, g in zip(params, grads)]
        f_grads = torch.cat([g.reshape(-1) for g in grads]).to(flat.dtype)
        jt = eng.jacobian(flat)
        G = self.gram(jt, op_idx, b_idx, n_b)
        damp = torch.minimum(loss.detach().reshape(()).to(G.dtype), G.new_zeros(()))
        G = damp * torch.eye(len(G), dtype=G.dtype, device=G.device) + G
        f = self.lstsq(G, f_grads)
        # every candidate's loss in one launch
        K = len(self.steps)
        steps = self.steps.to(device=flat.device, dtype=flat.dtype)
        cand = (flat.unsqueeze(0) - steps.view(K, 1) * f.unsqueeze(0)).contiguous()
        losses = torch.empty(2 * K, dtype=flat.dtype, device=flat.device)
        self._line_search(eng, search, cand, losses)
        losses = losses.view(K, 2)[:, 0]
        k = torch.argmin(losses)
        new = cand.index_select(0, k.view(1)).view(-1)
        with torch.no_grad():
            off = 0
            for p in params:
                n = p.numel()
                p.copy_(new[off:off + n].view_as(p))
                off += n
        self.last = dict(G=G, f=f, losses=losses, k=k, grad=f_grads, jt=jt)
        return loss
