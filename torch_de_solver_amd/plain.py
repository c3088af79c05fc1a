"""The plain path: the reference's per-step computation in stock torch ops.

This is the comparator the native engine is measured against ("plain GPU
path"), and it is what the reference itself executes: for 'NN' mode every
stencil entry of every factor evaluates the network on its own
materialised shifted grid (duplicates included, tedeous/derivative.py:47-55),
the residual and the zero-padded boundary matrix are assembled with torch
ops (tedeous/eval.py:55-87, 143-193, 283-461), the loss is
``mean(op^2) @ lambda + mean((bval - true)^2) @ lambda``
(tedeous/losses.py:84-135) and torch autograd differentiates it.  'mat' mode
differentiates with flattened ``torch.roll`` stencils and recomputes the
step with ``torch.unique`` on every application (tedeous/derivative.py:
199-323).  'autograd' mode differentiates the network by nested
``torch.autograd.grad(create_graph=True)`` calls, one chain per derivative
factor (tedeous/derivative.py:73-132).  It runs on whatever device its
tensors live on and never calls the native library.
"""

from typing import List, Optional

import numpy as np
import torch

from .compiler import ADProblem, MatProblem, NNProblem


class PlainNN:
    def __init__(self, prob: NNProblem, net: torch.nn.Module, device):
        self.net = net
        dev = torch.device(device)
        dt = prob.dtype
        self.eqs = []       # interior: list of terms
        self.bcs = []       # (column, target positions, terms)
        self.n_central = prob.n_central
        self.keys = prob.bc_keys
        self.max_len = prob.bc_max_len
        self.lam_op = None
        for g in prob.groups:
            bases = [torch.as_tensor(b, dtype=dt, device=dev) for b in g.bases]
            # one materialised grid per stencil entry, as the reference prepares
            def grid_of(slot):
                off = torch.as_tensor(g.slot_offset[slot], dtype=dt, device=dev)
                x = bases[g.slot_base[slot]].clone()
                for a in range(x.shape[1]):
                    x[:, a] = x[:, a] + off[a]
                return x
            for eq in g.equations:
                terms = []
                for t in eq.terms:
                    c = (float(t.coeff_scalar) if t.coeff_array is None else
                         torch.as_tensor(t.coeff_array, dtype=dt, device=dev).reshape(-1, 1))
                    facs = [(f.var, int(f.pow_val) if f.pow_kind == 0 else f.pow_val,
                             [(grid_of(s), w) for s, w in f.stencil]) for f in t.factors]
                    terms.append((c, facs))
                if eq.column < 0:
                    self.eqs.append((terms, eq.lam))
                else:
                    tgt = torch.as_tensor(eq.target, dtype=dt, device=dev)
                    self.bcs.append((eq.column, torch.as_tensor(eq.col_pos, device=dev), terms,
                                     tgt, eq.lam))
        self.lam_b = {}
        for col, pos, terms, tgt, lam in self.bcs:
            self.lam_b[col] = lam
        self.dtype = dt
        self.device = dev

    def _value(self, terms):
        total = None
        for c, facs in terms:
            der = 1.
            for var, p, entries in facs:
                s = 0.
                for x, w in entries:
                    s += self.net(x)[:, var].reshape(-1, 1) * w
                der = der * s ** p
            v = c * der
            total = v if total is None else total + v
        return total

    def loss(self):
        ops = [self._value(terms).reshape(-1, 1) for terms, _ in self.eqs]
        op = ops[0] if len(ops) == 1 else torch.cat(ops, 1)
        lam_op = torch.tensor([[lam for _, lam in self.eqs]], dtype=self.dtype, device=self.device)
        loss = torch.mean((op - torch.zeros_like(op)) ** 2, 0) @ lam_op.T
        if self.bcs:
            ncol = len(self.keys)
            bval = torch.zeros(self.max_len, ncol, dtype=self.dtype, device=self.device)
            true = torch.zeros_like(bval)
            for col, pos, terms, tgt, lam in self.bcs:
                v = self._value(terms).reshape(-1)
                bval = bval.index_put((pos, torch.full_like(pos, col)), v)
                true = true.index_put((pos, torch.full_like(pos, col)), tgt)
            lam_b = torch.tensor([[self.lam_b[c] for c in range(ncol)]], dtype=self.dtype,
                                 device=self.device)
            loss = loss + torch.mean((bval - true) ** 2, 0) @ lam_b.T
        return loss


class PlainMat:
    def __init__(self, prob: MatProblem, grid: torch.Tensor):
        self.prob = prob
        self.grid = grid
        self.ab = [float(x) for x in prob.alpha_back]
        self.af = [float(x) for x in prob.alpha_forw]
        dp = prob.dp
        self.back = list(range(-dp + 1, 1))
        self.forw = list(range(dp))
        self.back_idx = [-i for i in range(1, dp)]
        self.forw_idx = list(range(dp - 1))

    def _h(self, axis):
        flat = torch.vstack([self.grid[i].reshape(-1) for i in range(self.grid.shape[0])]).T.float()
        v = torch.unique(flat[:, axis])
        return abs(v[1] - v[0])

    def _d(self, u, axis):
        h = self._h(axis)
        if u.dim() == 1 or u.shape[0] == 1:
            shape = u.shape
            x = u.reshape(-1)
            pos = None
        else:
            pos = u.dim() - 1
            x = torch.transpose(u, pos, axis)
        b, f = 0, 0
        for i, (sb, sf) in enumerate(zip(self.back, self.forw)):
            b += torch.roll(x, -sb) * self.ab[i]
            f += torch.roll(x, -sf) * self.af[i]
        du = (b + f) / (2 * h)
        if pos is None:
            du[self.back_idx] = b[self.back_idx] / h
            du[self.forw_idx] = f[self.forw_idx] / h
            return du.reshape(shape)
        du[:, self.back_idx] = b[:, self.back_idx] / h
        du[:, self.forw_idx] = f[:, self.forw_idx] / h
        return torch.transpose(du, pos, axis)

    def _field(self, u, terms):
        total = None
        for t in terms:
            der = torch.zeros_like(u) + 1
            for f in t.factors:
                prod = u[f.var]
                for ax in f.axes:
                    prod = self._d(prod, ax)
                p = int(f.pow_val) if f.pow_kind == 0 else f.pow_val
                der = der * prod ** p
            c = t.coeff_scalar if t.coeff_field is None else torch.as_tensor(
                t.coeff_field, dtype=u.dtype, device=u.device)
            v = c * der
            total = v if total is None else total + v
        return total

    def loss(self, u):
        prob = self.prob
        cols = [self._field(u, eq).reshape(-1, 1) for eq in prob.equations]
        op = cols[0] if len(cols) == 1 else torch.cat(cols, 1)
        lam_op = torch.tensor([prob.lam_op], dtype=u.dtype, device=u.device)
        loss = torch.mean((op - torch.zeros_like(op)) ** 2, 0) @ lam_op.T
        if prob.bcs:
            vals = {}
            for bc in prob.bcs:
                if bc.terms is None:
                    parts = [u[bc.var].reshape(-1)[torch.as_tensor(p, device=u.device)]
                             for p in bc.positions]
                else:
                    fld = self._field(u, bc.terms)[bc.var].reshape(-1)
                    parts = [fld[torch.as_tensor(p, device=u.device)] for p in bc.positions]
                v = parts[0]
                for w in parts[1:]:
                    v = v - w
                vals.setdefault(bc.column, []).append(v)
            ncol = len(prob.bc_keys)
            mats = []
            for c in range(ncol):
                v = torch.cat(vals[c])
                mats.append(torch.nn.functional.pad(v, (0, prob.bc_max_len - v.numel())).reshape(-1, 1))
            bval = torch.hstack(mats)
            true = torch.as_tensor(prob.targets, dtype=u.dtype, device=u.device)
            lam_b = torch.tensor([prob.lam_bnd], dtype=u.dtype, device=u.device)
            loss = loss + torch.mean((bval - true) ** 2, 0) @ lam_b.T
        return loss


class PlainAD:
    """'autograd' mode as the reference computes it: every derivative factor
    of every term is its own chain of ``torch.autograd.grad`` calls on the
    whole point set (tedeous/derivative.py:73-132)."""

    def __init__(self, prob: ADProblem, net: torch.nn.Module, device):
        self.net = net
        dev = torch.device(device)
        dt = prob.dtype
        self.dtype, self.device = dt, dev
        self.eqs, self.bcs = [], []
        self.keys, self.max_len = prob.bc_keys, prob.bc_max_len
        for g in prob.groups:
            sets = [torch.as_tensor(s, dtype=dt, device=dev) for s in g.sets]
            for eq in g.equations:
                terms = []
                for t in eq.terms:
                    c = (float(t.coeff_scalar) if t.coeff_array is None else
                         torch.as_tensor(t.coeff_array, dtype=dt, device=dev).reshape(-1, 1))
                    facs = [(f.var, int(f.pow_val) if f.pow_kind == 0 else f.pow_val,
                             [(k, list(jet), w) for k, jet, w in f.stencil]) for f in t.factors]
                    terms.append((c, facs))
                if eq.column < 0:
                    self.eqs.append((sets, terms, eq.lam))
                else:
                    self.bcs.append((sets, terms, eq.column,
                                     torch.as_tensor(eq.col_pos, device=dev),
                                     torch.as_tensor(eq.target, dtype=dt, device=dev), eq.lam))

    def _derivative(self, pts, var, axes):
        if not axes:
            return self.net(pts)[:, var].reshape(-1, 1)
        pts = pts.clone().requires_grad_(True)
        fi = self.net(pts)[:, var].sum(0)
        for ax in axes:
            grads, = torch.autograd.grad(fi, pts, create_graph=True)
            fi = grads[:, ax].sum()
        return grads[:, axes[-1]].reshape(-1, 1)

    def _value(self, sets, terms):
        total = None
        for c, facs in terms:
            der = 1.
            for var, p, entries in facs:
                s = 0.
                for k, axes, w in entries:
                    s = s + self._derivative(sets[k], var, axes) * w
                der = der * s ** p
            v = c * der
            total = v if total is None else total + v
        return total

    def residuals(self):
        """Operator residuals ``[N, n_eq]``, padded boundary values and
        targets ``[max_len, n_types]`` and the boundary lambdas."""
        ops = [self._value(sets, terms).reshape(-1, 1) for sets, terms, _ in self.eqs]
        op = ops[0] if len(ops) == 1 else torch.cat(ops, 1)
        ncol = len(self.keys)
        bval = torch.zeros(self.max_len, max(ncol, 1), dtype=self.dtype, device=self.device)
        true = torch.zeros_like(bval)
        lam_b = [0.0] * max(ncol, 1)
        for sets, terms, col, pos, tgt, lam in self.bcs:
            v = self._value(sets, terms).reshape(-1)
            bval = bval.index_put((pos, torch.full_like(pos, col)), v)
            true = true.index_put((pos, torch.full_like(pos, col)), tgt)
            lam_b[col] = lam
        return op, bval[:, :ncol], true[:, :ncol], lam_b[:ncol]

    def loss(self):
        op, bval, true, lam_b = self.residuals()
        lam_op = torch.tensor([[lam for _, _, lam in self.eqs]], dtype=self.dtype,
                              device=self.device)
        loss = torch.mean(op ** 2, 0) @ lam_op.T
        if self.bcs:
            lb = torch.tensor([lam_b], dtype=self.dtype, device=self.device)
            loss = loss + torch.mean((bval - true) ** 2, 0) @ lb.T
        return loss


class PlainNGD:
    """Natural-gradient step as the reference computes it
    (tedeous/optimizers/ngd.py:29-55, 57-78, 104-126, 144-209) in stock
    torch ops on the device: one ``torch.autograd.grad`` per residual entry
    for the Gram matrices, the SVD pseudo-inverse, a loss evaluation per
    candidate step."""

    def __init__(self, plain: PlainAD, grid_steps_number: int = 20):
        self.plain = plain
        self.params = list(plain.net.parameters())
        self.steps = 0.5 ** torch.linspace(0, grid_steps_number, grid_steps_number + 1)

    def _gram(self, res):
        rows = []
        for r in res.reshape(-1):
            g = torch.autograd.grad(r, self.params, retain_graph=True, allow_unused=True)
            g = [torch.zeros_like(p) if x is None else x for p, x in zip(self.params, g)]
            rows.append(torch.cat([x.reshape(-1) for x in g]))
        J = torch.stack(rows)
        return 1.0 / len(rows) * J.T @ J

    @staticmethod
    def _lstsq(A, B):
        tol = torch.finfo(A.dtype).eps
        U, S, Vh = torch.linalg.svd(A, full_matrices=False)
        Spinv = torch.where(S > tol, 1 / S, torch.zeros_like(S))
        return Vh.adjoint() @ (Spinv * (U.adjoint() @ B))

    def step(self):
        from torch.nn.utils import parameters_to_vector, vector_to_parameters
        op, bval, true, lam_b = self.plain.residuals()
        loss = self.plain.loss()
        grads = torch.autograd.grad(loss, self.params, retain_graph=True, allow_unused=True)
        f_grads = torch.cat([(torch.zeros_like(p) if g is None else g).reshape(-1)
                             for p, g in zip(self.params, grads)])
        G = self._gram(op) + self._gram(bval - true)
        G = torch.minimum(loss.detach().reshape(()), torch.zeros_like(G[0, 0])) * \
            torch.eye(len(G), dtype=G.dtype, device=G.device) + G
        f = self._lstsq(G, f_grads)
        base = parameters_to_vector(self.params).detach().clone()
        losses = []
        for st in self.steps.tolist():
            vector_to_parameters(base - st * f, self.params)
            losses.append(self.plain.loss().detach().reshape(1))
        k = torch.argmin(torch.cat(losses))
        vector_to_parameters(base - self.steps.to(base.device, base.dtype)[k] * f, self.params)
        return loss
