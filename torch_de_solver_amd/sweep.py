"""Batched solves: many independent 'NN' problems trained in one plan.

The reference solves one problem per ``Model`` (tedeous/model.py:22-195);
a parameter sweep is a Python loop over models.  ``SweepSolver`` lowers all
problems into one grouped plan (one table set, one work list over every
(instance, tile)), keeps all networks' parameters in one flat buffer, and
trains them with the fused native step captured in a HIP graph: one graph
replay advances every instance by one Adam step.
"""

from typing import List, Optional, Sequence, Tuple

import torch

from .compiler import compile_nn
from .nn_engine import MLPArch, NNEngine, mlp_arch


class SweepSolver:
    def __init__(self, problems: Sequence[Tuple], device=None, tile_rows: int = 0,
                 chunk_tiles: int = 0, autotune: bool = False, persistent: int = 0):
        """``problems``: tuples ``(domain, equation, conditions, net, compile_kwargs)``
        as returned by the builders of ``workloads``; every net must have the
        same architecture.  ``autotune``: pick ``chunk_tiles`` by timing
        candidates (``autotune.tune_nn``); the choice is in ``self.tuning``.
        ``persistent``: ``tdx_nn_desc.persistent`` — 0 fuses forward and
        adjoint into one kernel where the network allows it, -1 keeps them
        separate, 1 runs the separate forward as a persistent grid."""
        self.device = torch.device(device or 'cuda')
        self.nets = [p[3] for p in problems]
        self.arch: MLPArch = mlp_arch(self.nets[0])
        compiled = []
        for dom, eq, conds, net, kw in problems:
            if mlp_arch(net) != self.arch:
                raise ValueError("all networks of a sweep need the same architecture")
            if kw.get('mode', 'NN') != 'NN':
                raise ValueError("a sweep holds 'NN' problems")
            grid = dom.build('NN')
            compiled.append(compile_nn(grid, eq.equation_lst, conds.build(dom.variable_dict),
                                       self.arch.n_out, h=kw.get('h', 0.001),
                                       inner_order=kw.get('inner_order', '1'),
                                       boundary_order=kw.get('boundary_order', '2'),
                                       lambda_operator=kw.get('lambda_operator', 1),
                                       lambda_bound=kw.get('lambda_bound', 1)))
        self.problems = compiled
        P = self.arch.n_params
        dtype = compiled[0].dtype
        self.params = torch.empty(len(problems) * P, dtype=dtype, device=self.device)
        for i, net in enumerate(self.nets):
            net.to(device=self.device, dtype=dtype)
            off = i * P
            for p in net.parameters():
                n = p.numel()
                self.params[off:off + n].copy_(p.detach().reshape(-1))
                p.data = self.params[off:off + n].view_as(p)
                off += n
        self.tuning = None
        if autotune and chunk_tiles == 0:
            from .autotune import tune_nn
            chunk_tiles, self.tuning = tune_nn(compiled, self.arch, self.device, self.params,
                                               tile_rows=tile_rows, persistent=persistent)
        self.engine = NNEngine(compiled, self.arch, self.device, tile_rows, chunk_tiles,
                               persistent)
        self.exp_avg = torch.zeros_like(self.params)
        self.exp_avg_sq = torch.zeros_like(self.params)
        self.step_count = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.loss_buf = torch.zeros(2 * len(problems), dtype=dtype, device=self.device)
        self._graph = None
        self._hyper = None

    @property
    def n_inst(self) -> int:
        return len(self.nets)

    def collocation_points(self) -> int:
        return sum(p.n_central for p in self.problems)

    def losses(self) -> torch.Tensor:
        """``[n_inst, 2]``: loss and normalized loss of every instance."""
        out = torch.empty_like(self.loss_buf)
        self.engine.forward(self.params, out)
        return out.view(-1, 2)

    def step(self, lr=1e-3, betas=(0.9, 0.999), eps=1e-8) -> None:
        """One fused Adam step of every instance (eager launch)."""
        self.engine.train_step(self.params, self.exp_avg, self.exp_avg_sq, self.step_count,
                               lr, betas[0], betas[1], eps, self.loss_buf)

    def capture(self, lr=1e-3, betas=(0.9, 0.999), eps=1e-8) -> None:
        """Capture one fused step into a HIP graph (runs one eager step first,
        which counts as a step)."""
        self.step(lr, betas, eps)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        with torch.cuda.graph(g):
            self.step(lr, betas, eps)
        self._graph = g
        self._hyper = (lr, betas, eps)

    def replay(self, n: int = 1) -> None:
        if self._graph is None:
            raise RuntimeError("capture() first")
        for _ in range(n):
            self._graph.replay()

    def train(self, steps: int, lr=1e-3, betas=(0.9, 0.999), eps=1e-8, graph: bool = True):
        if steps <= 0:
            return
        if graph and steps > 2:
            self.capture(lr, betas, eps)
            self.replay(steps - 1)
        else:
            for _ in range(steps):
                self.step(lr, betas, eps)
