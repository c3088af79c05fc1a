"""``Model``: compile and train loop (tedeous/model.py:22-195).

``compile`` lowers the problem once to native tables.  ``train`` keeps the
reference's loop semantics: one initial evaluation, then ``epochs - 1``
epochs (tedeous/model.py:154-188), callbacks around each epoch.

Mini-batching ('NN' mode, ``batch_size``): the reference draws batches of
the sorted grid with a DataLoader and takes one optimizer step per batch
(tedeous/model.py:178-183), but its 'NN' derivative ignores the batch's
points (tedeous/derivative.py:30-58 never reads ``grid_points``), so every
step evaluates the full operator.  An epoch here is therefore
``ceil(N / batch_size)`` full steps, and a ``batch_size`` above the number
of grid points turns batching off (tedeous/model.py:107-109).

When the optimizer is plain Adam with no scheduler and no callbacks, the
whole step (forward, adjoint, Adam update) runs as one native call per step
(``tdx_*_train_step``), captured once into a HIP graph and replayed.  Any
other optimizer goes through torch's optimizer and the native
``torch.autograd.Function`` of ``Solution.evaluate``.
"""

import datetime
from typing import List, Optional, Union

import torch

from .callbacks import CallbackList
from .data import Conditions, Domain, Equation
from .nn_engine import flatten_module_params, module_flat_view
from .optimizers import Closure, Optimizer
from .solution import Operator_bcond_preproc, Solution


class Model:
    def __init__(self, net: Union[torch.nn.Module, torch.Tensor], domain: Domain,
                 equation: Equation, conditions: Conditions, batch_size: int = None):
        self.net = net
        self.domain = domain
        self.equation = equation
        self.conditions = conditions
        self.batch_size = batch_size
        self.verbose = True
        self.use_graph = True

    def compile(self, mode: str, lambda_operator, lambda_bound, normalized_loss_stop: bool = False,
                h: float = 0.001, inner_order: str = '1', boundary_order: str = '2',
                derivative_points: int = 2, weak_form: List[callable] = None, tol: float = 0,
                removed_domains: list = None, exact_div: bool = True):
        self.mode = mode
        self.lambda_bound = lambda_bound
        self.lambda_operator = lambda_operator
        self.normalized_loss_stop = normalized_loss_stop
        self.weak_form = weak_form
        self.removed_domains = removed_domains
        grid = self.domain.build(mode=mode, removed_domains=removed_domains)
        dtype = grid.dtype
        if self.batch_size is not None and len(grid) < self.batch_size:
            self.batch_size = None
        # one optimizer step per DataLoader batch of the N sorted grid points
        self.n_batches = 1 if self.batch_size is None else -(-len(grid) // int(self.batch_size))
        if isinstance(self.net, torch.nn.Module):
            self.net.to(dtype)
            if torch.cuda.is_available():
                self.net.to('cuda')
        else:
            dev = 'cuda' if torch.cuda.is_available() else self.net.device
            self.net = self.net.to(dtype=dtype, device=dev).contiguous()
        bconds = self.conditions.build(self.domain.variable_dict)
        self.equation_cls = Operator_bcond_preproc(
            grid, self.equation.equation_lst, bconds, h=h, inner_order=inner_order,
            boundary_order=boundary_order).set_strategy(mode)
        self.solution_cls = Solution(grid, self.equation_cls, self.net, mode, weak_form,
                                     lambda_operator, lambda_bound, tol, derivative_points,
                                     batch_size=self.batch_size, exact_div=exact_div)

    # ------------------------------------------------------------------------
    def train(self, optimizer: Optimizer, epochs: int, info_string_every: Union[int, None] = None,
              mixed_precision: bool = False, save_model: bool = False,
              model_name: Union[str, None] = None, callbacks: Union[List, None] = None):
        epochs = int(epochs)
        self.t = 1
        self.stop_training = False
        if save_model:
            raise NotImplementedError("model caching is not provided by this engine")
        hyper = optimizer.fusable_adam()
        if hyper is not None and not callbacks and not mixed_precision:
            return self._train_fused(hyper, epochs)
        cbs = CallbackList(callbacks=callbacks, model=self)
        cbs.on_train_begin()
        self.net = self.solution_cls.model
        self.optimizer = optimizer.optimizer_choice(self.mode, self.net)
        closure = Closure(mixed_precision, self).get_closure(optimizer.optimizer)
        self.min_loss, _ = self.solution_cls.evaluate()
        self.cur_loss = self.min_loss
        if self.verbose:
            print('[{}] initial (min) loss is {}'.format(datetime.datetime.now(),
                                                         self.min_loss.item()))
        while self.t < epochs and self.stop_training is False:
            cbs.on_epoch_begin()
            self.optimizer.zero_grad()
            for _ in range(self.n_batches):
                self.optimizer.step(closure)
                if optimizer.gamma is not None and self.t % optimizer.decay_every == 0:
                    optimizer.scheduler.step()
            cbs.on_epoch_end()
            self.t += 1
        cbs.on_train_end()

    def _flat_state(self) -> torch.Tensor:
        if self.mode in ('NN', 'autograd'):
            flat = module_flat_view(self.net)
            if flat is None:
                flat = flatten_module_params(self.net)
            return flat
        return self.net.data

    def _train_fused(self, hyper, epochs: int):
        """Fused native Adam steps; same updates as torch.optim.Adam."""
        lr, b1, b2, eps = hyper
        sol = self.solution_cls
        eng = sol.engine
        params = self._flat_state()
        m = torch.zeros_like(params)
        v = torch.zeros_like(params)
        step = torch.zeros(1, dtype=torch.int64, device=params.device)
        loss_buf = torch.zeros(2 * getattr(eng, 'n_inst', 1), dtype=params.dtype,
                               device=params.device)
        self.min_loss, _ = sol.evaluate(save_graph=False)
        self.cur_loss = self.min_loss
        if self.verbose:
            print('[{}] initial (min) loss is {}'.format(datetime.datetime.now(),
                                                         self.min_loss.item()))
        n_steps = max(epochs - 1, 0) * self.n_batches
        self.loss_history_device = torch.zeros(0)

        # 'mat': the step reads one grid buffer and writes the other (the
        # kernel generated for the interior then applies Adam as it forms
        # each gradient); the buffers alternate and the result is copied
        # back into the model's tensor at the end
        pingpong = self.mode == 'mat' and hasattr(eng, 'train_step_to')
        bufs = [params, torch.empty_like(params)] if pingpong else [params]
        done = [0]

        def one_step(k=None):
            k = done[0] if k is None else k
            if pingpong:
                eng.train_step_to(bufs[k % 2], bufs[(k + 1) % 2], m, v, step, lr, b1, b2, eps,
                                  loss_buf)
            else:
                eng.train_step(params, m, v, step, lr, b1, b2, eps, loss_buf)

        remaining = n_steps
        if remaining > 0:
            one_step()  # first call sets kernel attributes outside any capture
            done[0] += 1
            remaining -= 1
        graph = None
        if self.use_graph and remaining > 4:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                one_step()
                done[0] += 1
            torch.cuda.current_stream().wait_stream(side)
            remaining -= 1
            graphs = [torch.cuda.CUDAGraph() for _ in bufs]
            for par, g in enumerate(graphs):
                with torch.cuda.graph(g):
                    one_step(par)  # recorded, not run
            for _ in range(remaining):
                graphs[done[0] % len(graphs)].replay()
                done[0] += 1
            graph = graphs[0]
        else:
            for _ in range(remaining):
                one_step()
                done[0] += 1
        if pingpong and done[0] % 2 == 1:
            params.copy_(bufs[1])
        self.t = epochs
        self.steps_done = int(step.item())
        self.cur_loss = loss_buf[0:1].clone()
        self.graph = graph
