"""``Optimizer`` and ``Closure`` (tedeous/optimizers/optimizer.py:11-73,
tedeous/optimizers/closure.py:5-140), for the optimizers this engine serves.

Natural-gradient descent ('NGD') runs on the native engine (``ngd.NGD``);
the swarm and Nystrom-Newton optimizers of the reference (PSO, CSO, NNCG)
are not provided.
"""

from typing import Union

import torch
from torch.optim.lr_scheduler import CosineAnnealingWarmRestarts, ExponentialLR

_TORCH = {
    'Adam': torch.optim.Adam,
    'AdamW': torch.optim.AdamW,
    'SGD': torch.optim.SGD,
    'LBFGS': torch.optim.LBFGS,
    'RMSprop': torch.optim.RMSprop,
}


class Optimizer:
    def __init__(self, optimizer: Union[str, dict], params: dict,
                 gamma: Union[float, None] = None, decay_every: Union[int, None] = None,
                 cosine_scheduler_patience: Union[float, None] = None):
        self.optimizer = optimizer
        self.params = params
        self.gamma = gamma
        self.decay_every = decay_every
        self.cosine_scheduler_patience = cosine_scheduler_patience
        self.scheduler = None

    def optimizer_choice(self, mode: str, model):
        """torch optimizer over the network parameters ('NN') or the grid
        tensor ('mat') (tedeous/optimizers/optimizer.py:25-73)."""
        if self.optimizer == 'NGD':
            if mode not in ('NN', 'autograd'):
                raise NotImplementedError("NGD is served for 'NN' and 'autograd' solutions")
            from .ngd import NGD
            cls = NGD
        elif self.optimizer not in _TORCH:
            raise NotImplementedError(f"optimizer {self.optimizer!r} is not provided")
        else:
            cls = _TORCH[self.optimizer]
        if mode == 'mat':
            opt = cls([model.requires_grad_()], **self.params)
        else:
            opt = cls(model.parameters(), **self.params)
        if self.gamma is not None:
            self.scheduler = ExponentialLR(opt, gamma=self.gamma)
        if self.cosine_scheduler_patience is not None:
            self.scheduler = CosineAnnealingWarmRestarts(opt, self.cosine_scheduler_patience)
        return opt

    def fusable_adam(self):
        """(lr, beta1, beta2, eps) when the settings are exactly what the fused
        native step implements (plain Adam, no decay/amsgrad/schedule)."""
        if self.optimizer != 'Adam' or self.gamma is not None or \
                self.cosine_scheduler_patience is not None:
            return None
        p = dict(self.params)
        lr = p.pop('lr', 1e-3)
        betas = p.pop('betas', (0.9, 0.999))
        eps = p.pop('eps', 1e-8)
        if p.pop('weight_decay', 0) != 0 or p.pop('amsgrad', False) or \
                p.pop('maximize', False):
            return None
        for k in ('foreach', 'fused', 'capturable', 'differentiable'):
            p.pop(k, None)
        if p:
            return None
        return float(lr), float(betas[0]), float(betas[1]), float(eps)


class Closure:
    """zero_grad -> evaluate -> backward (tedeous/optimizers/closure.py:49-64).

    ``mixed_precision=True`` keeps the reference's control flow on a GPU
    (tedeous/optimizers/closure.py:17-64): the evaluation under
    ``torch.autocast`` and a ``GradScaler`` that scales the loss, steps the
    optimizer inside the closure and updates its scale -- so, as in the
    reference, the optimizer's own ``step(closure)`` then applies a second
    update with the same gradient.  The kernels stay in FP32: autocast
    would lower the reference's network products to FP16, and nothing here
    is an autocast-eligible torch op.
    """

    def __init__(self, mixed_precision: bool, model):
        self.model = model
        self.optimizer = model.optimizer
        self.normalized_loss_stop = model.normalized_loss_stop
        self.mixed_precision = bool(mixed_precision)
        self.scaler = None
        if self.mixed_precision:
            if self.optimizer.__class__.__name__ == "LBFGS":
                raise NotImplementedError("AMP and the LBFGS optimizer are not compatible.")
            self.scaler = torch.amp.GradScaler('cuda', enabled=True)

    def _closure(self):
        self.optimizer.zero_grad()
        with torch.autocast(device_type='cuda', dtype=torch.float16,
                            enabled=self.mixed_precision):
            loss, loss_normalized = self.model.solution_cls.evaluate()
        if self.scaler is not None:
            self.scaler.scale(loss).backward()
            self.scaler.step(self.optimizer)
            self.scaler.update()
        else:
            loss.backward()
        self.model.cur_loss = loss_normalized if self.normalized_loss_stop else loss
        return loss

    def _closure_ngd(self):
        """(tedeous/optimizers/closure.py:98-117): loss and gradient, then the
        operator residuals and the boundary values, and the evaluate function
        the line search calls."""
        self.optimizer.zero_grad()
        sol = self.model.solution_cls
        loss, loss_normalized = sol.evaluate()
        loss.backward(retain_graph=True)
        self.model.cur_loss = loss_normalized if self.normalized_loss_stop else loss
        op, bval, true_bval, _, _ = sol._side()
        return op, bval, true_bval, loss, sol.evaluate

    def get_closure(self, _type: str):
        if _type in ('PSO', 'CSO', 'NNCG'):
            raise NotImplementedError(f"optimizer {_type} is not provided")
        if _type == 'NGD':
            return self._closure_ngd
        return self._closure
