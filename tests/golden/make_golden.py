"""Write golden vectors of the reference's end-to-end objective.

Runs ONLY in the build container, where the reference checkout is present
at /root/reference; it is never run by the tests (they read the .npz files
written here).  For each small problem of tests/problems.py it builds the
problem with the reference's own classes, evaluates
``Solution.evaluate()`` (tedeous/solution.py:129-168) and the gradient of
its loss, and records a few Adam steps of the reference's closure loop
(tedeous/optimizers/closure.py:49-64).  What is recorded: inputs (network
parameters or the initial grid tensor) and outputs (loss, normalized loss,
residual matrix, padded boundary matrices, column keys and lengths,
parameter gradients, losses along the Adam steps).

Usage: python tests/golden/make_golden.py
"""

import os
import sys

sys.dont_write_bytecode = True
HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))            # tests/
sys.path.insert(0, '/root/reference')

import numpy as np  # noqa: E402
import torch  # noqa: E402

import problems  # noqa: E402
from tedeous import data as ref_data  # noqa: E402
from tedeous.input_preprocessing import Operator_bcond_preproc  # noqa: E402
from tedeous.solution import Solution  # noqa: E402

ADAM_STEPS = 5
LR = 1e-3


def solution_for(builder, dtype, model):
    domain, eq, conds, kw = builder(ref_data, dtype=dtype)
    mode = kw['mode']
    grid = domain.build(mode)
    bconds = conds.build(domain.variable_dict)
    eqc = Operator_bcond_preproc(grid, eq.equation_lst, bconds, h=kw.get('h', 0.001),
                                 inner_order='1', boundary_order='2').set_strategy(mode)
    return Solution(grid, eqc, model, mode, None, kw['lambda_operator'], kw['lambda_bound'],
                    tol=0, derivative_points=kw.get('derivative_points', 2))


def record(sol, params):
    loss, norm = sol.evaluate()
    for p in params:
        p.grad = None
    loss.backward()
    out = dict(
        loss=loss.detach().numpy().reshape(-1), loss_norm=norm.detach().numpy().reshape(-1),
        op=sol.op.detach().numpy(), bval=sol.bval.detach().numpy(),
        true_bval=sol.true_bval.detach().numpy(),
        keys=np.array(sol.bval_keys), lens=np.array(sol.bval_length),
        grad=torch.cat([p.grad.reshape(-1) for p in params]).numpy())
    return out


def train(sol, params):
    opt = torch.optim.Adam(params, lr=LR)
    losses = []
    for _ in range(ADAM_STEPS):
        opt.zero_grad()
        loss, _ = sol.evaluate()
        loss.backward()
        losses.append(float(loss))
        opt.step()
    return np.array(losses)


def main():
    torch.set_default_device('cpu')
    for name, (builder, layers) in problems.NN_CASES.items():
        for dtype in ('float64', 'float32'):
            tdt = torch.float64 if dtype == 'float64' else torch.float32
            net = problems.randomize_biases(problems.make_mlp(layers, seed=0, dtype=tdt))
            params0 = torch.cat([p.detach().reshape(-1) for p in net.parameters()]).numpy().copy()
            sol = solution_for(builder, dtype, net)
            out = record(sol, list(net.parameters()))
            out['losses'] = train(sol, list(net.parameters()))
            out['params'] = params0
            path = os.path.join(HERE, f"{name}_{dtype}.npz")
            np.savez_compressed(path, **out)
            print(path, float(out['loss'][0]))
    for name, builder in problems.MAT_CASES.items():
        for dtype in ('float64', 'float32'):
            tdt = torch.float64 if dtype == 'float64' else torch.float32
            domain, eq, conds, kw = builder(ref_data, dtype=dtype)
            shape = [len(eq.equation_lst)] + list(domain.build('mat').shape[1:])
            u0 = problems.mat_init(shape, seed=0, dtype=tdt)
            model = u0.clone().requires_grad_()
            sol = solution_for(builder, dtype, model)
            out = record(sol, [model])
            m2 = u0.clone().requires_grad_()
            sol2 = solution_for(builder, dtype, m2)
            out['losses'] = train(sol2, [m2])
            out['params'] = u0.numpy().reshape(-1)
            path = os.path.join(HERE, f"{name}_{dtype}.npz")
            np.savez_compressed(path, **out)
            print(path, float(out['loss'][0]))


if __name__ == '__main__':
    main()
