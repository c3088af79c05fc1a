"""GPU parity of the native engine against the reference's golden vectors and
the CPU oracle (tests/golden, oracle/).

Tolerances:
* float64 kernels: loss, residuals and gradients to 1e-9 relative (the
  same mathematics in a different summation order).
* 'mat' float32 with exact division: residual field and boundary values
  bit-identical to the reference's; loss to 1e-6 and gradient to 1e-5
  relative (reduction order).
* 'NN' float32: the network is evaluated by MFMA with a different
  accumulation order than torch's GEMM, and the stencil divides those
  last-bit differences by (2h)^k.  The residual is compared at the tolerance
  that bound gives (see test docstrings); the loss of a smooth network to
  1e-5 relative when h = 1e-2.
"""
import math
import sys

import numpy as np
import pytest
import torch

import problems
from conftest import gpu_available
from oracle import tedeous_oracle as O

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

import torch_de_solver_amd as tdx  # noqa: E402
from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def nn_solution(name, dtype, params=None):
    builder, layers = problems.NN_CASES[name]
    d, eq, b, kw = builder(api, dtype=dtype)
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    net = problems.make_mlp(layers, dtype=tdt)
    if params is not None:
        torch.nn.utils.vector_to_parameters(torch.as_tensor(params, dtype=tdt), net.parameters())
    net = net.to(DEV)
    grid = d.build('NN')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict),
                                 h=kw.get('h', 0.001)).set_strategy('NN')
    sol = Solution(grid, pre, net, 'NN', None, kw['lambda_operator'], kw['lambda_bound'])
    return sol, net, kw


def mat_solution(name, dtype, u0, exact_div=True):
    d, eq, b, kw = problems.MAT_CASES[name](api, dtype=dtype)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    u = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, u, 'mat', None, kw['lambda_operator'], kw['lambda_bound'],
                   derivative_points=kw.get('derivative_points', 2), exact_div=exact_div)
    return sol, u, kw


# ---------------------------------------------------------------------------
# 'NN' mode
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('name', list(problems.NN_CASES))
def test_nn_float64_matches_reference(name, golden):
    _print(f"[nn f64] {name}")
    g = golden(f'{name}_float64')
    sol, net, _ = nn_solution(name, 'float64', g['params'])
    loss, norm = sol.evaluate()
    np.testing.assert_allclose(loss.detach().cpu().numpy(), g['loss'], rtol=1e-9)
    np.testing.assert_allclose(norm.cpu().numpy(), g['loss_norm'], rtol=1e-9)
    np.testing.assert_allclose(sol.op.cpu().numpy(), g['op'], rtol=1e-8,
                               atol=1e-9 * np.abs(g['op']).max())
    np.testing.assert_allclose(sol.bval.cpu().numpy(), g['bval'], rtol=1e-9, atol=1e-12)
    np.testing.assert_array_equal(sol.true_bval.cpu().numpy(), g['true_bval'])
    assert sol.bval_keys == list(g['keys']) and sol.bval_length == list(g['lens'])
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    np.testing.assert_allclose(grad, g['grad'], rtol=1e-8, atol=1e-10 * np.abs(g['grad']).max())


@pytest.mark.parametrize('name', list(problems.NN_CASES))
def test_nn_float32_against_oracle(name, golden):
    """FP32: the network output u agrees with torch's to ~1e-6 relative; the
    stencil of order k divides the last-bit differences by (2h)^k, so the
    residual is compared with atol = 8*eps*max|u|*sum|w| (w: stencil
    weights) and the loss/gradients at the relative bound that follows."""
    _print(f"[nn f32] {name}")
    g = golden(f'{name}_float32')
    sol, net, kw = nn_solution(name, 'float32', g['params'])
    loss, norm = sol.evaluate()
    loss.backward()
    grad = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu().numpy()
    h = kw.get('h', 0.001)
    # stencil amplification: second derivative 1/(4h^2) * 4 terms
    amp = {'legendre_nn': 1 / h ** 2, 'legendre_nn_opbc': 1 / h ** 2,
           'burgers_nn': 1 / h ** 2, 'kdv_nn': 1 / h ** 3}[name]
    eps = np.finfo(np.float32).eps
    umax = 2.0
    atol = 8 * eps * umax * amp
    op = sol.op.cpu().numpy()
    assert np.max(np.abs(op - g['op'])) <= atol, (np.max(np.abs(op - g['op'])), atol)
    rel = abs(float(loss) - float(g['loss'][0])) / float(g['loss'][0])
    assert rel < 2e-3, rel
    gn = np.linalg.norm(grad - g['grad']) / np.linalg.norm(g['grad'])
    assert gn < 2e-2, gn


def test_nn_float32_loss_1e5_at_h_1e2(golden):
    """With h = 1e-2 (the KdV setting) on a smooth 2-D problem the float32
    loss and gradient match the reference's float32 numbers to 1e-5 / 1e-4."""
    name = 'burgers_nn'
    g64 = golden(f'{name}_float64')
    builder, layers = problems.NN_CASES[name]
    out = {}
    for dtype in ('float64', 'float32'):
        d, eq, b, kw = builder(api, dtype=dtype)
        tdt = torch.float64 if dtype == 'float64' else torch.float32
        net = problems.make_mlp(layers, dtype=tdt)
        torch.nn.utils.vector_to_parameters(torch.as_tensor(g64['params'], dtype=tdt), net.parameters())
        orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst], h=1e-2)
        lo = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])[0]
        lo.backward()
        go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
        net2 = problems.make_mlp(layers, dtype=tdt)
        torch.nn.utils.vector_to_parameters(torch.as_tensor(g64['params'], dtype=tdt), net2.parameters())
        net2 = net2.to(DEV)
        grid = d.build('NN')
        pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict), h=1e-2).set_strategy('NN')
        sol = Solution(grid, pre, net2, 'NN', None, kw['lambda_operator'], kw['lambda_bound'])
        ln, _ = sol.evaluate()
        ln.backward()
        gn = torch.cat([p.grad.reshape(-1) for p in net2.parameters()]).cpu().numpy()
        out[dtype] = (float(lo), go, float(ln), gn)
    lo, go, ln, gn = out['float32']
    assert abs(ln - lo) / lo < 1e-5, (ln, lo)
    assert np.linalg.norm(gn - go) / np.linalg.norm(go) < 1e-4
    lo64, go64, ln64, gn64 = out['float64']
    assert abs(ln64 - lo64) / lo64 < 1e-10


def test_nn_gradient_central_difference():
    """float64 native gradient against central differences of the native loss."""
    name = 'kdv_nn'
    sol, net, _ = nn_solution(name, 'float64')
    problems.randomize_biases(net)
    loss, _ = sol.evaluate()
    loss.backward()
    params = list(net.parameters())
    grad = torch.cat([p.grad.reshape(-1) for p in params])
    flat = torch.cat([p.detach().reshape(-1) for p in params])
    rng = np.random.default_rng(0)
    idx = rng.choice(flat.numel(), 12, replace=False)
    eps = 1e-6
    for i in idx:
        for sgn, store in ((1, 'p'), (-1, 'm')):
            f = flat.clone()
            f[i] += sgn * eps
            torch.nn.utils.vector_to_parameters(f, params)
            val = float(sol.evaluate()[0])
            if store == 'p':
                lp = val
            else:
                lm = val
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - float(grad[i])) <= 1e-5 * max(1.0, abs(fd)), (i, fd, float(grad[i]))
    torch.nn.utils.vector_to_parameters(flat, params)


def test_nn_adam_steps_match_reference(golden):
    """Fused native steps (forward + adjoint + Adam in one call) reproduce the
    reference's losses along 5 Adam steps (float64)."""
    for name in ('burgers_nn', 'legendre_nn'):
        _print(f"[nn adam] {name}")
        g = golden(f'{name}_float64')
        sol, net, kw = nn_solution(name, 'float64', g['params'])
        from torch_de_solver_amd.nn_engine import flatten_module_params
        flat = flatten_module_params(net)
        m, v = torch.zeros_like(flat), torch.zeros_like(flat)
        step = torch.zeros(1, dtype=torch.int64, device=DEV)
        loss = torch.zeros(2, dtype=torch.float64, device=DEV)
        got = []
        for _ in range(len(g['losses'])):
            sol.engine.train_step(flat, m, v, step, 1e-3, 0.9, 0.999, 1e-8, loss)
            got.append(float(loss[0]))
        np.testing.assert_allclose(got, g['losses'], rtol=1e-9)
        assert int(step) == len(g['losses'])


def test_nn_multi_instance_equals_single():
    """A grouped plan of several instances gives each instance the loss and
    gradient of its own single-instance plan."""
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch
    probs, nets = [], []
    for i, (nx, nt, mu) in enumerate([(15, 9, 0.01), (21, 7, 0.003), (11, 12, 0.05)]):
        d, eq, b, kw = problems.burgers_nn(api, nx=nx, nt=nt, mu=mu)
        net = problems.randomize_biases(problems.make_mlp([2, 12, 12, 12, 1], seed=i), seed=i)
        probs.append(compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                                lambda_operator=1, lambda_bound=10))
        nets.append(net)
    arch = mlp_arch(nets[0])
    flat = torch.cat([torch.cat([p.detach().reshape(-1) for p in n.parameters()]) for n in nets]).to(DEV)
    grouped = NNEngine(probs, arch, DEV)
    fl = flat.clone().requires_grad_()
    lg = grouped.loss(fl)
    lg[:, 0].sum().backward()
    P = arch.n_params
    for i in range(3):
        single = NNEngine([probs[i]], arch, DEV)
        fi = flat[i * P:(i + 1) * P].clone().requires_grad_()
        ls = single.loss(fi)
        ls[0, 0].backward()
        np.testing.assert_allclose(lg[i].detach().cpu().numpy(), ls[0].detach().cpu().numpy(), rtol=1e-6)
        np.testing.assert_allclose(fl.grad[i * P:(i + 1) * P].cpu().numpy(), fi.grad.cpu().numpy(),
                                   rtol=1e-5, atol=1e-7)


def test_nn_deterministic():
    sol, net, _ = nn_solution('burgers_nn', 'float32')
    outs = []
    for _ in range(2):
        for p in net.parameters():
            p.grad = None
        loss, _ = sol.evaluate()
        loss.backward()
        outs.append((float(loss), torch.cat([p.grad.reshape(-1) for p in net.parameters()]).cpu()))
    assert outs[0][0] == outs[1][0]
    assert torch.equal(outs[0][1], outs[1][1])


# ---------------------------------------------------------------------------
# 'mat' mode
# ---------------------------------------------------------------------------

@pytest.mark.parametrize('name', list(problems.MAT_CASES))
@pytest.mark.parametrize('dtype', ['float64', 'float32'])
def test_mat_matches_reference(name, dtype, golden):
    _print(f"[mat] {name} {dtype}")
    g = golden(f'{name}_{dtype}')
    d, eq, b, kw = problems.MAT_CASES[name](api, dtype=dtype)
    tdt = torch.float64 if dtype == 'float64' else torch.float32
    shape = [len(eq.equation_lst)] + list(d.build('mat').shape[1:])
    u0 = torch.from_numpy(g['params']).reshape(shape).to(tdt)
    sol, u, kw = mat_solution(name, dtype, u0, exact_div=True)
    loss, norm = sol.evaluate()
    # residual field and boundary values: the reference's float operations
    np.testing.assert_array_equal(sol.op.cpu().numpy(), g['op'])
    np.testing.assert_array_equal(sol.bval.cpu().numpy(), g['bval'])
    np.testing.assert_array_equal(sol.true_bval.cpu().numpy(), g['true_bval'])
    rt = 1e-10 if dtype == 'float64' else 1e-6
    np.testing.assert_allclose(loss.detach().cpu().numpy(), g['loss'], rtol=rt)
    np.testing.assert_allclose(norm.cpu().numpy(), g['loss_norm'], rtol=rt)
    loss.backward()
    np.testing.assert_allclose(u.grad.cpu().numpy().reshape(-1), g['grad'],
                               rtol=1e-9 if dtype == 'float64' else 1e-5,
                               atol=(1e-12 if dtype == 'float64' else 1e-6) * np.abs(g['grad']).max())


@pytest.mark.parametrize('name', ['wave_mat', 'nonlinear_mat'])
def test_mat_fast_division_within_tolerance(name, golden):
    g = golden(f'{name}_float32')
    d, eq, b, kw = problems.MAT_CASES[name](api, dtype='float32')
    shape = [len(eq.equation_lst)] + list(d.build('mat').shape[1:])
    u0 = torch.from_numpy(g['params']).reshape(shape)
    sol, u, kw = mat_solution(name, 'float32', u0, exact_div=False)
    loss, _ = sol.evaluate()
    loss.backward()
    np.testing.assert_allclose(loss.detach().cpu().numpy(), g['loss'], rtol=1e-5)
    np.testing.assert_allclose(u.grad.cpu().numpy().reshape(-1), g['grad'], rtol=1e-4,
                               atol=1e-5 * np.abs(g['grad']).max())


def test_mat_adam_steps_match_reference(golden):
    for name in ('wave_mat', 'nonlinear_mat'):
        g = golden(f'{name}_float64')
        d, eq, b, kw = problems.MAT_CASES[name](api, dtype='float64')
        shape = [len(eq.equation_lst)] + list(d.build('mat').shape[1:])
        u0 = torch.from_numpy(g['params']).reshape(shape)
        sol, u, kw = mat_solution(name, 'float64', u0)
        w = u.detach()
        m, v = torch.zeros_like(w), torch.zeros_like(w)
        step = torch.zeros(1, dtype=torch.int64, device=DEV)
        loss = torch.zeros(2, dtype=torch.float64, device=DEV)
        got = []
        for _ in range(len(g['losses'])):
            sol.engine.train_step(w, m, v, step, 1e-3, 0.9, 0.999, 1e-8, loss)
            got.append(float(loss[0]))
        np.testing.assert_allclose(got, g['losses'], rtol=1e-9)


def test_mat_large_grid_against_oracle():
    """A 256 x 320 wave grid (several tiles, edges on every side) against the
    oracle: identical residual field, loss and gradient within tolerance."""
    d, eq, b, kw = problems.wave_mat(api, n=255)
    d2 = api.Domain()
    d2.variable('x', [0, 1], 255)
    d2.variable('t', [0, 1], 319)
    b2 = api.Conditions()
    x = d2.variable_dict['x']
    b2.dirichlet({'x': [0, 1], 't': 0}, value=torch.sin(np.pi * x))
    b2.dirichlet({'x': 0, 't': [0, 1]}, value=0)
    b2.dirichlet({'x': 1, 't': [0, 1]}, value=0)
    b2.operator({'x': [0, 1], 't': 0}, operator={'du/dt': {'coeff': 1, 'du/dt': [1], 'pow': 1}}, value=0)
    orc = O.MatOracle(dict(d2.variable_dict), eq.equation_lst, [dict(c) for c in b2.conditions_lst])
    u0 = problems.mat_init([1, 256, 320], seed=3)
    uo = u0.clone().requires_grad_()
    lo, no, op, bv, t, keys, lens = orc.evaluate(uo, 1, 100)
    lo.backward()
    grid = d2.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b2.build(d2.variable_dict)).set_strategy('mat')
    un = u0.to(DEV).requires_grad_()
    sol = Solution(grid, pre, un, 'mat', None, 1, 100)
    ln, nn = sol.evaluate()
    ln.backward()
    np.testing.assert_array_equal(sol.op.cpu().numpy(), op.detach().numpy())
    np.testing.assert_allclose(float(ln), float(lo), rtol=1e-5)
    np.testing.assert_allclose(un.grad.cpu().numpy(), uo.grad.numpy(), rtol=1e-4,
                               atol=1e-5 * float(uo.grad.abs().max()))


# ---------------------------------------------------------------------------
# drop-in Model API
# ---------------------------------------------------------------------------

def test_model_api_nn_fused_and_torch_optimizer_agree():
    """Model.compile/train: the fused graph-captured Adam path and torch's
    Adam over the native autograd Function reach the same parameters."""
    res = []
    for fused in (True, False):
        d, eq, b, kw = problems.burgers_nn(api, dtype='float64')
        net = problems.make_mlp([2, 12, 12, 12, 1], dtype=torch.float64)
        model = tdx.Model(net, d, eq, b)
        model.verbose = False
        model.compile('NN', lambda_operator=1, lambda_bound=10)
        opt = tdx.Optimizer('Adam', {'lr': 1e-3})
        cb = [] if fused else [object()]
        model.train(opt, 21, callbacks=cb or None)
        res.append(torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu())
    np.testing.assert_allclose(res[0].numpy(), res[1].numpy(), rtol=1e-9, atol=1e-12)


def test_model_api_nn_mini_batching():
    """batch_size in 'NN' mode: one full-operator Adam step per DataLoader
    batch (the reference's 'NN' derivative ignores the batch's points), so
    3 epochs of 3 batches reach the parameters of 9 unbatched steps, on the
    fused path and through torch's Adam."""
    def run(batch_size, epochs, fused):
        d, eq, b, kw = problems.burgers_nn(api, dtype='float64')
        net = problems.make_mlp([2, 12, 12, 12, 1], dtype=torch.float64)
        model = tdx.Model(net, d, eq, b, batch_size=batch_size)
        model.verbose = False
        model.compile('NN', lambda_operator=1, lambda_bound=10)
        n = len(d.build('NN'))
        model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), epochs,
                    callbacks=None if fused else [object()])
        return model, n, torch.cat([p.detach().reshape(-1) for p in net.parameters()]).cpu()

    _, n, _ = run(None, 1, True)
    bs = -(-n // 3)
    print(f"[case] burgers_nn float64, N={n}, batch_size={bs}", flush=True)
    mb, _, pb = run(bs, 4, True)
    assert mb.n_batches == 3
    _, _, pu = run(None, 10, True)
    np.testing.assert_array_equal(pb.numpy(), pu.numpy())
    _, _, pt = run(bs, 4, False)
    np.testing.assert_allclose(pt.numpy(), pu.numpy(), rtol=1e-9, atol=1e-12)
    mo, _, _ = run(n + 1, 2, True)  # larger than the grid: batching off
    assert mo.batch_size is None and mo.n_batches == 1


def test_model_api_mat_trains():
    d, eq, b, kw = problems.wave_mat(api, n=31)
    u = problems.mat_init([1, 32, 32]).to(DEV)
    model = tdx.Model(u, d, eq, b)
    model.verbose = False
    model.compile('mat', lambda_operator=1, lambda_bound=100)
    l0 = float(model.solution_cls.evaluate()[0])
    model.train(tdx.Optimizer('Adam', {'lr': 1e-3}), 200)
    l1 = float(model.solution_cls.evaluate()[0])
    assert l1 < 0.5 * l0


@pytest.mark.parametrize('name', ['burgers_nn', 'kdv_nn', 'legendre_nn_opbc'])
def test_nn_backward_kernels_agree(name, golden):
    """The wavefront-private backward (narrow nets) and the workgroup-
    cooperative backward (forced with tile_rows=-1) give the reference's
    float64 gradient."""
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch
    g = golden(f'{name}_float64')
    builder, layers = problems.NN_CASES[name]
    d, eq, b, kw = builder(api, dtype='float64')
    net = problems.make_mlp(layers, dtype=torch.float64)
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                      h=kw.get('h', 1e-3), lambda_operator=kw['lambda_operator'],
                      lambda_bound=kw['lambda_bound'])
    flat = torch.from_numpy(g['params']).to(DEV)
    grads = []
    for tile_rows, chunk in ((0, 0), (-1, 0), (-1, 3), (0, 1)):
        eng = NNEngine([prob], mlp_arch(net), DEV, tile_rows=tile_rows, chunk_tiles=chunk)
        p = flat.clone().requires_grad_()
        eng.loss(p)[0, 0].backward()
        grads.append(p.grad.cpu().numpy())
    for gr in grads:
        np.testing.assert_allclose(gr, g['grad'], rtol=1e-8, atol=1e-10 * np.abs(g['grad']).max())


@pytest.mark.parametrize("layers", [[2, 48, 48, 48, 1], [2, 64, 64, 64, 64, 1],
                                    [2, 100, 100, 100, 100, 1], [1, 100, 100, 1]])
def test_nn_wide_float32_backward(layers):
    """Wide float32 networks: the block-owning backward (default), the
    workgroup-cooperative one (tile_rows=-1) and work lists cut into 1- and
    3-tile items all give the float32 oracle's gradient (reference op order)
    to 1e-4 in norm at h = 1e-2."""
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch
    name = 'burgers_nn' if layers[0] == 2 else 'legendre_nn'
    builder, _ = problems.NN_CASES[name]
    d, eq, b, kw = builder(api, dtype='float32')
    net = problems.randomize_biases(problems.make_mlp(layers, dtype=torch.float32))
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=1e-2)
    lo = orc.evaluate(net, kw['lambda_operator'], kw['lambda_bound'])[0]
    lo.backward()
    go = torch.cat([p.grad.reshape(-1) for p in net.parameters()]).numpy()
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1, h=1e-2,
                      lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    flat = torch.cat([p.detach().reshape(-1) for p in net.parameters()]).to(DEV)
    for tile_rows, chunk in ((0, 0), (0, 1), (0, 3), (-1, 0)):
        _print(f"[wide backward {layers} tile_rows={tile_rows} chunk={chunk}]")
        eng = NNEngine([prob], mlp_arch(net), DEV, tile_rows=tile_rows, chunk_tiles=chunk)
        p = flat.clone().requires_grad_()
        ln = eng.loss(p)[0, 0]
        ln.backward()
        gn = p.grad.cpu().numpy()
        assert abs(float(ln) - float(lo)) / abs(float(lo)) < 1e-5, (float(ln), float(lo))
        err = np.linalg.norm(gn - go) / np.linalg.norm(go)
        assert err < 1e-4, (tile_rows, chunk, err)
