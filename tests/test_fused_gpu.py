"""The fused training step of narrow float32 networks (k_fused: forward,
residual program, adjoint and weight gradients in one kernel, no forward
recompute) against the separate kernels (k_forward, then the
wavefront-private k_backward_w), which the oracle tests pin
(test_full_size_gpu.py, test_native_gpu.py).

What must agree:
* the loss pairs of the step: bit for bit (the fused kernel evaluates the
  same network, residual program and seeds, and sums each 64-entry block's
  loss values in the forward kernel's pairwise tree);
* the gradient, read from the step itself (with beta1 = 0 torch's Adam
  stores m = 0 * m + 1 * g = g exactly): it is the sum over rows of
  seed * du/dparams, whose stencil weights (up to 1/h^2 = 1e6) cancel, so
  two float32 summation orders differ by far more than float32 epsilon
  (measured: up to 2e-3 of the largest entry).  Both kernels' gradients are
  therefore held against torch's float64 CPU backward of the same seeds (the
  separate kernels' seeds, the same bits as the fused kernel's) on the same
  float32 inputs, within four times the error torch's own float32 CPU
  backward makes there (or 1e-6), per instance in the L2 norm;
* two runs of the fused step: the same bits.
The plan with the separate kernels is requested with persistent = -1.
Which path ran is read from the launch profile (the fused step launches no
forward kernel) and from tdx_nn_step_kernel: plans of at least 2^17 network
rows compile the fused kernel with straight-line residual programs generated
from their tables (hipRTC), smaller ones interpret the tables;
TDX_NN_JIT=1 / 0 forces either.  The generated programs perform the
interpreter's operations in the same order: the two give the same bits.
"""
import copy
import sys

import numpy as np
import pytest
import torch

from conftest import gpu_available

pytestmark = pytest.mark.gpu

if not gpu_available():
    pytest.skip("no HIP device", allow_module_level=True)

import problems  # noqa: E402
from torch_de_solver_amd import data as api  # noqa: E402
from torch_de_solver_amd import workloads as W  # noqa: E402
from torch_de_solver_amd.compiler import compile_nn  # noqa: E402
from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch  # noqa: E402

DEV = torch.device('cuda')


def _print(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _flat(nets):
    return torch.cat([torch.cat([p.detach().reshape(-1) for p in n.parameters()])
                      for n in nets]).to(DEV)


def _step(eng, flat0, n_inst, steps=1, beta1=0.0):
    """Run `steps` train steps; returns (params, exp_avg, loss pairs of the
    last step, launches per kind of the last step)."""
    p = flat0.clone()
    m, v = torch.zeros_like(p), torch.zeros_like(p)
    step = torch.zeros(1, dtype=torch.int64, device=DEV)
    loss = torch.zeros(2 * n_inst, dtype=p.dtype, device=DEV)
    for k in range(steps):
        if k == steps - 1:
            eng.profile(4)
        eng.train_step(p, m, v, step, 1e-3, beta1, 0.999, 1e-8, loss)
    torch.cuda.synchronize()
    prof = eng.profile_read()
    eng.profile(0)
    counts = {k: c for k, (_, c) in prof.items()}
    return p.cpu(), m.cpu(), loss.cpu(), counts


def _row_inputs(prob):
    """float32 network input of every seed row of one problem, in row order
    (group by group, slot-major): base + fl(shift * h)."""
    xs = []
    for g in prob.groups:
        for sl, b in enumerate(g.slot_base):
            base = np.asarray(g.bases[b], dtype=np.float32).reshape(g.n_entries, -1)
            off = np.asarray(g.slot_offset[sl], dtype=np.float64).astype(np.float32)
            xs.append(base + off)
    return np.concatenate(xs, 0)


def _seed_backward(net, x, seeds, dt):
    model = copy.deepcopy(net).cpu().to(dt)
    out = model(torch.from_numpy(x).to(dt))
    torch.autograd.backward(out, seeds.to(dt))
    return torch.cat([q.grad.reshape(-1) for q in model.parameters()]).double().numpy()


def _compare(probs, nets, chunk_tiles=0, expect='fused'):
    arch = mlp_arch(nets[0])
    flat0 = _flat(nets)
    n = len(probs)
    fused = NNEngine(probs, arch, DEV, chunk_tiles=chunk_tiles)
    sep = NNEngine(probs, arch, DEV, chunk_tiles=chunk_tiles, persistent=-1)
    _print(f"  step kernels: {fused.step_kernel} (expected {expect}), {sep.step_kernel}")
    assert fused.step_kernel == expect and sep.step_kernel == 'separate'
    expect_fused = expect != 'separate'
    pf, gf, lf, cf = _step(fused, flat0, n)
    ps, gs, ls, cs = _step(sep, flat0, n)
    _print(f"  launches: fused {cf}, separate {cs}")
    assert cs['forward'] == 1 and cs['backward'] == 1
    if expect_fused:
        assert cf['forward'] == 0 and cf['backward'] == 1 and cf['update'] == 1
    else:
        assert cf == cs
    if expect.startswith('adjoint-seeds'):
        # the same per-entry loss values, summed per work item of whole-entry
        # tiles instead of per 64-entry block
        np.testing.assert_allclose(lf.numpy(), ls.numpy(), rtol=1e-6)
    else:
        assert torch.equal(lf, ls), (lf - ls).abs().max()
    # the seeds of every row (separate kernels; the fused kernel forms the
    # same bits) and the float64 / float32 CPU backward of them per instance
    u = sep.network_values(flat0)
    seeds = torch.empty_like(u)
    lv = torch.empty(2 * n, dtype=u.dtype, device=DEV)
    sep.forward_values(u, lv, None, seeds)
    seeds = seeds.cpu()
    P = arch.n_params
    r0 = 0
    worst = 0.0
    for i in range(n):
        x = _row_inputs(probs[i])
        S = seeds[r0:r0 + x.shape[0]]
        r0 += x.shape[0]
        g64 = _seed_backward(nets[i], x, S, torch.float64)
        g32 = _seed_backward(nets[i], x, S, torch.float32)
        nrm = np.linalg.norm(g64)
        e_f = np.linalg.norm(gf[i * P:(i + 1) * P].double().numpy() - g64) / nrm
        e_s = np.linalg.norm(gs[i * P:(i + 1) * P].double().numpy() - g64) / nrm
        e_c = np.linalg.norm(g32 - g64) / nrm
        worst = max(worst, e_f / max(e_c, 1e-12))
        _print(f"  instance {i}: gradient error vs float64 backward of the seeds: fused "
               f"{e_f:.2e}, separate {e_s:.2e}, torch float32 {e_c:.2e}")
        assert e_f <= max(4 * e_c, 1e-6), (i, e_f, e_c)
        assert e_s <= max(4 * e_c, 1e-6), (i, e_s, e_c)
    assert r0 == seeds.shape[0]
    _print(f"  largest fused error / torch float32 error: {worst:.2f}")
    # a second run of the fused step gives the same bits (three steps, Adam)
    p1, _, l1, _ = _step(fused, flat0, n, steps=3, beta1=0.9)
    p2, _, l2, _ = _step(fused, flat0, n, steps=3, beta1=0.9)
    assert torch.equal(p1, p2) and torch.equal(l1, l2)


def _sweep(idx):
    insts = W.sweep_instances(256)
    probs, nets = [], []
    for i in idx:
        d, eq, b, net, ckw = W.sweep_problem(insts[i])
        probs.append(compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                                h=ckw['h'], lambda_operator=ckw['lambda_operator'],
                                lambda_bound=ckw['lambda_bound']))
        nets.append(net)
    return probs, nets


@pytest.mark.parametrize('chunk_tiles', [0, 1, 3])
def test_fused_sweep_instances_match_separate_kernels(chunk_tiles):
    """Twelve sweep instances spread over (nu, nx, nt), MLP 2-32x3-1 (two
    16-column blocks, two 32 x 32 layers, 7 slots at the collocation points,
    1 at the boundary points); work items of 1, 3 and 4 blocks, so items end
    inside a group and at its ragged end."""
    idx = list(np.linspace(0, 255, 12).round().astype(int))
    _print(f"[fused sweep] instances {idx}, chunk_tiles {chunk_tiles}")
    probs, nets = _sweep(idx)
    _compare(probs, nets, chunk_tiles, expect='fused-generated')


def test_generated_programs_give_the_interpreters_bits(monkeypatch):
    """The sweep instances with the generated residual programs and with the
    interpreted ones: three Adam steps give the same parameters and loss
    pairs, bit for bit."""
    idx = list(np.linspace(0, 255, 12).round().astype(int))
    probs, nets = _sweep(idx)
    arch = mlp_arch(nets[0])
    flat0 = _flat(nets)
    out = {}
    for jit in ('1', '0'):
        monkeypatch.setenv('TDX_NN_JIT', jit)
        eng = NNEngine(probs, arch, DEV)
        _print(f"[generated vs interpreted] TDX_NN_JIT={jit}: {eng.step_kernel}")
        assert eng.step_kernel == ('fused-generated' if jit == '1' else 'fused')
        p, m, loss, _ = _step(eng, flat0, len(probs), steps=3, beta1=0.9)
        out[jit] = (p, m, loss)
    for a, b in zip(out['1'], out['0']):
        assert torch.equal(a, b)


def _small(builder, layers, seed, **kw):
    d, eq, b, ckw = builder(api, **kw)
    net = problems.randomize_biases(problems.make_mlp(layers, seed=seed), seed=seed + 1, scale=0.3)
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                      h=ckw.get('h', 0.001), lambda_bound=ckw.get('lambda_bound', 1))
    return prob, net


@pytest.mark.parametrize('jit', ['0', '1'])
@pytest.mark.parametrize('case', ['burgers_12x3', 'burgers_32x1', 'burgers_20x2',
                                  'legendre_16x2', 'legendre_9x1', 'kdv_10x3',
                                  'legendre_opbc_16x2'])
def test_fused_shapes_match_separate_kernels(case, jit, monkeypatch):
    """Every instantiated shape: one and two 16-column blocks (widths 9, 10,
    12, 16, 20, 32, padding features included), zero to two H x H layers,
    one and two inputs, periodic KdV conditions, an operator condition
    (one-sided stencil at t = 1) and per-entry coefficients (Legendre's
    1 - t^2, -2t); a grouped plan of three
    instances of different sizes; interpreted and generated residual
    programs."""
    monkeypatch.setenv('TDX_NN_JIT', jit)
    spec = {
        'burgers_12x3': (problems.burgers_nn, [2, 12, 12, 12, 1]),
        'burgers_32x1': (problems.burgers_nn, [2, 32, 1]),
        'burgers_20x2': (problems.burgers_nn, [2, 20, 20, 1]),
        'legendre_16x2': (problems.legendre_nn, [1, 16, 16, 1]),
        'legendre_9x1': (problems.legendre_nn, [1, 9, 1]),
        'kdv_10x3': (problems.kdv_nn, [2, 10, 10, 10, 1]),
        'legendre_opbc_16x2': (problems.legendre_nn_opbc, [1, 16, 16, 1]),
    }[case]
    builder, layers = spec
    _print(f"[fused shape] {case}, TDX_NN_JIT={jit}")
    probs, nets = [], []
    for k, size in enumerate((None, 'mid', 'big')):
        kw = {}
        if size and builder in (problems.burgers_nn, problems.kdv_nn):
            kw = {'nx': 23 if size == 'mid' else 41, 'nt': 11 if size == 'mid' else 17}
        elif size:
            kw = {'n': 37 if size == 'mid' else 150}
        prob, net = _small(builder, layers, seed=10 + k, **kw)
        probs.append(prob)
        nets.append(net)
    _compare(probs, nets, expect='fused-generated' if jit == '1' else 'fused')


@pytest.mark.parametrize('jit', ['0', '1'])
@pytest.mark.parametrize('case', ['burgers_100x4', 'kdv_64x4', 'legendre_opbc_48x2'])
def test_wide_step_forms_its_own_seeds(case, jit, monkeypatch):
    """Wide float networks (48 to 128 features): the training step's
    block-owning adjoint evaluates the network on tiles of whole entries,
    runs the residual program there and forms its own seeds (no forward
    kernel); against the separate kernels.  A grouped plan of three
    instances of different sizes (7, 2 and 1 slots per group; tiles of 9,
    32 and 64 entries, ragged ends); interpreted and generated residual
    programs."""
    monkeypatch.setenv('TDX_NN_JIT', jit)
    builder, layers = {
        'burgers_100x4': (problems.burgers_nn, [2, 100, 100, 100, 100, 1]),
        'kdv_64x4': (problems.kdv_nn, [2, 64, 64, 64, 64, 1]),
        'legendre_opbc_48x2': (problems.legendre_nn_opbc, [1, 48, 48, 1]),
    }[case]
    _print(f"[wide step] {case}, TDX_NN_JIT={jit}")
    probs, nets = [], []
    for k, size in enumerate((None, 'mid', 'big')):
        kw = {}
        if size and builder in (problems.burgers_nn, problems.kdv_nn):
            kw = {'nx': 23 if size == 'mid' else 41, 'nt': 11 if size == 'mid' else 17}
        elif size:
            kw = {'n': 37 if size == 'mid' else 150}
        prob, net = _small(builder, layers, seed=20 + k, **kw)
        probs.append(prob)
        nets.append(net)
    for chunk in (0, 1):
        _compare(probs, nets, chunk_tiles=chunk,
                 expect='adjoint-seeds-generated' if jit == '1' else 'adjoint-seeds')


def _mixed(layers, seed):
    """u_xx + u_tt + u_xt / 2 + 1 = 0 on a 9 x 7 grid: nine shifted copies
    per collocation point (centre, four axis neighbours, four corners)."""
    d = api.Domain()
    d.variable('x', [0, 1], 9)
    d.variable('t', [0, 1], 7)
    b = api.Conditions()
    b.dirichlet({'x': [0, 1], 't': 0}, value=0.0)
    b.dirichlet({'x': 0, 't': [0, 1]}, value=0.0)
    eq = api.Equation()
    eq.add({'d2u/dx2': {'coeff': 1, 'term': [0, 0], 'pow': 1},
            'd2u/dt2': {'coeff': 1, 'term': [1, 1], 'pow': 1},
            'd2u/dxdt': {'coeff': 0.5, 'term': [0, 1], 'pow': 1},
            'f': {'coeff': 1.0, 'term': [None], 'pow': 0}})
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1, h=0.001)
    net = problems.randomize_biases(problems.make_mlp(layers, seed=seed), seed=seed + 1, scale=0.3)
    return prob, net


def test_slots_beyond_the_fused_limit_use_separate_kernels():
    """A mixed second derivative reads nine shifted copies per point, beyond
    the fused kernel's eight slots: the plan keeps the separate kernels
    (same launches, same bits)."""
    prob, net = _mixed([2, 16, 16, 1], seed=3)
    slots = max(len(g.slot_base) for g in prob.groups)
    _print(f"[fused fallback] largest slot count {slots}")
    assert slots > 8
    _compare([prob], [net], expect='separate')
