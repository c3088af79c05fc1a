#!/usr/bin/env python3
"""Training-throughput benchmark (one JSON line on stdout).

Default workload: the sweep of 256 independent Burgers solves (each with its
own 2-32-32-32-1 tanh network, nu = 10^U(-3,-1), nx in [64, 512],
nt in [32, 128]) trained together by the fused native step (forward,
hand-written adjoint and Adam of every instance in one grouped launch
sequence, replayed from a HIP graph).  A step is one Adam step of every
instance.  Reported next to it, measured in the same run:

* plain_gpu: the same sweep trained by the reference's algorithm in stock
  torch ops on the same GPU, one instance after the other (the per-instance
  path a reference user runs);
* cpu_baseline: the CPU oracle (an op-for-op restatement of the reference in
  torch CPU ops) timed on the host cores on a bounded sample of instances,
  scaled to the full sweep by collocation points;
* roofline: FP32-MFMA roofline of the dominant kernel (the backward kernel),
  from HIP-event timing of its launches during a timed pass;
* accuracy: a fresh copy of the workload trained by the fused path and by
  the plain path from the same initial weights for the same number of Adam
  steps, each instance's rel-L2 error on its collocation points against the
  FP64 host solution (Cole-Hopf quadrature for Burgers, closed forms for
  KdV and Legendre), for a spread of instances.

Other workloads (--workload): burgers, kdv, legendre ('NN', one instance),
wave (the 128 x 128 'mat' problem), wave16k (the wave operator on a
16384 x 16384 field, the HBM-roofline case of the stencil kernels) and
poisson_ad (the 'autograd' mode: a 2-D Poisson problem on 201 x 201 points
with the 2-100x5-1 network of the reference's Poisson examples; roofline of
the whole fused step against FP32 matrix peak from the algorithmic FLOPs of
its linear jet maps) and ngd_lotka (natural-gradient descent: NGD steps/s
on the reference's Lotka-Volterra NGD example in 'autograd' mode, 501
points, 1-32-32-2 network; the reference's NGD algorithm in stock torch on
the same GPU and the CPU oracle's NGD step beside it).  The
'mat' step reads one grid buffer and writes the other (as Model.train does),
so the kernel generated for the grid interior applies Adam as it forms each
gradient; its roofline is that whole step at 24 B per point, the residual +
adjoint pass alone is reported beside it at 8 B per point.

--calibrate runs two microbenchmarks instead: a stream copy of 1 GiB
(out = -in, float32: 1 GiB read + 1 GiB written per launch; its kernel is
also the known-size reference for scaling rocprofv3's FETCH_SIZE and
WRITE_SIZE counters) and an FP32 GEMM (torch.mm, 8192^3) for the attainable
FP32-matrix rate.
"""

import argparse
import json
import math
import os
import sys
import time

import numpy as np
import torch

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP32_MFMA_PEAK = 157.3e12     # MI355X dense FP32 matrix, FLOP/s
HBM_PEAK = 8.0e12             # MI355X HBM3E, B/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def profiled_traffic(name: str):
    """HBM bytes per launch of a kernel measured by rocprofv3 FETCH_SIZE and
    WRITE_SIZE passes (each scaled by what it reads for a copy kernel of
    known size), from the summary ``profiles/<name>_traffic.txt``, whose
    ``traffic_bytes_per_launch:`` line holds the figure; None when absent."""
    path = os.path.join(ROOT, 'profiles', name + '_traffic.txt')
    try:
        with open(path) as f:
            for line in f:
                if line.startswith('traffic_bytes_per_launch:'):
                    return float(line.split(':', 1)[1])
    except OSError:
        pass
    return None


def mlp_flops_per_row(arch) -> int:
    """2 * multiply-adds of one network evaluation."""
    H, d, o, L = arch.width, arch.d_in, arch.n_out, arch.n_layers
    return 2 * (d * H + (L - 2) * H * H + H * o)


def time_cuda(fn, n):
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(n):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / n


def cpu_cores() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except Exception:
        return os.cpu_count() or 1


# ---------------------------------------------------------------------------
# 'NN' workloads
# ---------------------------------------------------------------------------

def nn_problems(workload):
    from torch_de_solver_amd import workloads as W
    if workload == 'sweep':
        insts = W.sweep_instances(256)
        return [W.sweep_problem(i) for i in insts], insts
    if workload == 'burgers':
        return [W.burgers()], None
    if workload == 'kdv':
        return [W.kdv()], None
    if workload == 'legendre':
        return [W.legendre()], None
    raise ValueError(workload)


def plain_nn_step_ms(problems, nets0, n_steps, max_instances, lr):
    """Per-instance plain path on the GPU: ms of one sweep step (sum over
    instances), plus the sample description."""
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.plain import PlainNN
    dev = torch.device('cuda')
    total_ms, pts_done, pts_all = 0.0, 0, 0
    sel = list(range(len(problems)))[:max_instances]
    comp = []
    for i, (dom, eq, conds, net, kw) in enumerate(problems):
        prob = compile_nn(dom.build('NN'), eq.equation_lst, conds.build(dom.variable_dict),
                          1, h=kw.get('h', 1e-3), lambda_operator=kw['lambda_operator'],
                          lambda_bound=kw['lambda_bound'])
        pts_all += prob.n_central
        comp.append(prob)
    for i in sel:
        net = nets0[i].to(dev)
        pl = PlainNN(comp[i], net, dev)
        opt = torch.optim.Adam(net.parameters(), lr=lr)

        def step():
            opt.zero_grad()
            loss = pl.loss()
            loss.backward()
            opt.step()
        step()
        # median of three timed repeats after the warm-up step
        total_ms += float(np.median([time_cuda(step, n_steps) for _ in range(3)]))
        pts_done += comp[i].n_central
    scale = pts_all / max(pts_done, 1)
    return total_ms * scale, len(sel), scale


def cpu_nn_step_ms(problems, nets0, budget_s, lr):
    """CPU oracle step time for the sweep, from a bounded sample of
    instances scaled by collocation points."""
    from oracle import tedeous_oracle as O
    torch.set_num_threads(cpu_cores())
    spent, pts_done, n_inst, steps_done, ms_sum = 0.0, 0, 0, 0, 0.0
    pts_all = 0
    from torch_de_solver_amd.workloads import collocation_points
    for dom, eq, conds, net, kw in problems:
        axes = list(dom.variable_dict.values())
        pts_all += collocation_points(len(axes[0]), len(axes[1])) if len(axes) == 2 else len(axes[0])
    t_start = time.time()
    for i, (dom, eq, conds, net, kw) in enumerate(problems):
        if time.time() - t_start > budget_s and n_inst > 0:
            break
        orc = O.NNOracle(dict(dom.variable_dict), eq.equation_lst,
                         [dict(c) for c in conds.conditions_lst], h=kw.get('h', 1e-3))
        netc = nets0[i].to('cpu')
        opt = torch.optim.Adam(netc.parameters(), lr=lr)

        def step():
            t0 = time.perf_counter()
            opt.zero_grad()
            loss = orc.evaluate(netc, kw['lambda_operator'], kw['lambda_bound'])[0]
            loss.backward()
            opt.step()
            return (time.perf_counter() - t0) * 1e3
        step()  # first-call overhead stays out of the sample
        ms_sum += float(np.median([step() for _ in range(3)]))
        pts_done += len(orc.central)
        n_inst += 1
    return ms_sum * pts_all / pts_done, n_inst, pts_done, pts_all


def exact_solution(workload, inst):
    """FP64 host solution u*(points [n, d]) of an 'NN' workload instance."""
    from torch_de_solver_amd import exact as X
    if workload == 'sweep':
        return lambda p: X.burgers(p[:, 0], p[:, 1], inst.nu)
    if workload == 'burgers':
        return lambda p: X.burgers(p[:, 0], p[:, 1], 0.01 / math.pi)
    if workload == 'kdv':
        return lambda p: X.kdv_soliton(p[:, 0], p[:, 1])
    if workload == 'legendre':
        return lambda p: X.legendre_p3(p[:, 0])
    raise ValueError(workload)


def accuracy_nn(args, lr):
    """Rel-L2 of fused-trained and plain-trained solutions, same init, same
    number of Adam steps, against the FP64 host solution."""
    import numpy as np
    from torch_de_solver_amd import exact as X
    from torch_de_solver_amd.plain import PlainNN
    from torch_de_solver_amd.sweep import SweepSolver
    problems, insts = nn_problems(args.workload)
    n = len(problems)
    sel = sorted(set(int(round(v)) for v in np.linspace(0, n - 1, min(args.acc_instances, n))))
    solver = SweepSolver(problems)
    dev, dt = solver.device, solver.params.dtype
    pts, uref = {}, {}
    for i in sel:
        grid = problems[i][0].build('NN')
        P = grid[torch.as_tensor(solver.problems[i].central_index)]
        pts[i] = P.to(dev, dt)
        uref[i] = exact_solution(args.workload, insts[i] if insts else None)(
            P.double().numpy())

    def rel(net, i):
        with torch.no_grad():
            u = net(pts[i])[:, 0].double().cpu().numpy()
        return X.relative_l2(u, uref[i])

    untrained = [rel(solver.nets[i], i) for i in sel]
    t0 = time.time()
    solver.train(args.acc_steps, lr)
    torch.cuda.synchronize()
    fused = [rel(solver.nets[i], i) for i in sel]
    log(f"[bench] accuracy: fused {args.acc_steps} steps in {time.time() - t0:.1f}s")
    fresh, _ = nn_problems(args.workload)
    plain = []
    t0 = time.time()
    for i in sel:
        net = fresh[i][3].to(dev, dt)
        pl = PlainNN(solver.problems[i], net, dev)
        opt = torch.optim.Adam(net.parameters(), lr=lr)
        for _ in range(args.acc_steps):
            opt.zero_grad()
            pl.loss().backward()
            opt.step()
        plain.append(rel(net, i))
    torch.cuda.synchronize()
    log(f"[bench] accuracy: plain {len(sel)} x {args.acc_steps} steps in {time.time() - t0:.1f}s")
    ratio = [f / p for f, p in zip(fused, plain)]
    return {"steps": args.acc_steps, "instances": sel,
            "reference_solution": "FP64 host: " + {
                'sweep': "Cole-Hopf quadrature of viscous Burgers",
                'burgers': "Cole-Hopf quadrature of viscous Burgers",
                'kdv': "KdV soliton 0.5 sech^2(0.5(x-t-10)), periodic images",
                'legendre': "P3(x) = (5x^3-3x)/2"}[args.workload],
            "points": "collocation ('central') points of each instance",
            "rel_l2_fused": fused, "rel_l2_plain": plain, "rel_l2_untrained": untrained,
            "fused_over_plain": ratio,
            "within_10pct": sum(1 for r in ratio if abs(r - 1.0) <= 0.10),
            "of": len(sel)}


def bench_nn(args):
    from torch_de_solver_amd.nn_engine import mlp_arch
    from torch_de_solver_amd.sweep import SweepSolver
    problems, insts = nn_problems(args.workload)
    import copy
    nets0 = [copy.deepcopy(p[3]) for p in problems]
    t0 = time.time()
    solver = SweepSolver(problems, autotune=not args.no_tune, tile_rows=args.nn_tile_rows,
                         chunk_tiles=args.chunk_tiles, persistent=args.persistent)
    log(f"[bench] compiled {solver.n_inst} instance(s) in {time.time() - t0:.1f}s"
        f" (tuning: {solver.tuning})")
    eng = solver.engine
    arch = solver.arch
    n_rows = eng.tables.n_rows
    colloc = solver.collocation_points()
    lr = 1e-3
    # warm-up: one eager step, capture, W replays
    solver.capture(lr)
    solver.replay(args.warmup)
    torch.cuda.synchronize()
    log(f"[bench] warm-up done ({args.warmup} replays)")
    ms = time_cuda(lambda: solver.replay(1), args.steps)
    loss_now = solver.loss_buf.view(-1, 2)[:, 0].sum().item()
    log(f"[bench] fused: {ms:.3f} ms/step")
    # roofline pass: eager steps with event timing of every kernel launch
    eng.profile(args.steps)
    for _ in range(args.steps):
        solver.step(lr)
    torch.cuda.synchronize()
    prof = eng.profile_read()
    eng.profile(0)
    F = mlp_flops_per_row(arch)
    # the fused step (k_fused: forward and adjoint in one kernel) launches no
    # separate forward kernel; its kernel is profiled under 'backward'
    fused = prof['forward'][1] == 0
    bwd_ms = prof['backward'][0] / max(prof['backward'][1], 1)
    fwd_ms = prof['forward'][0] / max(prof['forward'][1], 1) if not fused else None
    upd_ms = prof['update'][0] / max(prof['update'][1], 1)
    # algorithmic FLOPs: F per row forward, 2F per row adjoint (a recompute
    # of the forward inside the adjoint is not counted)
    bwd_flops = (3.0 if fused else 2.0) * F * n_rows
    fwd_flops = 1.0 * F * n_rows
    achieved = bwd_flops / (bwd_ms * 1e-3)
    if fused:
        log(f"[bench] kernels: fused forward+adjoint {bwd_ms:.3f} ms, update {upd_ms:.3f} ms")
    else:
        log(f"[bench] kernels: forward {fwd_ms:.3f} ms, backward {bwd_ms:.3f} ms, "
            f"update {upd_ms:.3f} ms")
    kern = {'fused-generated': 'tdx_fused_jit', 'fused': 'k_fused',
            'adjoint-seeds-generated': 'tdx_block_jit',
            'adjoint-seeds': 'k_backward_c'}.get(eng.step_kernel, 'k_backward')
    kname = {'tdx_fused_jit': "tdx_fused_jit (forward + generated residual programs + adjoint "
                              "+ weight gradients, compiled at plan creation)",
             'k_fused': "k_fused (forward + interpreted residual programs + adjoint + weight "
                        "gradients)",
             'tdx_block_jit': "tdx_block_jit (block-owning: forward + generated residual "
                              "programs + adjoint + weight gradients, compiled at plan creation)",
             'k_backward_c': "k_backward_c (block-owning: forward + interpreted residual "
                             "programs + adjoint + weight gradients)",
             'k_backward': "k_backward (recompute + adjoint + weight gradients)"}[kern]
    tname = f"{kern}_{args.workload}"
    out = {
        "metric": "train steps/s & collocation pts/s, fused+adjoint vs plain GPU, at equal rel-L2",
        "value": 1e3 / ms,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload if args.workload != 'sweep' else
                   "sweep: 256 Burgers solves, nu=10^U(-3,-1), nx~U[64,512], nt~U[32,128], "
                   "MLP 2-32-32-32-1 tanh each, Adam lr 1e-3, h=1e-3",
                   "instances": solver.n_inst, "collocation_points": colloc,
                   "network_rows_per_step": n_rows, "mlp": f"{arch.d_in}-{arch.width}x{arch.n_layers - 1}-{arch.n_out}",
                   "parallelism": "single GPU",
                   "autotune": solver.tuning},
        "colloc_pts_per_s": colloc * 1e3 / ms,
        "loss_sum_after": loss_now,
        "roofline": {"bound": "mfma", "kernel": kname,
                     "achieved": achieved / 1e12, "peak": FP32_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / FP32_MFMA_PEAK,
                     "traffic": profiled_traffic(tname) if args.workload == 'sweep' else None,
                     "traffic_source": f"profiles/{tname}_traffic.txt (rocprofv3 "
                                       "FETCH_SIZE/WRITE_SIZE, copy-calibrated)",
                     "algorithmic_flops_per_launch": bwd_flops,
                     "avg_launch_ms": bwd_ms},
        "kernel_ms": ({"fused": bwd_ms, "reduce_adam": upd_ms} if fused else
                      {"forward": fwd_ms, "backward": bwd_ms, "reduce_adam": upd_ms,
                       "forward_tflops": fwd_flops / (fwd_ms * 1e-3) / 1e12}),
    }
    if not args.no_plain:
        plain_ms, n_sel, scale = plain_nn_step_ms(problems, nets0, args.plain_steps,
                                                  args.plain_instances, lr)
        out["plain_gpu"] = {"value": 1e3 / plain_ms, "unit": "steps/s", "ms_per_step": plain_ms,
                            "instances_timed": n_sel, "scaled_by_points": scale,
                            "steps_per_instance": args.plain_steps,
                            "repeats": "median of 3 timed repeats per instance after a warm-up step"}
        out["speedup_vs_plain_gpu"] = plain_ms / ms
        log(f"[bench] plain GPU: {plain_ms:.1f} ms/step ({n_sel} instances timed)")
    if not args.no_cpu:
        cms, n_inst, pts, pts_all = cpu_nn_step_ms(problems, nets0, args.cpu_seconds, lr)
        out["cpu_baseline"] = {"value": 1e3 / cms, "unit": "steps/s", "cores": cpu_cores(),
                               "kind": "port",
                               "sample": f"{n_inst} instance(s): per instance one warm-up Adam "
                                         f"step of the CPU oracle, then the median of 3 timed "
                                         f"steps ({pts} of {pts_all} collocation points), "
                                         f"scaled by points"}
        log(f"[bench] cpu oracle: {cms:.0f} ms/step (scaled)")
    if args.acc_steps > 0:
        out["accuracy"] = accuracy_nn(args, lr)
    return out


# ---------------------------------------------------------------------------
# 'mat' workloads
# ---------------------------------------------------------------------------

def bench_mat(args):
    import numpy as np
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.compiler import compile_mat
    from torch_de_solver_amd.mat_engine import MatEngine
    from torch_de_solver_amd.plain import PlainMat
    n = 16384 if args.workload == 'wave16k' else 128
    dom, eq, conds, u0, kw = W.wave(n=n)
    t0 = time.time()
    grid = dom.build('mat')
    prob = compile_mat(grid, eq.equation_lst, conds.build(dom.variable_dict), 1,
                       lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    dev = torch.device('cuda')
    u = u0.to(dev).contiguous()
    tuning = None
    tile_rows = args.tile_rows
    if not args.no_tune and tile_rows == 0:
        from torch_de_solver_amd.autotune import tune_mat
        tile_rows, tuning = tune_mat(prob, dev, u, exact_div=not args.fast_div)
    eng = MatEngine(prob, dev, exact_div=not args.fast_div, tile_rows=tile_rows,
                    specialize=args.specialize, band_rows=args.band_rows)
    spec_pts = eng.specialized_points
    log(f"[bench] compiled {n}x{n} in {time.time() - t0:.1f}s (tuning: {tuning}; "
        f"specialized interior points: {spec_pts})")
    m, v = torch.zeros_like(u), torch.zeros_like(u)
    step = torch.zeros(1, dtype=torch.int64, device=dev)
    loss = torch.zeros(2, device=dev)
    lr = 1e-3

    # the training step reads one grid buffer and writes the other (as
    # Model.train does): with the grid read-only the generated interior
    # kernel applies Adam as it forms each gradient
    bufs = [u, torch.empty_like(u)]
    done = [0]

    def one(k=None):
        k = done[0] if k is None else k
        eng.train_step_to(bufs[k % 2], bufs[(k + 1) % 2], m, v, step, lr, 0.9, 0.999, 1e-8,
                          loss)
        done[0] += 1
    one()
    one()
    torch.cuda.synchronize()
    graphs = [torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()]
    for par, g in enumerate(graphs):
        with torch.cuda.graph(g):
            eng.train_step_to(bufs[par], bufs[1 - par], m, v, step, lr, 0.9, 0.999, 1e-8, loss)

    def replay():
        graphs[done[0] % 2].replay()
        done[0] += 1
    for _ in range(args.warmup):
        replay()
    ms = time_cuda(replay, args.steps)
    N = n * n
    # roofline of the fused step (every launch of tdx_mat_train_step_to: the
    # generated interior kernel with its Adam update, the general kernel's
    # edge tiles, the boundary kernels and the edge points' Adam update):
    # 24 B/point (read u, m, v; write the new u, m, v)
    step_ms = time_cuda(one, args.steps)
    achieved = 24.0 * N / (step_ms * 1e-3)
    # the residual + adjoint pass alone (tdx_mat_backward): 8 B/point
    dl = torch.ones(1, device=dev)
    grad = torch.empty_like(u)
    bwd_ms = time_cuda(lambda: eng.backward(bufs[done[0] % 2], dl, grad), args.steps)
    out = {
        "metric": "train steps/s & collocation pts/s, fused+adjoint vs plain GPU, at equal rel-L2",
        "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"wave u_tt=4u_xx 'mat' {n}x{n}, Adam lr 1e-3",
                   "collocation_points": N, "parallelism": "single GPU",
                   "tile_rows": tile_rows, "autotune": tuning,
                   "specialized_points": spec_pts, "band_rows": args.band_rows},
        "colloc_pts_per_s": N * 1e3 / ms,
        "roofline": {"bound": "hbm",
                     "kernel": ("fused train step: tdx_mat_jit_step (generated for the program: "
                                "residual + loss + adjoint + Adam on the interior) + k_mat_tile "
                                "(edge tiles) + boundary kernels + k_mat_frame_adam (edge points)")
                     if spec_pts else "fused train step: k_mat_tile + k_mat_frame_adam",
                     "achieved": achieved / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK,
                     "traffic": profiled_traffic(f"mat_step_{args.workload}") if spec_pts
                     else None,
                     "traffic_source": f"profiles/mat_step_{args.workload}_traffic.txt "
                                       "(rocprofv3 FETCH_SIZE/WRITE_SIZE summed over the step's "
                                       "kernels, copy-calibrated)",
                     "algorithmic_bytes_per_launch": 24.0 * N, "avg_launch_ms": step_ms},
        "adjoint_pass": {"kernel": "tdx_mat_backward (residual + loss + adjoint, no update)",
                         "ms": bwd_ms, "GB_per_s": 8.0 * N / (bwd_ms * 1e-3) / 1e9,
                         "frac": 8.0 * N / (bwd_ms * 1e-3) / HBM_PEAK},
    }
    if not args.no_plain and n <= 4096:
        pl = PlainMat(prob, grid.to(dev))
        up = u0.to(dev).clone().requires_grad_()
        opt = torch.optim.Adam([up], lr=lr)

        def pstep():
            opt.zero_grad()
            l = pl.loss(up)
            l.backward()
            opt.step()
        pstep()
        pms = time_cuda(pstep, max(args.plain_steps, 3))
        out["plain_gpu"] = {"value": 1e3 / pms, "unit": "steps/s", "ms_per_step": pms}
        out["speedup_vs_plain_gpu"] = pms / ms
    if not args.no_cpu and n <= 4096:
        from oracle import tedeous_oracle as O
        torch.set_num_threads(cpu_cores())
        orc = O.MatOracle(dict(dom.variable_dict), eq.equation_lst,
                          [dict(c) for c in conds.conditions_lst])
        uc = u0.clone().requires_grad_()
        opt = torch.optim.Adam([uc], lr=lr)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 10) or k < 2:
            opt.zero_grad()
            l = orc.evaluate(uc, kw['lambda_operator'], kw['lambda_bound'])[0]
            l.backward()
            opt.step()
            k += 1
        cms = (time.perf_counter() - t0) * 1e3 / k
        out["cpu_baseline"] = {"value": 1e3 / cms, "unit": "steps/s", "cores": cpu_cores(),
                               "kind": "port", "sample": f"{k} Adam steps of the CPU oracle"}
    return out


# ---------------------------------------------------------------------------
# 'autograd' workload
# ---------------------------------------------------------------------------

def ad_flops_per_step(eng) -> float:
    """Algorithmic FLOPs of one fused 'autograd' step: per point and jet
    component, every layer's linear map forward (2 din dout), its transpose
    backward for layers above the first (2 din dout) and its weight-gradient
    product (2 (din + 1) dout)."""
    a = eng.arch
    dims = [(a.d_in if l == 0 else a.width, a.n_out if l == a.n_layers - 1 else a.width)
            for l in range(a.n_layers)]
    rows = eng.tables.n_points * eng.tables.n_jets
    fwd = sum(2 * di * do for di, do in dims)
    bwd = sum(2 * di * do for di, do in dims[1:])
    wg = sum(2 * (di + 1) * do for di, do in dims)
    return float(rows) * (fwd + bwd + wg)


def bench_ad(args):
    import copy
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.ad_engine import ADEngine
    from torch_de_solver_amd.compiler import compile_ad
    from torch_de_solver_amd.nn_engine import flatten_module_params, mlp_arch
    from torch_de_solver_amd.plain import PlainAD
    dom, eq, conds, net0, kw = W.poisson_ad(args.grid_res)
    grid = dom.build('autograd')
    prob = compile_ad(grid, eq.equation_lst, conds.build(dom.variable_dict), 1,
                      lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    dev = torch.device('cuda')
    arch = mlp_arch(net0)
    eng = ADEngine(prob, arch, dev)
    net = copy.deepcopy(net0).to(dev)
    params = flatten_module_params(net)
    m, v = torch.zeros_like(params), torch.zeros_like(params)
    step = torch.zeros(1, dtype=torch.int64, device=dev)
    loss = torch.zeros(2, device=dev)
    lr = 1e-3

    def one():
        eng.train_step(params, m, v, step, lr, 0.9, 0.999, 1e-8, loss)
    one()
    one()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        one()
    for _ in range(args.warmup):
        g.replay()
    ms = time_cuda(g.replay, args.steps)
    # passes of the step, timed on their own (the stream they run on)
    dl = torch.ones(1, device=dev)
    grad = torch.empty_like(params)
    fwd_ms = time_cuda(lambda: eng.forward(params, loss), args.steps)
    bwd_ms = time_cuda(lambda: eng.backward(params, dl, grad), args.steps)
    step_ms = time_cuda(one, args.steps)
    flops = ad_flops_per_step(eng)
    N = prob.n_grid
    out = {
        "metric": "train steps/s & collocation pts/s, fused+adjoint vs plain GPU, at equal rel-L2",
        "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"Poisson 'autograd' {args.grid_res + 1}x{args.grid_res + 1}, "
                               "MLP 2-100x5-1, Adam lr 1e-3",
                   "collocation_points": N, "points_with_conditions": eng.tables.n_points,
                   "jets": [list(j) for j in prob.jets], "parallelism": "single GPU"},
        "colloc_pts_per_s": N * 1e3 / ms,
        "roofline": {"bound": "mfma",
                     "kernel": {
                         'fused-generated': "fused train step: k_ad_pack, tdx_ad_step_jit "
                                            "(forward jets, generated residual programs and "
                                            "jet maps, adjoint, weight gradients; compiled at "
                                            "plan creation), k_ad_step_sum",
                         'fused': "fused train step: k_ad_pack, k_ad_step (forward jets, "
                                  "interpreted residual programs and jet maps, adjoint, weight "
                                  "gradients), k_ad_step_sum",
                     }.get(eng.step_kernel,
                           "train step: k_ad_fwd, k_ad_resid, k_ad_bwd, k_ad_wgrad, "
                           "k_ad_wreduce, k_adam (linear jet maps on FP32 MFMA)"),
                     "achieved": flops / (step_ms * 1e-3) / 1e12,
                     "peak": FP32_MFMA_PEAK / 1e12, "unit": "TFLOP/s",
                     "frac": flops / (step_ms * 1e-3) / FP32_MFMA_PEAK,
                     "traffic": (profiled_traffic('tdx_ad_step_jit_poisson')
                                 if eng.step_kernel == 'fused-generated' and args.grid_res == 200
                                 else None),
                     "traffic_source": "profiles/tdx_ad_step_jit_poisson_traffic.txt (rocprofv3 "
                                       "FETCH_SIZE/WRITE_SIZE, copy-calibrated; the step kernel "
                                       "of the 201x201 problem)",
                     "algorithmic_flops_per_launch": flops, "avg_launch_ms": step_ms},
        "forward_pass_ms": fwd_ms, "backward_pass_ms": bwd_ms,
    }
    if not args.no_plain:
        pnet = copy.deepcopy(net0).to(dev)
        pl = PlainAD(prob, pnet, dev)
        opt = torch.optim.Adam(pnet.parameters(), lr=lr)

        def pstep():
            opt.zero_grad()
            l = pl.loss()
            l.backward()
            opt.step()
        pstep()
        pms = time_cuda(pstep, max(args.plain_steps, 3))
        out["plain_gpu"] = {"value": 1e3 / pms, "unit": "steps/s", "ms_per_step": pms,
                            "what": "the reference's nested torch.autograd.grad chains in "
                                    "stock torch ops on the same GPU"}
        out["speedup_vs_plain_gpu"] = pms / ms
    if not args.no_cpu:
        from oracle import tedeous_oracle as O
        torch.set_num_threads(cpu_cores())
        orc = O.ADOracle(dict(dom.variable_dict), eq.equation_lst,
                         [dict(c) for c in conds.conditions_lst])
        cnet = copy.deepcopy(net0)
        opt = torch.optim.Adam(cnet.parameters(), lr=lr)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 20) or k < 1:
            opt.zero_grad()
            l = orc.evaluate(cnet, kw['lambda_operator'], kw['lambda_bound'])[0]
            l.backward()
            opt.step()
            k += 1
        cms = (time.perf_counter() - t0) * 1e3 / k
        out["cpu_baseline"] = {"value": 1e3 / cms, "unit": "steps/s", "cores": cpu_cores(),
                               "kind": "port",
                               "sample": f"{k} full Adam steps of the CPU oracle (nested autograd)"}
    return out


# ---------------------------------------------------------------------------
# natural-gradient workload
# ---------------------------------------------------------------------------

def bench_ngd(args):
    """NGD steps (tedeous/optimizers/ngd.py) of the Lotka-Volterra 'autograd'
    system through Model/Optimizer('NGD'): the Jacobian kernel, the Gram
    GEMMs, the eigensolver's pseudo-inverse and the candidate forwards."""
    import copy
    import torch_de_solver_amd as tdx
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.compiler import compile_ad
    from torch_de_solver_amd.optimizers import Closure
    from torch_de_solver_amd.plain import PlainAD, PlainNGD
    dom, eq, conds, net0, kw = W.lotka_ad(args.ngd_grid_res)
    dev = torch.device('cuda')
    gsn = 20
    net = copy.deepcopy(net0).to(dev)
    model = tdx.Model(net, dom, eq, conds)
    model.verbose = False
    model.compile('autograd', lambda_operator=kw['lambda_operator'],
                  lambda_bound=kw['lambda_bound'])
    opt = tdx.Optimizer('NGD', {'grid_steps_number': gsn})
    model.optimizer = opt.optimizer_choice('autograd', model.solution_cls.model)
    closure = Closure(False, model).get_closure('NGD')

    def one():
        model.optimizer.step(closure)
    one()
    one()
    for _ in range(args.warmup):
        one()
    ms = time_cuda(one, args.steps)
    loss_after = float(model.solution_cls.evaluate()[0])
    eng = model.solution_cls.engine
    flat = torch.cat([p.detach().reshape(-1) for p in model.solution_cls.model.parameters()])
    jac_ms = time_cuda(lambda: eng.jacobian(flat), args.steps)
    n_res, P = eng.tables.n_resid, eng.n_params
    out = {
        "metric": "NGD steps/s", "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"Lotka-Volterra 'autograd' {args.ngd_grid_res + 1} points, "
                               "MLP 1-32-32-2, NGD grid_steps_number 20",
                   "parameters": P, "residual_entries": n_res, "parallelism": "single GPU"},
        "jacobian_ms": jac_ms, "loss_after": loss_after,
    }
    if not args.no_plain:
        prob = compile_ad(dom.build('autograd'), eq.equation_lst,
                          conds.build(dom.variable_dict), 2,
                          lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
        pnet = copy.deepcopy(net0).to(dev)
        pl = PlainNGD(PlainAD(prob, pnet, dev), gsn)
        pl.step()
        pms = time_cuda(pl.step, max(args.plain_steps, 1))
        out["plain_gpu"] = {"value": 1e3 / pms, "unit": "steps/s", "ms_per_step": pms,
                            "what": "the reference's NGD step in stock torch ops on the same "
                                    "GPU (one autograd.grad per residual entry, SVD "
                                    "pseudo-inverse, a loss per candidate)"}
        out["speedup_vs_plain_gpu"] = pms / ms
    if not args.no_cpu:
        from oracle import tedeous_oracle as O
        torch.set_num_threads(cpu_cores())
        orc = O.ADOracle(dict(dom.variable_dict), eq.equation_lst,
                         [dict(c) for c in conds.conditions_lst])
        cnet = copy.deepcopy(net0)
        t0 = time.perf_counter()
        k = 0
        while time.perf_counter() - t0 < min(args.cpu_seconds, 20) or k < 1:
            O.ngd_step(cnet, orc, kw['lambda_operator'], kw['lambda_bound'], gsn)
            k += 1
        cms = (time.perf_counter() - t0) * 1e3 / k
        out["cpu_baseline"] = {"value": 1e3 / cms, "unit": "steps/s", "cores": cpu_cores(),
                               "kind": "port",
                               "sample": f"{k} NGD steps of the CPU oracle"}
    return out


def calibrate(args):
    dev = torch.device('cuda')
    n = 1 << 28                      # 2^28 float32 = 1 GiB
    a = torch.randn(n, device=dev)
    b = torch.empty_like(a)
    torch.neg(a, out=b)
    ms = time_cuda(lambda: torch.neg(a, out=b), args.steps)
    del a, b
    m = 8192
    x = torch.randn(m, m, device=dev)
    y = torch.randn(m, m, device=dev)
    z = torch.mm(x, y)
    gms = time_cuda(lambda: torch.mm(x, y, out=z), max(args.steps // 4, 3))
    return {"copy": {"bytes_read": 4 * n, "bytes_written": 4 * n, "ms": ms,
                     "GB_per_s": 8.0 * n / (ms * 1e-3) / 1e9, "peak_GB_per_s": HBM_PEAK / 1e9},
            "gemm_fp32": {"m": m, "ms": gms, "TFLOP_per_s": 2.0 * m ** 3 / (gms * 1e-3) / 1e12,
                          "peak_TFLOP_per_s": FP32_MFMA_PEAK / 1e12}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=20)
    ap.add_argument('--warmup', type=int, default=3)
    ap.add_argument('--workload', default='sweep',
                    choices=['sweep', 'burgers', 'kdv', 'legendre', 'wave', 'wave16k',
                             'poisson_ad', 'ngd_lotka'])
    ap.add_argument('--grid-res', type=int, default=200,
                    help="'poisson_ad': intervals per axis (points per axis minus one)")
    ap.add_argument('--ngd-grid-res', type=int, default=500,
                    help="'ngd_lotka': intervals in t (points minus one)")
    ap.add_argument('--plain-steps', type=int, default=2)
    ap.add_argument('--plain-instances', type=int, default=256)
    ap.add_argument('--cpu-seconds', type=float, default=20.0)
    ap.add_argument('--no-plain', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--fast-div', action='store_true')
    ap.add_argument('--acc-steps', type=int, default=500,
                    help="Adam steps of the accuracy comparison (0: skip it)")
    ap.add_argument('--acc-instances', type=int, default=8)
    ap.add_argument('--no-tune', action='store_true',
                    help="default tile shapes instead of the autotuner's choice")
    ap.add_argument('--tile-rows', type=int, default=0, help="'mat': output rows per tile")
    ap.add_argument('--band-rows', type=int, default=0,
                    help="'mat': rows per wavefront of the generated kernel (0: plan's choice)")
    ap.add_argument('--specialize', type=int, default=0,
                    help="'mat': kernel generated for the program on the interior (0 auto, "
                         "1 always, -1 never)")
    ap.add_argument('--persistent', type=int, default=0,
                    help="'NN': 0 fuses forward and adjoint into one kernel where the "
                         "network allows it (three launches per step); -1 keeps separate "
                         "forward and adjoint kernels (four launches); 1 runs the separate "
                         "forward as a persistent grid (three launches)")
    ap.add_argument('--chunk-tiles', type=int, default=0,
                    help="'NN': work per backward item: 64-row tiles (separate kernels) or "
                         "64-entry blocks (fused kernel) (0: autotuned)")
    ap.add_argument('--nn-tile-rows', type=int, default=0,
                    help="'NN': backward tile rows (32: two tiles per wavefront iteration)")
    ap.add_argument('--calibrate', action='store_true',
                    help="stream-copy and FP32-GEMM microbenchmarks only")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(--gpus must be 1)", file=sys.stderr)
        sys.exit(2)
    if not torch.cuda.is_available():
        print("bench.py needs a HIP device", file=sys.stderr)
        sys.exit(2)
    if args.calibrate:
        out = calibrate(args)
    elif args.workload in ('wave', 'wave16k'):
        out = bench_mat(args)
    elif args.workload == 'poisson_ad':
        out = bench_ad(args)
    elif args.workload == 'ngd_lotka':
        out = bench_ngd(args)
    else:
        out = bench_nn(args)
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
