"""Tile-shape autotuner for the native plans.

The reference has no tunable launch shapes (it runs stock torch ops).  Here
two knobs decide how work maps onto the GPU:

* 'NN' plans: ``chunk_tiles``, the work of one gradient workgroup — 64-row
  tiles for the separate adjoint kernel, 64-entry blocks for the fused
  forward-and-adjoint kernel (more: fewer gradient partials to reduce;
  fewer: a finer, better balanced grid over uneven instances);
* 'mat' plans: ``tile_rows``, the output rows of one stencil tile (more rows:
  less halo recomputed per point; fewer rows: less LDS per workgroup, so more
  workgroups in flight per CU).

``tune_nn`` / ``tune_mat`` build one plan per candidate, time the dominant
launches with HIP events after a warm-up, keep the fastest, and cache the
choice in a JSON file keyed by the plan's shape (``TDX_TUNE_CACHE``, default
``~/.cache/torch_de_solver_amd/tuning.json``; ``TDX_TUNE_CACHE=off`` keeps
nothing), so a later plan of the same shape skips the search.
"""

import json
import os
import threading
from typing import Dict, List, Optional, Sequence, Tuple

import torch

NN_CHUNK_CANDIDATES = (1, 2, 4, 8, 16, 32, 64, 128, 256)
MAT_ROW_CANDIDATES = (4, 8, 12, 16, 20, 28, 36, 44, 60)

_lock = threading.Lock()


def _cache_path() -> Optional[str]:
    p = os.environ.get('TDX_TUNE_CACHE')
    if p == 'off':
        return None
    if p:
        return p
    return os.path.join(os.path.expanduser('~'), '.cache', 'torch_de_solver_amd', 'tuning.json')


def _load() -> Dict[str, dict]:
    path = _cache_path()
    if not path or not os.path.exists(path):
        return {}
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def _store(key: str, value: dict) -> None:
    path = _cache_path()
    if not path:
        return
    with _lock:
        data = _load()
        data[key] = value
        try:
            os.makedirs(os.path.dirname(path), exist_ok=True)
            tmp = path + f'.{os.getpid()}.tmp'
            with open(tmp, 'w') as f:
                json.dump(data, f, indent=1, sort_keys=True)
            os.replace(tmp, path)
        except OSError:
            pass


def _time_ms(fn, reps: int) -> float:
    fn()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        fn()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) / reps


def nn_key(problems, arch, dtype) -> str:
    rows = []
    for p in problems:
        for g in p.groups:
            rows.append(g.n_entries * len(g.slot_base))
    return (f"nn|{arch.d_in}-{arch.width}x{arch.n_layers - 1}-{arch.n_out}|{dtype}|"
            f"inst{len(problems)}|rows{sum(rows)}|groups{len(rows)}")


def tune_nn(problems, arch, device, params: torch.Tensor,
            candidates: Sequence[int] = NN_CHUNK_CANDIDATES, reps: int = 3,
            use_cache: bool = True, tile_rows: int = 0,
            persistent: int = 0) -> Tuple[int, dict]:
    """Pick ``chunk_tiles`` for an 'NN' plan over ``problems`` (compiled) by
    timing the training step (``tdx_nn_train_step`` at learning rate 0, on a
    copy of ``params``: the fused kernel where the plan has it, else the
    separate forward and adjoint kernels).  Returns (chunk_tiles, report)."""
    from .nn_engine import NNEngine
    key = nn_key(problems, arch, params.dtype) + f"|tr{tile_rows}|ps{persistent}|step"
    if use_cache:
        hit = _load().get(key)
        if hit and hit.get('chunk_tiles') in candidates:
            return int(hit['chunk_tiles']), dict(hit, cached=True)
    loss = torch.empty(2 * len(problems), dtype=params.dtype, device=params.device)
    work = params.detach().clone()
    m, v = torch.zeros_like(work), torch.zeros_like(work)
    step = torch.zeros(1, dtype=torch.int64, device=params.device)
    times = {}
    for ct in candidates:
        eng = NNEngine(problems, arch, device, tile_rows, ct, persistent)

        def run():
            eng.train_step(work, m, v, step, 0.0, 0.9, 0.999, 1e-8, loss)
        times[ct] = _time_ms(run, reps)
        del eng
    best = min(times, key=times.get)
    rep = {'chunk_tiles': best, 'ms': {str(k): v for k, v in times.items()}}
    _store(key, rep)
    return best, dict(rep, cached=False)


def mat_key(prob, exact_div: bool) -> str:
    shape = 'x'.join(str(int(v)) for v in prob.shape)
    return f"mat|{shape}|eq{prob.n_eq}|dp{prob.dp}|{prob.dtype}|div{int(exact_div)}|{_mat_program_sig(prob)}"


def _mat_program_sig(prob) -> str:
    sig = []
    for eq in prob.equations:
        for t in eq:
            sig.append('.'.join(''.join(str(a) for a in f.axes) or 'u' for f in t.factors))
    return '+'.join(sig)


def tune_mat(prob, device, u: torch.Tensor, exact_div: bool = True,
             candidates: Sequence[int] = MAT_ROW_CANDIDATES, reps: int = 3,
             use_cache: bool = True) -> Tuple[int, dict]:
    """Pick ``tile_rows`` for a 'mat' plan by timing its residual + adjoint
    launch on ``u``.  Returns (tile_rows, report); 1-D grids have no choice."""
    from .mat_engine import MatEngine
    if len(prob.shape) < 2 or int(prob.shape[0]) == 1:
        return 0, {'tile_rows': 0, 'ms': {}, 'cached': False}
    key = mat_key(prob, exact_div)
    if use_cache:
        hit = _load().get(key)
        if hit and hit.get('tile_rows') in candidates:
            return int(hit['tile_rows']), dict(hit, cached=True)
    one = torch.ones(1, dtype=u.dtype, device=u.device)
    grad = torch.empty_like(u)
    times = {}
    for tr in candidates:
        eng = MatEngine(prob, device, exact_div=exact_div, tile_rows=tr)
        times[tr] = _time_ms(lambda: eng.backward(u, one, grad), reps)
        del eng
    best = min(times, key=times.get)
    rep = {'tile_rows': best, 'ms': {str(k): v for k, v in times.items()}}
    _store(key, rep)
    return best, dict(rep, cached=False)
