"""CPU tests of the host side: the library loads and exports every symbol of
include/tdx.h, the compiler's tables against closed forms and against the
oracle's own preparation, and plan validation (error codes, no launches)."""
import ctypes
import os
import re

import numpy as np
import pytest
import torch

import problems
from oracle import tedeous_oracle as O
from torch_de_solver_amd import data as api
from torch_de_solver_amd import native
from torch_de_solver_amd.compiler import compile_mat, compile_nn
from torch_de_solver_amd.finite_diffs import mat_alphas, nn_scheme
from torch_de_solver_amd.points_type import point_types, types_by_hull

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    text = open(os.path.join(ROOT, 'include', 'tdx.h')).read()
    return sorted(set(re.findall(r'\b(tdx_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_every_declared_symbol():
    lib = native.lib()
    syms = header_symbols()
    assert syms, "no declarations found"
    for s in syms:
        assert hasattr(lib, s), f"missing export {s}"
    assert set(syms) == set(native.EXPORTS)
    assert lib.tdx_version() == 3


def test_plan_validation_rejects_bad_tables():
    """Malformed tables give an error code and a message, never a launch."""
    lib = native.lib()
    d = native.NNDesc()
    d.dtype, d.d_in, d.n_out, d.width, d.n_layers, d.n_inst = 0, 2, 1, 16, 3, 1
    groups = np.array([0, 10, 1, 0, 0, 1, 1, 0], dtype=np.int32)  # slot table ok
    ints = np.array([999999], dtype=np.int32)                    # coordinates out of range
    reals = np.zeros(4, dtype=np.float32)
    data = np.zeros(4, dtype=np.float32)
    d.n_groups = 1
    d.groups = groups.ctypes.data_as(native.c_i32p)
    d.n_ints, d.ints = ints.size, ints.ctypes.data_as(native.c_i32p)
    d.n_reals, d.reals = reals.size, reals.ctypes.data
    d.n_data, d.data = data.size, data.ctypes.data
    h = ctypes.c_void_p()
    rc = lib.tdx_nn_plan_create(ctypes.byref(d), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b'out of range' in lib.tdx_last_error()
    d.width = 1000
    assert lib.tdx_nn_plan_create(ctypes.byref(d), ctypes.byref(h)) == -1
    assert lib.tdx_nn_plan_create(None, ctypes.byref(h)) == -1


@pytest.mark.parametrize('term,ptype,label', [
    ([0], 'central', '1'), ([0, 0], 'central', '1'), ([0, 0, 0], 'central', '1'),
    ([0, 1], 'central', '1'), ([1], 'fb', '1'), ([0], 'fb', '1'), ([0, 1], 'bf', '2'),
    ([1, 1], 'ff', '2'), ([0], 'bb', '2'),
])
def test_nn_scheme_matches_oracle(term, ptype, label):
    h = 1e-3
    mine = nn_scheme(term, 2, ptype, label, h)
    sh, sg = O.scheme(term, 2, ptype, label, h)
    assert [tuple(s) for s in sh] == [s for s, _ in mine]
    assert [w for _, w in mine] == list(sg)   # same double arithmetic: identical


def test_inner_order_2_is_rejected():
    with pytest.raises(ValueError):
        nn_scheme([0], 1, 'central', '2', 1e-3)


@pytest.mark.parametrize('dp', [2, 3, 5])
def test_mat_alphas_match_oracle(dp):
    ab, af, back, forw = mat_alphas(dp)
    b2, f2, ab2, af2 = O.mat_alphas(dp)
    np.testing.assert_allclose(ab, ab2, rtol=1e-14, atol=1e-15)
    np.testing.assert_allclose(af, af2, rtol=1e-14, atol=1e-15)


def test_point_types_closed_form_and_hull():
    x = torch.linspace(0, 1, 7)
    t = torch.linspace(0, 2, 5)
    g = torch.cartesian_prod(x, t).numpy()
    box = point_types(g)
    hull = types_by_hull(g)
    assert list(box) == list(hull)
    central = np.nonzero(box == 'central')[0]
    assert len(central) == 5 * 3                 # strict interior
    ref = O.point_types(torch.from_numpy(g))
    assert list(box) == ref
    # 1-D: every point is central
    assert set(point_types(np.linspace(0, 1, 9).reshape(-1, 1))) == {'central'}


@pytest.mark.parametrize('name', list(problems.NN_CASES))
def test_nn_tables_reproduce_oracle_shifted_grids(name):
    """Every (factor, stencil entry) of the compiled interior group points at
    the slot whose input equals the oracle's materialised shifted grid."""
    builder, layers = problems.NN_CASES[name]
    d, eq, b, kw = builder(api)
    grid = d.build('NN')
    prob = compile_nn(grid, eq.equation_lst, b.build(d.variable_dict), 1, h=kw.get('h', 1e-3),
                      lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    orc = O.NNOracle(dict(d.variable_dict), eq.equation_lst, [dict(c) for c in b.conditions_lst],
                     h=kw.get('h', 1e-3))
    g0 = prob.groups[0]
    base = torch.from_numpy(g0.bases[0])
    assert torch.equal(base, orc.central)
    op = orc.ops[0]
    for t_mine, term in zip(g0.equations[0].terms, op.values()):
        key = list(term.keys())[1]
        for f_mine, grids, signs in zip(t_mine.factors, term[key][0], term[key][1]):
            assert len(f_mine.stencil) == len(grids)
            for (slot, w), x_ref, w_ref in zip(f_mine.stencil, grids, signs):
                off = torch.tensor(g0.slot_offset[slot], dtype=torch.float64)
                x = base.clone()
                for a in range(x.shape[1]):
                    x[:, a] = x[:, a] + float(off[a])
                assert torch.equal(x, x_ref)
                assert np.float32(w) == np.float32(w_ref)


def test_bc_columns_padding_and_targets(golden):
    """Column order, lengths and padded targets equal the reference's."""
    for name in ('kdv_nn', 'legendre_nn_opbc', 'burgers_nn'):
        g = golden(f'{name}_float32')
        builder, layers = problems.NN_CASES[name]
        d, eq, b, kw = builder(api)
        prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                          h=kw.get('h', 1e-3), lambda_bound=kw['lambda_bound'])
        assert prob.bc_keys == list(g['keys'])
        assert prob.bc_lengths == list(g['lens'])
        true = np.zeros_like(g['true_bval'])
        for grp in prob.groups[1:]:
            for eq_ in grp.equations:
                true[eq_.col_pos, eq_.column] = eq_.target
        np.testing.assert_array_equal(true, g['true_bval'])


def test_unsupported_features_raise():
    d, eq, b, kw = problems.burgers_nn(api)
    grid = d.build('NN')
    bc = b.build(d.variable_dict)
    bad = [dict(eq.equation_lst[0])]
    bad[0] = {k: dict(v) for k, v in bad[0].items()}
    # a callable power that is not a chain of one-argument operations; any
    # callable power in 'mat' mode
    bad[0]['du/dt']['pow'] = lambda x: x * torch.sin(x)
    with pytest.raises(NotImplementedError):
        compile_nn(grid, bad, bc, 1)
    bad[0]['du/dt']['pow'] = lambda x: x
    compile_nn(grid, bad, bc, 1)
    dm, eqm, bm, kwm = problems.wave_mat(api)
    badm = [{k: dict(v) for k, v in eqm.equation_lst[0].items()}]
    key = next(k for k, v in badm[0].items() if 'pow' in v)
    badm[0][key]['pow'] = lambda x: x
    with pytest.raises(NotImplementedError, match="'NN' and 'autograd'"):
        compile_mat(dm.build('mat'), badm, bm.build(dm.variable_dict), 1)
    with pytest.raises(NotImplementedError):
        compile_nn(grid, eq.equation_lst, bc + [{'type': 'robin', 'bnd': bc[0]['bnd'],
                                                 'bop': None, 'bval': bc[0]['bval'], 'var': 0}], 1)


def test_mini_batching_limits():
    """'mat' mini-batching and batching with a causal loss raise before any
    device work; a batch larger than the grid turns batching off."""
    import torch_de_solver_amd as tdx
    d, eq, b, kw = problems.wave_mat(api, n=15)
    u = problems.mat_init([1, 16, 16])
    m = tdx.Model(u, d, eq, b, batch_size=1)
    with pytest.raises(NotImplementedError):
        m.compile('mat', lambda_operator=1, lambda_bound=100)
    d, eq, b, kw = problems.burgers_nn(api)
    net = problems.make_mlp([2, 8, 1])
    m = tdx.Model(net, d, eq, b, batch_size=16)
    with pytest.raises(NotImplementedError):
        m.compile('NN', lambda_operator=1, lambda_bound=10, tol=0.1)


def test_native_forward_requires_device_tensors():
    """The product path has no CPU fallback: without a HIP device the
    engines refuse to build."""
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from torch_de_solver_amd.nn_engine import NNEngine, mlp_arch
    d, eq, b, kw = problems.burgers_nn(api)
    net = problems.make_mlp([2, 12, 12, 12, 1])
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1)
    with pytest.raises(native.NativeUnavailable):
        NNEngine([prob], mlp_arch(net), torch.device('cpu'))


def test_quotient_step_check():
    """The 'mat' plans' exact division: the one-step correction of the
    reciprocal product is the correctly rounded quotient for every float32
    significand for the grid steps of the configs (wave 1/127, Legendre
    2/99, 1e-3, 0.02); the reciprocal product alone is not, so the check
    can fail; bad arguments give -1."""
    lib = native.lib()
    for h in (np.float32(1.0 / 127), np.float32(2.0 / 99), np.float32(1e-3), np.float32(0.02)):
        assert lib.tdx_quotient_check(float(h), 0) == 0, h
    assert lib.tdx_quotient_check(float(np.float32(1.0 / 127)), 1) > 0
    assert lib.tdx_quotient_check(0.0, 0) == -1
    assert lib.tdx_quotient_check(1.0, 7) == -1


def test_autotune_cache_roundtrip(tmp_path, monkeypatch):
    """Tuning choices are stored per plan shape in the JSON cache and read
    back; TDX_TUNE_CACHE=off keeps nothing."""
    from torch_de_solver_amd import autotune as A
    path = tmp_path / 'tuning.json'
    monkeypatch.setenv('TDX_TUNE_CACHE', str(path))
    A._store('nn|key', {'chunk_tiles': 32, 'ms': {'32': 1.0}})
    A._store('mat|key', {'tile_rows': 16, 'ms': {'16': 2.0}})
    data = A._load()
    assert data['nn|key']['chunk_tiles'] == 32 and data['mat|key']['tile_rows'] == 16
    monkeypatch.setenv('TDX_TUNE_CACHE', 'off')
    assert A._load() == {}
    A._store('x', {'tile_rows': 4})
    assert not (tmp_path / 'x').exists()


def _spec_tables(name, dtype=np.float32):
    from torch_de_solver_amd.mat_engine import MatTables
    from torch_de_solver_amd.solution import Operator_bcond_preproc, _operators
    builder, nx, nt, dp_, _ = problems.SPECIALIZED_CASES[name]
    d, eq, b, dp, n_eq = builder(api, nx, nt, dp_)
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    ops = _operators(pre.operator)
    prob = compile_mat(grid, ops, pre.bconds, n_eq=len(ops), derivative_points=dp,
                       lambda_operator=1, lambda_bound=10)
    return MatTables(prob, dtype)


@pytest.mark.parametrize('name', list(problems.SPECIALIZED_CASES))
def test_specialized_kernel_generates_and_compiles(name):
    """The kernel a float 2-D 'mat' plan generates for its program compiles
    for gfx950 (hipRTC, no device), for every division mode; each variable
    gets one gradient store and each equation one residual store per row."""
    t = _spec_tables(name)
    src = t.specialized_source(1, compile=True)
    assert src
    n_eq = t.n_eq
    adj = src[src.index('void tdx_mat_jit_adj'):src.index('void tdx_mat_jit_fwd')]
    steps = adj.count('const int gr = t - ')
    assert steps >= 1
    assert adj.count('a.grad[') == steps * n_eq
    assert adj.count('a.op[') == steps * n_eq
    for mode in (0, 2):
        s2 = t.specialized_source(mode)
        assert s2 and s2 != src


def test_specialized_kernel_not_for_1d_or_double():
    t = _spec_tables('wave_dp2', np.float64)
    assert t.specialized_source(1) == ''
    assert 'float 2-D' in native.lib().tdx_last_error().decode()
    d, eq, b, kw = problems.legendre_mat(api)
    from torch_de_solver_amd.mat_engine import MatTables
    from torch_de_solver_amd.solution import Operator_bcond_preproc, _operators
    grid = d.build('mat')
    pre = Operator_bcond_preproc(grid, eq.equation_lst, b.build(d.variable_dict)).set_strategy('mat')
    ops = _operators(pre.operator)
    prob = compile_mat(grid, ops, pre.bconds, n_eq=1, derivative_points=2,
                       lambda_operator=1, lambda_bound=10)
    assert MatTables(prob, np.float32).specialized_source(1) == ''


def _sweep_tables(idx, dtype=np.float32):
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNTables, mlp_arch
    insts = W.sweep_instances(256)
    probs, nets = [], []
    for i in idx:
        d, eq, b, net, ckw = W.sweep_problem(insts[i])
        probs.append(compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1,
                                h=ckw['h'], lambda_operator=ckw['lambda_operator'],
                                lambda_bound=ckw['lambda_bound']))
        nets.append(net)
    return NNTables(probs, mlp_arch(nets[0]), dtype), mlp_arch(nets[0])


def test_fused_step_kernel_generates_and_compiles():
    """The fused 'NN' kernel with residual programs generated from a grouped
    sweep plan's tables compiles for gfx950 (hipRTC, no device).  Instances
    that differ only in their coefficient (nu) and grid share generated code:
    one program for the Burgers operator (7 slots, every slot seeded), one for
    the Dirichlet conditions (1 slot, a target array)."""
    t, arch = _sweep_tables([0, 100, 255])
    src = t.fused_source(arch, compile=True)
    assert src
    gen = src[src.index('struct GenResid'):]
    cases = [c for c in gen.split('    case ')[1:]]
    assert len(cases) == 2
    op = next(c for c in cases if 'sdv[6]' in c)
    assert all(f'sdv[{k}]' in op for k in range(7)) and 'sdv[7]' not in op
    bc = next(c for c in cases if c is not op)
    assert 'sdv[0]' in bc and 'sdv[1]' not in bc and 'data[(int64_t)I[' in bc
    # fixed positions of the group's own tables, no loop over records
    assert 'for (' not in op.replace('for (int s = 0; s < FUSE_MAX_SLOTS', '')


def test_fused_step_kernel_not_for_double_or_many_slots():
    t, arch = _sweep_tables([0], np.float64)
    assert t.fused_source(arch) == ''
    assert 'fused step' in native.lib().tdx_last_error().decode()
    # a narrow network with nine shifted copies per point (mixed derivative)
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNTables, mlp_arch
    d = api.Domain()
    d.variable('x', [0, 1], 9)
    d.variable('t', [0, 1], 7)
    b = api.Conditions()
    b.dirichlet({'x': [0, 1], 't': 0}, value=0.0)
    eq = api.Equation()
    eq.add({'d2u/dx2': {'coeff': 1, 'term': [0, 0], 'pow': 1},
            'd2u/dxdt': {'coeff': 0.5, 'term': [0, 1], 'pow': 1},
            'd2u/dt2': {'coeff': 1, 'term': [1, 1], 'pow': 1}})
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1, h=0.001)
    net = problems.make_mlp([2, 16, 16, 1])
    assert NNTables([prob], mlp_arch(net), np.float32).fused_source(mlp_arch(net)) == ''


def test_wide_step_kernel_generates_and_compiles():
    """Wide float networks (here 2-100x4-1 on Burgers, two instances): the
    block-owning kernel that forms its own seeds, with generated residual
    programs, compiles for gfx950 (hipRTC, no device)."""
    from torch_de_solver_amd.compiler import compile_nn
    from torch_de_solver_amd.nn_engine import NNTables, mlp_arch
    d, eq, b, kw = problems.burgers_nn(api)
    prob = compile_nn(d.build('NN'), eq.equation_lst, b.build(d.variable_dict), 1, h=0.001)
    net = problems.make_mlp([2, 100, 100, 100, 100, 1])
    src = NNTables([prob, prob], mlp_arch(net), np.float32).fused_source(mlp_arch(net),
                                                                        compile=True)
    assert 'tdx_block_jit' in src and 'block_body<float, 7, 3, true' in src
    gen = src[src.index('struct GenBlockResid'):]
    assert gen.count('    case ') == 2  # the operator and the Dirichlet conditions


def test_ad_step_kernel_generates_and_compiles():
    """The fused 'autograd' step kernel for the Poisson workload (jets (),
    x, y, xx, yy; 2-100x5-1) with generated jet maps and residual programs
    compiles for gfx950 (hipRTC, no device); the generated jet maps hold the
    Faa di Bruno terms of the second derivatives; float64 plans and plans
    with more than 16 rows per entry are left out."""
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.ad_engine import ad_step_source
    from torch_de_solver_amd.compiler import compile_ad
    from torch_de_solver_amd.nn_engine import mlp_arch
    dom, eq, conds, net, kw = W.poisson_ad(20)
    prob = compile_ad(dom.build('autograd'), eq.equation_lst, conds.build(dom.variable_dict), 1,
                      lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    src = ad_step_source(prob, mlp_arch(net), compile=True)
    assert 'tdx_ad_step_jit' in src and 'ad_step_body<7, 4, tdx::GenJets' in src
    jets = src[src.index('struct GenJets'):src.index('struct GenADResid')]
    assert 'float v = ((s.s2 * z1) * z1);' in jets and 'v = v + (s.s1 * z3);' in jets
    gen = src[src.index('struct GenADResid'):]
    assert gen.count('    case ') == len(prob.groups)
    d, eq, b, kw = problems.heat3_ad(api, dtype='float32')
    prob3 = compile_ad(d.build('autograd'), eq.equation_lst, b.build(d.variable_dict), 1,
                       lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    assert ad_step_source(prob3, mlp_arch(problems.make_mlp([3, 20, 20, 1])))
    # callable powers: the generated residual function calls chain_term
    d, eq, b, kw = problems.sin_source_ad(api, dtype='float32')
    probc = compile_ad(d.build('autograd'), eq.equation_lst, b.build(d.variable_dict), 1,
                       lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    srcc = ad_step_source(probc, mlp_arch(problems.make_mlp([2, 16, 16, 1])), compile=True)
    genc = srcc[srcc.index('struct GenADResid'):]
    assert genc.count('chain_term(a.ints') == 6  # three terms: value and seeds
    d, eq, b, kw = problems.wave_ad(api, dtype='float64')
    prob64 = compile_ad(d.build('autograd'), eq.equation_lst, b.build(d.variable_dict), 1,
                        lambda_operator=kw['lambda_operator'], lambda_bound=kw['lambda_bound'])
    assert ad_step_source(prob64, mlp_arch(problems.make_mlp([2, 12, 12, 1]))) == ''
    assert 'fused step' in native.lib().tdx_last_error().decode()


def _ad_desc(name='wave_ad'):
    from torch_de_solver_amd.ad_engine import ADTables
    from torch_de_solver_amd.compiler import compile_ad
    builder, layers = problems.AD_CASES[name]
    d, eq, b, kw = builder(api)
    prob = compile_ad(d.build('autograd'), eq.equation_lst, b.build(d.variable_dict),
                      n_out=layers[-1], lambda_operator=kw['lambda_operator'],
                      lambda_bound=kw['lambda_bound'])
    t = ADTables(prob, np.float64)
    desc = native.ADDesc()
    desc.dtype, desc.d_in, desc.n_out, desc.width, desc.n_layers = 1, layers[0], layers[-1], \
        layers[1], len(layers) - 1
    desc.n_jets, desc.jet_i, desc.pptr_i, desc.part_i = t.n_jets, t.jet_i, t.pptr_i, t.part_i
    desc.n_points, desc.points = t.n_points, t.points.ctypes.data
    desc.n_groups, desc.groups_i = len(prob.groups), t.groups_i
    desc.n_ints, desc.ints = t.ints.size, t.ints.ctypes.data_as(native.c_i32p)
    desc.n_reals, desc.reals = t.reals.size, t.reals.ctypes.data_as(native.c_dp)
    desc.n_data, desc.data = t.data.size, t.data.ctypes.data
    desc.n_resid = t.n_resid
    return prob, t, desc


def test_ad_jets_and_partitions():
    """Jets are closed under sub-multisets and sorted by order; each jet's
    partitions are its set partitions (Bell numbers 1, 2, 5, 15) with blocks
    that multiply back to the jet."""
    import itertools
    for name in problems.AD_CASES:
        prob, t, _ = _ad_desc(name)
        jets = prob.jets
        assert jets[0] == ()
        assert [len(j) for j in jets] == sorted(len(j) for j in jets)
        for j in jets:
            for r in range(len(j)):
                for sub in itertools.combinations(j, r):
                    assert tuple(sorted(sub)) in jets
        for j, parts in zip(jets, prob.partitions()):
            assert len(parts) == [1, 1, 2, 5, 15][len(j)] if j else parts == []
            for blocks in parts:
                merged = tuple(sorted(a for b in blocks for a in jets[b]))
                assert merged == j
    prob, _, _ = _ad_desc('beam_ad')
    assert prob.jets[-1] == (0, 0, 0, 0)


def test_ad_plan_validation_rejects_bad_tables():
    """The 'autograd' plan checks every table index before any device work."""
    lib = native.lib()
    prob, t, desc = _ad_desc()
    h = ctypes.c_void_p()
    keep = t.ints.copy()
    # a stencil entry pointing past the jet list
    bad = keep.copy()
    g0 = bad[t.groups_i: t.groups_i + 5]
    eq0 = bad[g0[4]: g0[4] + 6]
    term0 = bad[eq0[1]: eq0[1] + 4]
    fac0 = bad[term0[1]: term0[1] + 5]
    bad[fac0[1] + 1] = t.n_jets
    desc.ints = bad.ctypes.data_as(native.c_i32p)
    assert lib.tdx_ad_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1 and not h.value
    assert b'stencil entry' in lib.tdx_last_error()
    # a point claimed by two entries
    bad = keep.copy()
    g1 = t.groups_i + 5
    bad[bad[g1 + 2]] = 0
    desc.ints = bad.ctypes.data_as(native.c_i32p)
    assert lib.tdx_ad_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert b'two entries' in lib.tdx_last_error()
    desc.ints = keep.ctypes.data_as(native.c_i32p)
    desc.width = 300
    assert lib.tdx_ad_plan_create(ctypes.byref(desc), ctypes.byref(h)) == -1
    assert lib.tdx_ad_plan_create(None, ctypes.byref(h)) == -1


@pytest.mark.parametrize('name', ['lotka_weak_nn', 'burgers_weak_nn'])
def test_weak_tables_reproduce_oracle_integration(name):
    """The compiled weak-form weights give the oracle's weak residual
    (tedeous/eval.py:13-52, 195-221) for random residuals on the problem's
    'central' points."""
    from torch_de_solver_amd.compiler import weak_tables
    builder, _ = problems.WEAK_CASES[name]
    d, eq, b, kw, weak = builder(api)
    grid = d.build('NN')
    types = point_types(grid.numpy())
    central = grid[torch.from_numpy(np.nonzero(types == 'central')[0])]
    lv, ctf2, cx, gstart = weak_tables(central, weak)
    rng = np.random.default_rng(7)
    op = torch.from_numpy(rng.standard_normal((len(central), 1)))
    want = float(O.weak_operator(op, central, weak))
    r2 = op[:, 0].numpy() ** 2
    runs = np.array([np.sum(ctf2[a:b_] * r2[a:b_]) for a, b_ in zip(gstart[:-1], gstart[1:])])
    got = float(runs[0]) if lv == 1 else float(np.sum(cx * runs ** 2))
    np.testing.assert_allclose(got, want, rtol=1e-12)
    assert lv == (1 if central.shape[1] == 1 else 2)


def test_poisson_autograd_flop_model():
    """The algorithmic FLOPs behind the 'autograd' roofline: per (point, jet)
    row, forward 2 din dout per layer, the transposed maps above the first
    layer, and 2 (din + 1) dout for the weight gradients."""
    import types
    import bench
    from torch_de_solver_amd import workloads as W
    from torch_de_solver_amd.ad_engine import ADTables
    from torch_de_solver_amd.compiler import compile_ad
    from torch_de_solver_amd.nn_engine import mlp_arch
    dom, eq, conds, net, kw = W.poisson_ad(200)
    prob = compile_ad(dom.build('autograd'), eq.equation_lst, conds.build(dom.variable_dict), 1)
    eng = types.SimpleNamespace(arch=mlp_arch(net), tables=ADTables(prob, np.float32))
    rows = (201 * 201 + 4 * 201) * 5            # grid + boundary points, jets (), x, y, xx, yy
    dims = [(2, 100)] + [(100, 100)] * 4 + [(100, 1)]
    fwd = sum(2 * a * b for a, b in dims)
    bwd = sum(2 * a * b for a, b in dims[1:])
    wg = sum(2 * (a + 1) * b for a, b in dims)
    assert eng.tables.n_points * eng.tables.n_jets == rows
    assert bench.ad_flops_per_step(eng) == rows * (fwd + bwd + wg) == 49940872050


def test_lambda_values_compare_by_value():
    """A lambda given as a number, a list or the reference's [1, n] tensor is
    the same weight; a changed value is a change (the plan is rebuilt)."""
    from torch_de_solver_amd.solution import _lambda_key
    assert _lambda_key(10) == _lambda_key(10.0) == _lambda_key([10]) == \
        _lambda_key(torch.tensor([[10.0]], dtype=torch.float64))
    assert _lambda_key([1, 2]) == _lambda_key(torch.tensor([[1.0, 2.0]]))
    assert _lambda_key([1, 2]) != _lambda_key([1, 3])
    assert _lambda_key(100) != _lambda_key(10)


def test_no_conditions_raise_like_the_reference():
    """Without boundary conditions the reference fails building its boundary
    matrix (max() of an empty sequence, tedeous/eval.py:75) in every mode."""
    from torch_de_solver_amd.solution import Operator_bcond_preproc, Solution
    d, eq, b, kw = problems.burgers_nn(api)
    grid = d.build('NN')
    for mode in ('NN', 'mat', 'autograd'):
        pre = Operator_bcond_preproc(grid, eq.equation_lst, []).set_strategy(mode)
        with pytest.raises(ValueError):
            Solution(grid, pre, problems.make_mlp([2, 8, 1]), mode, None, 1, 10)


def test_parse_operator_one_factor_variable_list():
    """A one-factor term written with 'var': [k] (the reference's
    Lotka-Volterra NGD example) addresses output k, as the reference's list
    index net(x)[:, [k]] does; a two-factor term keeps one index per
    factor."""
    from torch_de_solver_amd.equations import parse_operator
    d, eq, b, kw = problems.lotka_ad(api)
    t0 = parse_operator(eq.equation_lst[0])
    t1 = parse_operator(eq.equation_lst[1])
    assert [f.var for f in t0[0].factors] == [0] and t0[0].factors[0].axes == [0]
    assert [f.var for f in t1[0].factors] == [1] and t1[0].factors[0].axes == [0]
    assert [f.var for f in t1[1].factors] == [1] and t1[1].factors[0].axes is None
    assert [f.var for f in t0[2].factors] == [0, 1]
    with pytest.raises(ValueError):
        parse_operator({'bad': {'coeff': 1, 'term': [0], 'pow': 1, 'var': [0, 1]}})
