"""Device engine vs the oracle (needs an MI355X).

Bit-exact parity: identical fixpoint domains, identical wipe-out detection,
identical MAC node counts, node sequences (with domain hashes) and solutions
under dom-min / min-value ordering, on the reference's fixtures (golden
files), on seeded random networks checked against the oracle at run time,
and at full benchmark sizes through size-independent properties.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from conftest import golden_names, load_golden
from cpim_amd import GModelAdapter, GModelOptions, GModelSolver, workload
from cpim_amd import _native as N
from cpim_amd.model import domains_to_sets, from_tuples

pytestmark = pytest.mark.gpu

GAC_MODES = {"plain": N.CPIM_GAC_PLAIN, "persistent": N.CPIM_GAC_PERSISTENT,
             "block": N.CPIM_GAC_BLOCK, "async": N.CPIM_GAC_ASYNC}


def popcount(dom):
    return int(sum(bin(int(x)).count("1") for x in dom.reshape(-1)))


def all_tests(dom, desc):
    return np.array([(v, a) for v, vals in enumerate(domains_to_sets(dom, desc.dom_size))
                     for a in vals], np.int32).reshape(-1, 2)


# ---------------------------------------------------------------------------
# Golden fixtures (the reference's own instances)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("mode", sorted(GAC_MODES))
@pytest.mark.parametrize("name", golden_names())
def test_golden_root_gac(gpu, name, mode):
    desc, g = load_golden(name)
    with GModelAdapter.Build(desc) as m:
        before = popcount(m.ReadDomains(0))
        st = m._gac(-1, GAC_MODES[mode])
        assert st.inconsistent == (not bool(g["root_ok"]))
        if g["root_ok"]:
            dom = m.ReadDomains(0)
            assert (dom == g["root_dom"]).all()
            assert st.deletions == before - popcount(dom)
            assert st.converged
            sizes = [m.GetDomainSize(v, 0) for v in range(desc.num_vars)]
            assert sizes == [len(s) for s in domains_to_sets(dom, desc.dom_size)]


SAC_MODES = ["auto", "workgroup", "sliced"]


@pytest.mark.parametrize("mode", SAC_MODES)
@pytest.mark.parametrize("name", golden_names())
def test_golden_sac(gpu, name, mode):
    desc, g = load_golden(name)
    with GModelAdapter.Build(desc) as m:
        st = m.EnforceSAC(mode=mode)
        assert st.consistent == bool(g["sac_ok"])
        if g["sac_ok"]:
            assert (m.ReadDomains(0) == g["sac_dom"]).all()


@pytest.mark.parametrize("name", ["test", "queens-4_ext", "queens-12_ext",
                                  "composed-25-1-2-2_ext"])
def test_golden_sac_plain_path(gpu, name):
    desc, g = load_golden(name)
    with GModelAdapter.Build(desc) as m:
        st = m.EnforceSAC(batched=False)
        assert st.consistent == bool(g["sac_ok"])
        if g["sac_ok"]:
            assert (m.ReadDomains(0) == g["sac_dom"]).all()


@pytest.mark.parametrize("name", golden_names())
def test_golden_mac(gpu, name):
    desc, g = load_golden(name)
    cap = len(g["mac_trace"])
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=int(g["node_limit"]), trace_capacity=cap)
        assert (st.num_solutions, st.num_positive, st.num_negative, int(st.time_out)) == \
            (int(g["mac_num_sol"]), int(g["mac_pos"]), int(g["mac_neg"]), int(g["mac_timeout"]))
        assert np.array(trace, dtype=np.uint64).reshape(-1, 5).tolist() == g["mac_trace"].tolist()
        if st.num_solutions:
            assert list(sol) == list(g["mac_solution"])


def test_known_answers_on_device(gpu):
    desc, _ = load_golden("test")
    with GModelAdapter.Build(desc) as m:
        s = GModelSolver(m)
        st = s.Solve()
        assert s.GetSolution() == [0, 0, 1] and (st.num_positive, st.num_negative) == (3, 0)
    desc, _ = load_golden("queens-4_ext")
    with GModelAdapter.Build(desc) as m:
        s = GModelSolver(m)
        s.Solve()
        assert s.GetSolution() == [1, 3, 0, 2]
    desc, _ = load_golden("composed-25-1-2-2_ext")
    with GModelAdapter.Build(desc) as m:
        st = GModelSolver(m).Solve()
        assert st.num_solutions == 0 and not st.time_out


def test_dway_solver_reference_semantics(gpu):
    # GModelSolver's d-way search on queens-4: 2 positive nodes (V0=0 fails,
    # V0=1 solves), solution [1,3,0,2] (docs/GPU_PERFORMANCE_ANALYSIS.md:28,99)
    desc, _ = load_golden("queens-4_ext")
    with GModelAdapter.Build(desc) as m:
        s = GModelSolver(m, branching="dway")
        st = s.Solve()
        assert s.GetSolution() == [1, 3, 0, 2]
        assert (st.num_positive, st.num_negative) == (2, 1)
        assert m.GetCurrentLevel() == 0


# ---------------------------------------------------------------------------
# Level / trail API: reference return conventions (src/GModel.cu:1397-1701)
# ---------------------------------------------------------------------------
def test_level_api_conventions(gpu):
    desc, _ = load_golden("queens-4_ext")
    with GModelAdapter.Build(desc) as m:
        assert m.GetCurrentLevel() == 0 and m.max_depth == 5
        assert m.GetMinDomainVar(0) == 0
        assert m.GetFirstValue(1, 0) == 0 and m.GetNextValue(1, 0, 0) == 1
        assert m.GetNextValue(1, 3, 0) == -1
        L = m.CreateNewLevel()
        assert L == 1 and (m.ReadDomains(1) == m.ReadDomains(0)).all()
        assert m.AssignValue(0, 2, L)
        assert m.GetDomainSize(0, L) == 1 and m.IsAssigned(0, L) and m.GetAssignedValue(0, L) == 2
        assert m.GetAssignedValue(1, L) == -1
        assert not m.RemoveValue(0, 1, L)          # absent value
        assert m.RemoveValue(1, 3, L) and m.GetDomainSize(1, L) == 3
        assert not m.RemoveValue(0, 2, L)          # domain becomes empty
        st = m.EnforceGAC(False, 1)
        assert st.inconsistent                     # var 0 is empty
        m.BackToLevel(0)
        assert m.GetCurrentLevel() == 0 and m.GetDomainSize(0, 0) == 4
        with pytest.raises(RuntimeError, match="Cannot back to future level"):
            m.BackToLevel(1)
        with pytest.raises(RuntimeError, match="Invalid var"):
            m.GetDomainSize(7, 0)
        with pytest.raises(RuntimeError, match="Invalid level"):
            m.GetDomainSize(0, 9)
        for _ in range(4):
            m.CreateNewLevel()
        with pytest.raises(RuntimeError, match="Max depth reached"):
            m.CreateNewLevel()
        assert not m.IsFullyAssigned(0)


def test_incremental_gac_from_assignment(gpu):
    desc = workload.random_binary(40, 12, 160, 0.38, 31)
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC(False)
        assert (m.ReadDomains(0) == root).all()
        for mode in GAC_MODES.values():
            L = m.CreateNewLevel()
            m.AssignValue(3, int(domains_to_sets(root, desc.dom_size)[3][0]), L)
            st = m._gac(3, mode)
            init = root.copy()
            a = domains_to_sets(root, desc.dom_size)[3][0]
            init[3] = 0
            init[3, a // 64] = np.uint64(1) << np.uint64(a % 64)
            bok, bdom = oracle.bf_gac(desc.with_init_dom(init))
            assert st.inconsistent == (not bok)
            if bok:
                assert (m.ReadDomains(L) == bdom).all()
            m.BackToLevel(0)


# ---------------------------------------------------------------------------
# Seeded random networks vs the oracle (single and multi-word domains)
# ---------------------------------------------------------------------------
RANDOM = [  # (n, d, e, p, preremove)
    (30, 16, 120, 0.40, 0.0), (60, 10, 400, 0.42, 0.0), (50, 64, 300, 0.25, 0.3),
    (40, 100, 150, 0.80, 0.0), (30, 130, 60, 0.93, 0.0), (25, 256, 50, 0.97, 0.2),
    (80, 20, 600, 0.55, 0.0),
]


@pytest.mark.parametrize("case", range(len(RANDOM)))
def test_random_gac_all_modes(gpu, case):
    n, d, e, p, pre = RANDOM[case]
    desc = workload.random_binary(n, d, e, p, 500 + case, preremove=pre)
    ok, dom, _ = oracle.Oracle(desc).root_ac()
    for mode, code in GAC_MODES.items():
        with GModelAdapter.Build(desc) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok), mode
            if ok:
                assert (m.ReadDomains(0) == dom).all(), mode


@pytest.mark.parametrize("case", range(len(RANDOM)))
def test_random_sac_and_singletons(gpu, case):
    n, d, e, p, pre = RANDOM[case]
    desc = workload.random_binary(n, d, e, p, 900 + case, preremove=pre)
    ok, dom, _ = oracle.Oracle(desc).sac1()
    with GModelAdapter.Build(desc) as m:
        st = m.EnforceSAC()
        assert st.consistent == ok
        if ok:
            assert (m.ReadDomains(0) == dom).all()
    o = oracle.Oracle(desc)
    rok, root, _ = o.root_ac()
    if rok:
        tests = all_tests(root, desc)
        want, _ = o.singleton_tests(tests, nthreads=4)
        with GModelAdapter.Build(desc) as m:
            m.EnforceGAC(False)
            for mode in SAC_MODES:
                got, _ = m.SingletonTests(tests, mode=mode)
                assert (got == want).all(), mode
            got2, _ = m.SingletonTests(tests[:64], batched=False)
            assert (got2 == want[:64]).all()
    for mode in ("workgroup", "sliced"):
        with GModelAdapter.Build(desc) as m:
            st = m.EnforceSAC(mode=mode)
            assert st.consistent == ok, mode
            if ok:
                assert (m.ReadDomains(0) == dom).all(), mode


# Arbitrary test lists for the bit-sliced kernel: unsorted, repeated tests,
# values outside the current domain (the oracle reduces D(x) to {a} all the
# same), several variables per group, one-word to sixteen-word domains.
SLICED = [  # (n, d, e, p, preremove, tests)
    (40, 7, 200, 0.30, 0.2, 500), (30, 64, 150, 0.75, 0.3, 900), (20, 65, 60, 0.85, 0.0, 700),
    (15, 1000, 30, 0.97, 0.1, 400), (60, 3, 300, 0.15, 0.0, 300),
]


# (kernel, masks, nibbles, bytes) knobs: default, each mask width with and
# without nibble tables, byte tables in front of them, and the
# one-stage-per-wave kernel
SLICE_VARIANTS = [(0, 0, 0, 0), (0, 1, 1, 0), (0, 2, 1, 0), (0, 4, 1, 0), (0, 1, 2, 0), (0, 2, 2, 0),
                  (0, 4, 2, 0), (0, 1, 1, 6), (0, 2, 1, 4), (0, 4, 1, 4), (0, 4, 1, 6), (1, 0, 0, 0)]
SLICE_KNOBS = ("sac_slice_kernel", "sac_slice_masks", "sac_slice_nibbles", "sac_slice_bytes")


def set_slice_knobs(values):
    for k, v in zip(SLICE_KNOBS, values):
        N.tune_set(k, v)


@pytest.fixture(params=SLICE_VARIANTS, ids=lambda v: f"kernel{v[0]}-masks{v[1]}-nib{v[2]}-bytes{v[3]}")
def slice_variant(request):
    set_slice_knobs(request.param)
    yield request.param
    set_slice_knobs((0, 0, 0, 0))


@pytest.mark.parametrize("name", golden_names())
def test_golden_sac_slice_variants(gpu, name, slice_variant):
    desc, g = load_golden(name)
    with GModelAdapter.Build(desc) as m:
        st = m.EnforceSAC(mode="sliced")
        assert st.consistent == bool(g["sac_ok"])
        if g["sac_ok"]:
            assert (m.ReadDomains(0) == g["sac_dom"]).all()


@pytest.mark.parametrize("case", range(len(SLICED)))
def test_sliced_arbitrary_test_lists(gpu, case, slice_variant):
    n, d, e, p, pre, k = SLICED[case]
    desc = workload.random_binary(n, d, e, p, 1700 + case, preremove=pre)
    o = oracle.Oracle(desc)
    rok, root, _ = o.root_ac()
    rng = np.random.default_rng(case)
    tests = np.stack([rng.integers(0, n, k), rng.integers(0, d, k)], 1).astype(np.int32)
    tests[k // 2:k // 2 + 20] = tests[:20]  # repeats
    want, _ = o.singleton_tests(tests, nthreads=4)
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC(False)
        if rok:
            assert (m.ReadDomains(0) == root).all()
        for mode in ("sliced", "workgroup"):
            got, st = m.SingletonTests(tests, mode=mode)
            assert st.singleton_tests == k
            assert (got == want).all(), mode
        # a list shorter than one group, and a single test
        got, _ = m.SingletonTests(tests[:5], mode="sliced")
        assert (got == want[:5]).all()
        got, _ = m.SingletonTests(tests[7:8], mode="sliced")
        assert (got == want[7:8]).all()


def test_sliced_long_runs_split_across_groups(gpu, slice_variant):
    # a sorted list (no permutation) whose variables have more tests than a
    # workgroup holds (64, 128 or 256 by the mask count), with repeated values
    desc = workload.random_binary(30, 64, 150, 0.75, 1701, preremove=0.3)
    o = oracle.Oracle(desc)
    rng = np.random.default_rng(99)
    runs = [(0, 300), (3, 40), (7, 700), (8, 1), (29, 257)]
    tests = np.concatenate([np.stack([np.full(k, x), rng.integers(0, 64, k)], 1) for x, k in runs]).astype(np.int32)
    want, _ = o.singleton_tests(tests, nthreads=4)
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC(False)
        got, st = m.SingletonTests(tests, mode="sliced")
        assert st.singleton_tests == len(tests)
        assert (got == want).all()


def test_sliced_rejects_table_networks(gpu):
    desc = ternary(1)
    with GModelAdapter.Build(desc, GModelOptions(skip_non_binary=False)) as m:
        with pytest.raises(N.CpimError, match="bit-sliced"):
            m.SingletonTests(np.array([[0, 0]], np.int32), mode="sliced")
        # auto falls back to one test per workgroup
        m.SingletonTests(np.array([[0, 0]], np.int32))


@pytest.mark.parametrize("case", range(len(RANDOM)))
def test_random_mac_traces(gpu, case):
    n, d, e, p, pre = RANDOM[case]
    desc = workload.random_binary(n, d, e, p, 1300 + case, preremove=pre)
    r = oracle.Oracle(desc).mac(node_limit=3000, trace_cap=6000)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=3000, trace_capacity=6000)
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, r.time_out)
    assert trace == r.trace
    if r.num_sol:
        assert list(sol) == list(r.solution)


@pytest.mark.parametrize("case", range(len(RANDOM)))
def test_random_mac_traces_dom_deg(gpu, case):
    # dom/deg ordering (MAC.cpp:371-384) node for node against the oracle
    n, d, e, p, pre = RANDOM[case]
    desc = workload.random_binary(n, d, e, p, 1300 + case, preremove=pre)
    with oracle.var_heuristic(1):
        r = oracle.Oracle(desc).mac(node_limit=3000, trace_cap=6000)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=3000, trace_capacity=6000, heuristic="dom/deg")
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, r.time_out)
    assert trace == r.trace
    if r.num_sol:
        assert list(sol) == list(r.solution)


@pytest.mark.parametrize("name", golden_names())
def test_golden_mac_dom_deg(gpu, name):
    desc, g = load_golden(name)
    limit = int(g["node_limit"]) or 3000
    with oracle.var_heuristic(1):
        r = oracle.Oracle(desc).mac(node_limit=limit, trace_cap=2 * limit + 16)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=limit, trace_capacity=2 * limit + 16, heuristic="dom/deg")
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, r.time_out)
    assert trace == r.trace


@pytest.mark.parametrize("case", range(len(RANDOM)))
def test_random_mac_traces_dom_wdeg(gpu, case):
    # dom/wdeg (MAC.cpp:386-411): fixpoints in AC3's order, weights raised on
    # each wipe-out; node for node and weight for weight against the oracle
    n, d, e, p, pre = RANDOM[case]
    desc = workload.random_binary(n, d, e, p, 1300 + case, preremove=pre)
    with oracle.var_heuristic(2):
        r = oracle.Oracle(desc).mac(node_limit=3000, trace_cap=6000)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=3000, trace_capacity=6000, heuristic="dom/wdeg")
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, r.time_out)
    assert trace == r.trace
    if r.num_sol:
        assert list(sol) == list(r.solution)


@pytest.mark.parametrize("name", golden_names())
def test_golden_mac_dom_wdeg(gpu, name):
    desc, g = load_golden(name)
    limit = int(g["node_limit"]) or 3000
    with oracle.var_heuristic(2):
        r = oracle.Oracle(desc).mac(node_limit=limit, trace_cap=2 * limit + 16)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=limit, trace_capacity=2 * limit + 16, heuristic="dom/wdeg")
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, r.time_out)
    assert trace == r.trace


def test_dom_wdeg_triangle_on_device(gpu):
    # 0 != 1, 1 != 2, 0 != 2 over {0, 1}: 0=0 fails, its refutation fails
    desc = from_tuples([2, 2, 2], [(0, 1), (1, 2), (0, 2)],
                       [[(a, b) for a in range(2) for b in range(2) if a != b]] * 3)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(trace_capacity=16, heuristic="dom/wdeg")
    assert [t[:4] for t in trace] == [(1, 0, 0, 0), (0, 0, 0, 0)]
    assert (st.num_solutions, st.num_positive, st.num_negative) == (0, 1, 1)


def test_dom_deg_isolated_variable_on_device(gpu):
    # a variable in no constraint has key -1 and is decided first
    dom = [4, 3, 3, 3, 5]
    scopes, lists = [], []
    for x, y in [(0, 1), (0, 2), (0, 3), (1, 2)]:
        scopes.append((x, y))
        lists.append([(a, b) for a in range(dom[x]) for b in range(dom[y]) if a != b])
    desc = from_tuples(dom, scopes, lists)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(trace_capacity=64, heuristic="dom/deg")
    assert [t[1:3] for t in trace] == [(4, 0), (0, 0), (1, 1), (2, 2), (3, 1)]
    assert st.num_solutions == 1


def test_unknown_heuristic_rejected(gpu):
    desc = workload.random_binary(6, 4, 6, 0.4, 1)
    with GModelAdapter.Build(desc) as m:
        with pytest.raises(ValueError):
            m.Solve(heuristic="wdeg")
        o = N.CSolveOpts()
        o.var_heuristic = 7
        st = N.CSearchStats()
        sol = np.zeros(6, np.int32)
        tl = C.c_int32(0)
        with pytest.raises(N.CpimError, match="heuristic"):
            N.check(N.gpu_lib().cpim_gmodel_solve(m.handle, C.byref(o), C.byref(st),
                                                  sol.ctypes.data_as(N.p_i32), None, C.byref(tl)))


def test_mac_across_launch_slices(gpu):
    # more nodes than one launch runs: the device state is resumed
    desc = workload.random_binary(23, 23, 253, 0.247, 4)
    r = oracle.Oracle(desc).mac(node_limit=40000)
    with GModelAdapter.Build(desc) as m:
        st, sol, _ = m.Solve(node_limit=40000)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)


# ---------------------------------------------------------------------------
# Positive tables (arity 3) when skip_non_binary is off
# ---------------------------------------------------------------------------
def ternary(seed):
    base = workload.random_binary(14, 6, 25, 0.3, seed)
    scopes = [list(s) for s in base.scopes()]
    lists = list(base.tuple_lists())
    rng = np.random.default_rng(seed)
    for k in range(5):
        scopes.append(sorted(rng.choice(14, 3, replace=False).tolist()))
        lists.append(workload.random_table(seed * 17 + k, 3, 6, 60))
    return from_tuples(base.dom_size, scopes, lists)


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_tables_gac_and_mac(gpu, seed):
    desc = ternary(seed)
    opts = GModelOptions(skip_non_binary=False)
    ok, dom, _ = oracle.Oracle(desc, oracle.AC3).root_ac()
    for mode, code in GAC_MODES.items():
        with GModelAdapter.Build(desc, opts) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok), mode
            if ok:
                assert (m.ReadDomains(0) == dom).all(), mode
    r = oracle.Oracle(desc, oracle.AC3).mac(node_limit=2000, trace_cap=4000)
    with GModelAdapter.Build(desc, opts) as m:
        st, sol, trace = m.Solve(node_limit=2000, trace_capacity=4000)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)
    assert trace == r.trace
    # dom/deg counts the neighbours through the tables as well; dom/wdeg
    # revises one table position at a time in AC3's order
    for varh, name in ((1, "dom/deg"), (2, "dom/wdeg")):
        with oracle.var_heuristic(varh):
            r = oracle.Oracle(desc, oracle.AC3).mac(node_limit=2000, trace_cap=4000)
        with GModelAdapter.Build(desc, opts) as m:
            st, sol, trace = m.Solve(node_limit=2000, trace_capacity=4000, heuristic=name)
        assert (st.num_solutions, st.num_positive, st.num_negative) == \
            (r.num_sol, r.num_positive, r.num_negative), name
        assert trace == r.trace, name


@pytest.mark.parametrize("seed", [1, 2, 3, 4])
def test_large_tables_gac_sac_mac(gpu, seed):
    # tables of arity 6-8 with 4500-5000 tuples: the revision walks 4096-tuple
    # chunks; every GAC mode, SAC, and the MAC under all three orderings
    from conftest import large_table_model
    desc = large_table_model(seed)
    opts = GModelOptions(skip_non_binary=False)
    ok, dom, _ = oracle.Oracle(desc, oracle.AC3).root_ac()
    for mode, code in GAC_MODES.items():
        with GModelAdapter.Build(desc, opts) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok), mode
            if ok:
                assert (m.ReadDomains(0) == dom).all(), mode
    sok, sdom, _ = oracle.Oracle(desc, oracle.AC3).sac1()
    with GModelAdapter.Build(desc, opts) as m:
        st = m.EnforceSAC()
        assert st.consistent == sok
        if sok:
            assert (m.ReadDomains(0) == sdom).all()
    for varh, name in ((0, "dom"), (1, "dom/deg"), (2, "dom/wdeg")):
        with oracle.var_heuristic(varh):
            r = oracle.Oracle(desc, oracle.AC3).mac(node_limit=2000, trace_cap=4000)
        with GModelAdapter.Build(desc, opts) as m:
            st, sol, trace = m.Solve(node_limit=2000, trace_capacity=4000, heuristic=name)
        assert (st.num_solutions, st.num_positive, st.num_negative) == \
            (r.num_sol, r.num_positive, r.num_negative), name
        assert trace == r.trace, name
        if r.num_sol:
            assert list(sol) == list(r.solution), name


def test_table_limits_reported(gpu):
    # arity above 8 is refused with a clear error
    from cpim_amd import workload
    dom = [3] * 9
    desc = from_tuples(dom, [list(range(9))], [workload.random_table(1, 9, 3, 50)])
    with pytest.raises(N.CpimError, match="arity above 8"):
        GModelAdapter.Build(desc, GModelOptions(skip_non_binary=False))


# ---------------------------------------------------------------------------
# Full benchmark sizes: properties + sampled oracle checks
# ---------------------------------------------------------------------------
def test_single_large_full_size(gpu):
    desc = workload.single_large()
    ok, dom, _ = oracle.Oracle(desc).root_ac()
    results = {}
    for mode, code in GAC_MODES.items():
        with GModelAdapter.Build(desc) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok)
            results[mode] = m.ReadDomains(0)
            assert (results[mode] == dom).all(), mode
            # idempotence: a second fixpoint from every constraint changes nothing
            st2 = m._gac(-1, code)
            assert st2.deletions == 0 and (m.ReadDomains(0) == dom).all()


def test_sliced_pass_with_many_variables_per_group(gpu):
    # 5,000 variables of 16 values: 77,546 tests in about 300 groups of 16
    # variables each, placed on CUs by their summed degree; a sample of the
    # outcomes (a fifth of them fail) against the oracle
    desc = workload.random_binary(5000, 16, 60000, 0.65, 77)
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    tests = all_tests(root, desc)
    rng = np.random.default_rng(5)
    pick = np.sort(rng.choice(len(tests), 64, replace=False))
    want, _ = o.singleton_tests(tests[pick], nthreads=8)
    assert 0 < int(want.sum()) < 64
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC_Auto()
        allf, st = m.SingletonTests(tests)
        assert st.singleton_tests == len(tests)
        assert (allf[pick] == want).all()
        # the same tests in reverse order (unsorted input) give the same outcomes
        rev, _ = m.SingletonTests(tests[::-1].copy())
        assert (rev[::-1] == allf).all()


def test_batched_sac_full_size_sample(gpu):
    desc = workload.batched_sac()
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    tests = all_tests(root, desc)
    rng = np.random.default_rng(3)
    sample = tests[np.sort(rng.choice(len(tests), 96, replace=False))]
    want, _ = o.singleton_tests(sample, nthreads=8)
    with GModelAdapter.Build(desc) as m:
        st = m.EnforceGAC_Auto()
        assert not st.inconsistent and (m.ReadDomains(0) == root).all()
        for mode in SAC_MODES:
            got, sst = m.SingletonTests(sample, mode=mode)
            assert (got == want).all(), mode
        for variant in SLICE_VARIANTS:
            set_slice_knobs(variant)
            try:
                got, sst = m.SingletonTests(sample, mode="sliced")
            finally:
                set_slice_knobs((0, 0, 0, 0))
            assert (got == want).all(), variant
        # every singleton test of the full pass, in one launch
        allf, sst = m.SingletonTests(tests)
        assert sst.singleton_tests == len(tests)
        assert (allf[np.searchsorted(tests[:, 0] * 256 + tests[:, 1],
                                     sample[:, 0] * 256 + sample[:, 1])] == want).all()


# (n, d, e, p, seed): networks whose first dom/wdeg search finds a solution
# after some failures, and whose second search, started from the raised
# weights, takes another path
WDEG_AGAIN = [(30, 16, 120, 0.40, 1300), (30, 16, 120, 0.42, 5021), (30, 16, 120, 0.42, 5038)]


@pytest.mark.parametrize("case", range(len(WDEG_AGAIN)))
def test_dom_wdeg_weights_persist_across_solves(gpu, case):
    # the weights live in the model as Tabular::weight lives in the network
    # (Network.cpp:303, AC3.cpp:122-126): a second solve continues from the
    # weights the first one raised, as the oracle's second search on the same
    # network does; after a reset the first search repeats.  (Both searches
    # end at a solution: after an exhausted or node-limited search the CPU
    # network and the device model are left in different states, so a
    # second search there would not compare weights alone.)
    n, d, e, p, seed = WDEG_AGAIN[case]
    desc = workload.random_binary(n, d, e, p, seed)
    with oracle.var_heuristic(2):
        o = oracle.Oracle(desc)
        r1 = o.mac(node_limit=5000, trace_cap=10000)
        w1 = o.weights().copy()
        r2 = o.mac(node_limit=5000, trace_cap=10000)
        w2 = o.weights().copy()
    assert r1.num_sol and r2.num_sol
    assert (w1 > 1).any() and (w2 != w1).any() and r1.trace != r2.trace
    with GModelAdapter.Build(desc) as m:
        assert (m.Weights() == 1).all() and len(m.Weights()) == len(w1)
        st, sol, trace = m.Solve(node_limit=5000, trace_capacity=10000, heuristic="dom/wdeg")
        assert trace == r1.trace and (m.Weights() == w1).all()
        st, sol, trace = m.Solve(node_limit=5000, trace_capacity=10000, heuristic="dom/wdeg")
        assert (st.num_positive, st.num_negative) == (r2.num_positive, r2.num_negative)
        assert trace == r2.trace and (m.Weights() == w2).all()
        assert list(sol) == list(r2.solution)
        assert (m.Weights(reset=True) == w2).all()
        assert (m.Weights() == 1).all()
        st, sol, trace = m.Solve(node_limit=5000, trace_capacity=10000, heuristic="dom/wdeg")
        assert trace == r1.trace and (m.Weights() == w1).all()
