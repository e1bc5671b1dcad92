"""Edge cases on the device vs the oracle (needs an MI355X): empty and
degenerate networks, singleton and maximum-size domains, empty test lists,
and tests on values outside the current domain."""
import numpy as np
import pytest

import oracle
from cpim_amd import GModelAdapter, GModelOptions, workload
from cpim_amd import _native as N
from cpim_amd.model import domains_to_sets, from_tuples

pytestmark = pytest.mark.gpu

MODES = {"plain": N.CPIM_GAC_PLAIN, "block": N.CPIM_GAC_BLOCK, "async": N.CPIM_GAC_ASYNC}


def check_all(desc, node_limit=5000):
    ok, dom, _ = oracle.Oracle(desc).root_ac()
    for name, code in MODES.items():
        with GModelAdapter.Build(desc) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok), name
            if ok:
                assert (m.ReadDomains(0) == dom).all(), name
    sok, sdom, _ = oracle.Oracle(desc).sac1()
    for mode in ("auto", "workgroup", "sliced"):
        with GModelAdapter.Build(desc) as m:
            st = m.EnforceSAC(mode=mode)
            assert st.consistent == sok, mode
            if sok:
                assert (m.ReadDomains(0) == sdom).all(), mode
    r = oracle.Oracle(desc).mac(node_limit=node_limit, trace_cap=2 * node_limit)
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=node_limit, trace_capacity=2 * node_limit)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)
    assert trace == r.trace
    if r.num_sol:
        assert list(sol) == list(r.solution)


def test_no_constraints(gpu):
    desc = from_tuples([5, 1, 7], [], [])
    check_all(desc)


def test_single_variable(gpu):
    check_all(from_tuples([9], [], []))


def test_singleton_domains_chain(gpu):
    # x0 = x1 = x2 with singleton domains {0}: consistent, one solution
    eq = np.array([[0, 0]], np.int32)
    check_all(from_tuples([1, 1, 1], [(0, 1), (1, 2)], [eq, eq]))


def test_unsatisfiable_singletons(gpu):
    ne = np.zeros((0, 2), np.int32)
    check_all(from_tuples([1, 1], [(0, 1)], [ne]))


@pytest.mark.parametrize("d", [1024, 1000, 65, 64, 63, 2])
def test_domain_sizes(gpu, d):
    n = 12 if d > 256 else 20
    desc = workload.random_binary(n, d, 30, 0.9 if d > 256 else 0.5, 2100 + d, preremove=0.1)
    check_all(desc, node_limit=300)


def test_empty_and_out_of_domain_tests(gpu):
    desc = workload.random_binary(20, 40, 60, 0.55, 77, preremove=0.3)
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    sets = domains_to_sets(root, desc.dom_size)
    # values that the root fixpoint removed, tested anyway
    outside = [(v, a) for v in range(20) for a in range(40) if a not in sets[v]][:50]
    tests = np.array(outside, np.int32)
    want, _ = o.singleton_tests(tests)
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC(False)
        for mode in ("sliced", "workgroup", "plain"):
            got, st = m.SingletonTests(tests, mode=mode)
            assert (got == want).all(), mode
            empty, st = m.SingletonTests(np.zeros((0, 2), np.int32), mode=mode)
            assert len(empty) == 0 and st.singleton_tests == 0


def test_invalid_tests_rejected(gpu):
    desc = workload.random_binary(10, 8, 12, 0.3, 5)
    with GModelAdapter.Build(desc) as m:
        with pytest.raises(N.CpimInvalidArgument, match="unknown variable"):
            m.SingletonTests(np.array([[10, 0]], np.int32))
        with pytest.raises(N.CpimInvalidArgument, match="outside the domain"):
            m.SingletonTests(np.array([[0, 8]], np.int32))


def padded_pigeonhole(k=6, d=5, npad=3470):
    """Pigeonhole core (k variables pairwise different over d < k values,
    unsatisfiable, arc consistent) plus a chain of npad variables of size 8
    under universal relations: every ordering decides the core first, and the
    3,476 variables push the search state past the LDS budget."""
    ne = np.array([(a, b) for a in range(d) for b in range(d) if a != b], np.int32)
    univ = np.array([(a, b) for a in range(8) for b in range(8)], np.int32)
    scopes, lists = [], []
    for i in range(k):
        for j in range(i + 1, k):
            scopes.append((i, j))
            lists.append(ne)
    for v in range(k, k + npad - 1):
        scopes.append((v, v + 1))
        lists.append(univ)
    return from_tuples([d] * k + [8] * npad, scopes, lists)


@pytest.mark.parametrize("heuristic,varh", [("dom", 0), ("dom/deg", 1), ("dom/wdeg", 2)])
def test_mac_in_global_scratch(gpu, heuristic, varh):
    # the device MAC keeps domains, queues and stamps in global scratch here;
    # about 140 failures each, so dom/wdeg's weights change all along
    desc = padded_pigeonhole()
    with oracle.var_heuristic(varh):
        r = oracle.Oracle(desc).mac(node_limit=400, trace_cap=800)
    assert r.num_negative > 100 and not r.num_sol
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=400, trace_capacity=800, heuristic=heuristic)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)
    assert trace == r.trace


@pytest.mark.parametrize("heuristic,varh", [("dom", 0), ("dom/deg", 1), ("dom/wdeg", 2)])
def test_mac_stops_at_an_empty_unconstrained_domain(gpu, heuristic, varh):
    # 53 variables, 20 constraints, a fifth of the values removed: variable
    # 40 has no constraints and an empty initial domain, which no revision
    # reaches; the search ends without a solution when it selects it
    desc = workload.random_binary(53, 3, 20, 0.46, 922837970, preremove=0.2)
    with oracle.var_heuristic(varh):
        r = oracle.Oracle(desc, oracle.AC3BIT).mac(node_limit=1500, trace_cap=3000)
    assert r.num_sol == 0 and not r.time_out
    with GModelAdapter.Build(desc) as m:
        st, sol, trace = m.Solve(node_limit=1500, trace_capacity=3000, heuristic=heuristic)
    assert (st.num_solutions, st.num_positive, st.num_negative, st.time_out) == \
        (r.num_sol, r.num_positive, r.num_negative, 0)
    assert trace == r.trace
