"""Enumerating device MAC (find-all / max-solutions) against the oracle (needs an MI355X).

GModelSolver::SetFindAllSolutions / SetMaxSolutions (include/GModelSolver.h:
48-54) in the binary search: every solution found is counted and kept, then
treated as a failed node, until the tree is exhausted or max_solutions are
found.  The oracle's orc_mac_all does the same on the CPU; its enumeration is
pinned on the CPU against exhaustive enumeration (tests/test_oracle.py).
"""
import numpy as np
import pytest

import oracle
from conftest import load_golden
from cpim_amd import GModelAdapter, GModelSolver, workload

pytestmark = pytest.mark.gpu

HEUR = {"dom": 0, "dom/deg": 1, "dom/wdeg": 2}


def device_all(desc, k, cap=4096, trace=4096, heuristic="dom"):
    with GModelAdapter.Build(desc) as m:
        return m.SolveAll(k, cap, trace_capacity=trace, heuristic=heuristic)


@pytest.mark.parametrize("heuristic", sorted(HEUR))
@pytest.mark.parametrize("seed", range(6))
def test_find_all_matches_oracle(gpu, seed, heuristic):
    n, d, e, p = [(6, 3, 8, 0.3), (7, 3, 10, 0.35), (5, 4, 7, 0.4), (8, 2, 6, 0.2), (9, 3, 12, 0.3),
                  (10, 4, 20, 0.35)][seed]
    desc = workload.random_binary(n, d, e, p, 900 + seed)
    with oracle.var_heuristic(HEUR[heuristic]):
        want = oracle.mac_all(oracle.Oracle(desc), 10**6, sol_cap=4096, trace_cap=4096)
    st, sols, trace = device_all(desc, 10**6, heuristic=heuristic)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (want.num_sol, want.num_positive, want.num_negative)
    assert st.unsolvable == (want.num_sol == 0)
    assert trace == want.trace
    assert (sols == want.solutions).all()


@pytest.mark.parametrize("k", [1, 2, 7])
def test_max_solutions_matches_oracle(gpu, k):
    desc = workload.random_binary(9, 3, 10, 0.3, 77)
    want = oracle.mac_all(oracle.Oracle(desc), k, sol_cap=64, trace_cap=4096)
    st, sols, trace = device_all(desc, k, cap=64)
    assert st.num_solutions == want.num_sol == k
    assert (st.num_positive, st.num_negative) == (want.num_positive, want.num_negative)
    assert trace == want.trace and (sols == want.solutions).all()


def test_find_all_across_launches_and_capacity(gpu):
    # about 70,000 search nodes: several resumable launches of 16,384 nodes;
    # only the first 100 solutions are kept, all are counted
    desc = workload.random_binary(14, 3, 14, 0.25, 5)
    count, _ = oracle.bf_count_solutions(desc)
    want = oracle.mac_all(oracle.Oracle(desc), 10**7, sol_cap=100)
    assert want.num_sol == count and want.num_positive + want.num_negative > 3 * 16384
    st, sols, _ = device_all(desc, 10**7, cap=100, trace=0)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (want.num_sol, want.num_positive, want.num_negative)
    assert sols.shape == (100, 14) and (sols == want.solutions).all()


def test_solver_facade_find_all_queens4(gpu):
    desc, _ = load_golden("queens-4_ext")
    with GModelAdapter.Build(desc) as m:
        s = GModelSolver(m)
        st = s.Solve()
        assert st.num_solutions == 1 and s.GetSolution() == [1, 3, 0, 2]
        s.SetFindAllSolutions(True)
        st = s.Solve()
        assert st.num_solutions == 2 and not st.unsolvable
        assert s.GetAllSolutions() == [[1, 3, 0, 2], [2, 0, 3, 1]]
        s.SetMaxSolutions(1)
        st = s.Solve()
        assert st.num_solutions == 1 and s.GetAllSolutions() == [[1, 3, 0, 2]]


def test_solver_facade_max_solutions_needs_find_all(gpu):
    # as in the reference, SetMaxSolutions alone does not enumerate
    desc = workload.random_binary(9, 3, 10, 0.3, 77)
    with GModelAdapter.Build(desc) as m:
        s = GModelSolver(m)
        s.SetMaxSolutions(5)
        assert s.Solve().num_solutions == 1
        s.SetFindAllSolutions(True)
        s.SetMaxSolutions(5)
        st = s.Solve()
        want = oracle.mac_all(oracle.Oracle(desc), 5, sol_cap=5)
        assert st.num_solutions == 5 and np.array_equal(np.array(s.GetAllSolutions()), want.solutions)


def test_find_all_unsat(gpu):
    desc, _ = load_golden("composed-25-1-2-2_ext")
    st, sols, _ = device_all(desc, 10**6, trace=0)
    assert st.num_solutions == 0 and st.unsolvable and sols.shape[0] == 0


def test_solve_leaves_level0_at_the_root_closure(gpu):
    # a search that backtracks at the root rewrites its depth-0 trail; level 0
    # must hold the root closure again afterwards, so a second solve on the
    # same model repeats the first, node for node
    desc = workload.random_binary(9, 3, 10, 0.3, 77)
    ok, root, _ = oracle.Oracle(desc).root_ac()
    want = oracle.mac_all(oracle.Oracle(desc), 10**6, sol_cap=4096, trace_cap=4096)
    assert want.trace and any(r[0] == 0 for r in want.trace)  # negative decisions happen
    with GModelAdapter.Build(desc) as m:
        for _ in range(2):
            st, sols, trace = m.SolveAll(10**6, 4096, trace_capacity=4096)
            assert (st.num_solutions, st.num_positive, st.num_negative) == \
                (want.num_sol, want.num_positive, want.num_negative)
            assert trace == want.trace and (sols == want.solutions).all()
            assert (m.ReadDomains(0) == root).all()
            assert [m.GetDomainSize(v, 0) for v in range(desc.num_vars)] == \
                [bin(int(w)).count("1") for w in root[:, 0]]
