"""Randomized parity sweep (needs an MI355X): seeded random networks of mixed
sizes, domain sizes (2 to 200 values, across word boundaries), tightness,
pre-removed values and positive tables of arity 3 to 6, each checked against
the oracle in every GAC mode, every SAC mode, and under the three MAC
orderings (node sequences with domain hashes)."""
import numpy as np
import pytest

import oracle
from cpim_amd import GModelAdapter, GModelOptions, workload
from cpim_amd import _native as N
from cpim_amd.model import from_tuples

pytestmark = pytest.mark.gpu

GAC_MODES = {"plain": N.CPIM_GAC_PLAIN, "persistent": N.CPIM_GAC_PERSISTENT, "block": N.CPIM_GAC_BLOCK,
             "async": N.CPIM_GAC_ASYNC}
CASES = 40


def random_case(k):
    rng = np.random.default_rng(1000 + k)
    n = int(rng.integers(5, 80))
    d = int(rng.choice([2, 3, 5, 8, 16, 33, 64, 65, 100, 128, 200]))
    e = int(rng.integers(1, max(2, min(n * (n - 1) // 2, 4 * n))))
    p = float(rng.uniform(0.05, 0.9))
    pre = float(rng.choice([0.0, 0.0, 0.2]))
    seed = int(rng.integers(1, 1 << 30))
    desc = workload.random_binary(n, d, e, p, seed, preremove=pre)
    if rng.integers(0, 3) == 0 and n >= 4 and d <= 16:
        scopes = [list(s) for s in desc.scopes()]
        lists = list(desc.tuple_lists())
        for t in range(int(rng.integers(1, 4))):
            ar = int(rng.integers(3, min(6, n) + 1))
            scopes.append(sorted(rng.choice(n, ar, replace=False).tolist()))
            lists.append(workload.random_table(seed + t, ar, d, int(min(d ** ar, rng.integers(1, 5000)))))
        return from_tuples(desc.dom_size, scopes, lists).with_init_dom(desc.init_dom), True
    return desc, False


@pytest.mark.parametrize("k", range(CASES))
def test_random_network(gpu, k):
    desc, tables = random_case(k)
    algo = oracle.AC3 if tables else oracle.AC3BIT
    opts = GModelOptions(skip_non_binary=False)
    ok, dom, _ = oracle.Oracle(desc, algo).root_ac()
    for name, code in GAC_MODES.items():
        with GModelAdapter.Build(desc, opts) as m:
            st = m._gac(-1, code)
            assert st.inconsistent == (not ok), name
            if ok:
                assert (m.ReadDomains(0) == dom).all(), name
    sok, sdom, _ = oracle.Oracle(desc, algo).sac1()
    for mode in ["auto", "workgroup"] + ([] if tables else ["sliced"]):
        with GModelAdapter.Build(desc, opts) as m:
            st = m.EnforceSAC(mode=mode)
            assert st.consistent == sok, mode
            if sok:
                assert (m.ReadDomains(0) == sdom).all(), mode
    for varh, name in ((0, "dom"), (1, "dom/deg"), (2, "dom/wdeg")):
        with oracle.var_heuristic(varh):
            r = oracle.Oracle(desc, algo).mac(node_limit=1500, trace_cap=3000)
        with GModelAdapter.Build(desc, opts) as m:
            st, sol, trace = m.Solve(node_limit=1500, trace_capacity=3000, heuristic=name)
        assert (st.num_solutions, st.num_positive, st.num_negative) == \
            (r.num_sol, r.num_positive, r.num_negative), name
        assert trace == r.trace, name
