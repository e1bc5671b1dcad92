"""The device MAC with the network (adjacency, constraint metadata, support
planes and tables) copied into LDS with the value stride cut from 64 W to the
largest domain, and with the network read from device memory (tuning knob
mac_lds_net = 1): both give the oracle's node sequence, with domain hashes,
under every ordering.  The cases cover one-word domains, strides that are
not a multiple of 64 (65, 100, 200 values: the target-side revision, which
loads whole 64-value chunks of a row), positive tables of one and of several
4096-tuple chunks, and the batch kernel beside units without an image."""
import numpy as np
import pytest

import oracle
from cpim_amd import Batch, GModelAdapter, GModelOptions, workload
from cpim_amd import _native as N
from cpim_amd.model import from_tuples

pytestmark = pytest.mark.gpu


@pytest.fixture(params=[0, 1], ids=["lds-net", "hbm-net"])
def lds_net(request):
    old = N.tune_get("mac_lds_net")
    N.tune_set("mac_lds_net", request.param)
    yield request.param
    N.tune_set("mac_lds_net", old)


def with_tables(desc, seed, tables):
    rng = np.random.default_rng(seed)
    scopes = [list(s) for s in desc.scopes()]
    lists = list(desc.tuple_lists())
    n = desc.num_vars
    for t, (ar, count) in enumerate(tables):
        scopes.append(sorted(rng.choice(n, ar, replace=False).tolist()))
        d = int(desc.dom_size[scopes[-1][0]])
        lists.append(workload.random_table(seed + t, ar, d, count))
    return from_tuples(desc.dom_size, scopes, lists).with_init_dom(desc.init_dom)


def cases():
    # tightness near the phase transition of each shape: searches of 20 to
    # 600 nodes with failures, some SAT and some not
    yield "d16", workload.random_binary(30, 16, 110, 0.504, 3), False
    yield "d65", workload.random_binary(14, 65, 30, 0.815, 1), False
    yield "d100", workload.random_binary(12, 100, 20, 0.937, 1), False
    yield "d200", workload.random_binary(6, 200, 9, 0.971, 1), False
    yield "d16-tables", with_tables(workload.random_binary(24, 16, 70, 0.46, 15), 15,
                                    [(3, 400), (4, 900), (5, 1500)]), True
    yield "d20-chunked-table", with_tables(workload.random_binary(16, 20, 30, 0.76, 16), 16,
                                           [(3, 5000)]), True


CASES = list(cases())


@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
@pytest.mark.parametrize("varh,name", [(0, "dom"), (1, "dom/deg"), (2, "dom/wdeg")])
def test_mac_network_in_lds(gpu, lds_net, k, varh, name):
    _, desc, tables = CASES[k]
    algo = oracle.AC3 if tables else oracle.AC3BIT
    with oracle.var_heuristic(varh):
        r = oracle.Oracle(desc, algo).mac(node_limit=600, trace_cap=1200)
    assert r.num_negative > 0 or r.num_positive > 20
    with GModelAdapter.Build(desc, GModelOptions(skip_non_binary=False)) as m:
        st, sol, trace = m.Solve(node_limit=600, trace_capacity=1200, heuristic=name)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)
    assert trace == r.trace
    if r.num_sol:
        assert (np.asarray(sol) == r.solution).all()


def test_batch_units_with_and_without_image(gpu, lds_net):
    # small units (image in LDS) beside uneven instances whose image is too
    # large for it: every unit's first nodes equal the oracle's
    models = [CASES[0][1], workload.uneven_instance(5), CASES[2][1], workload.uneven_instance(3000),
              CASES[4][1]]
    b = Batch(models, GModelOptions(skip_non_binary=False))
    try:
        res, _ = b.run(48, trace_capacity=96)
        for i, (m, got) in enumerate(zip(models, res)):
            algo = oracle.AC3 if i == 4 else oracle.AC3BIT
            r = oracle.Oracle(m, algo).mac(node_limit=48, trace_cap=96)
            assert (got.num_positive, got.num_negative) == (r.num_positive, r.num_negative), i
            assert got.trace == r.trace, i
    finally:
        b.close()
