"""Batched instances and the serving layer vs the oracle (needs an MI355X).

Uneven batch: every instance's root fixpoint domains and its first 64 MAC
nodes (decision, value, outcome and domain hash of each node) must equal the
oracle's MAC run on that instance alone, whether all instances share one
launch or run one launch each.  Serving: every request's result (status,
node counts, solution) is the same as solving it alone, whatever else is
resident, and the same as the oracle's.
"""
import numpy as np
import pytest

import oracle
from cpim_amd import Batch, GModelAdapter, GModelOptions, Server, workload

pytestmark = pytest.mark.gpu

NODES = 64


def oracle_status(r, node_limit):
    if r.num_sol:
        return 1
    return 0 if r.time_out else 2


@pytest.fixture(scope="module")
def uneven():
    # 40 instances of the 4096-instance uneven batch (n in [20, 200], d in {8..128})
    idx = list(range(0, 4096, 103))
    models = [workload.uneven_instance(i) for i in idx]
    want = []
    for m in models:
        o = oracle.Oracle(m)
        ok, root, _ = o.root_ac()
        r = oracle.Oracle(m).mac(node_limit=NODES, trace_cap=2 * NODES)
        want.append((ok, root, r))
    return models, want


@pytest.mark.parametrize("persistent", [True, False], ids=["one-launch", "per-instance"])
def test_uneven_batch_first_64_nodes(gpu, uneven, persistent):
    models, want = uneven
    b = Batch(models)
    try:
        res, ms = b.run(NODES, persistent=persistent, trace_capacity=2 * NODES)
        assert ms > 0
        for i, (m, (ok, root, r), got) in enumerate(zip(models, want, res)):
            assert got.root_consistent == ok, i
            if ok:
                assert (b.root_domains(i) == root).all(), i
            assert (got.num_positive, got.num_negative) == (r.num_positive, r.num_negative), i
            assert got.status == oracle_status(r, NODES), i
            assert got.trace == r.trace, i
    finally:
        b.close()


def test_uneven_batch_rerun_is_identical(gpu, uneven):
    models, _ = uneven
    b = Batch(models[:12])
    try:
        first, _ = b.run(NODES, trace_capacity=2 * NODES)
        again, _ = b.run(NODES, trace_capacity=2 * NODES)
        assert [(x.status, x.num_positive, x.num_negative, x.trace) for x in first] == \
               [(x.status, x.num_positive, x.num_negative, x.trace) for x in again]
    finally:
        b.close()


def test_uneven_batch_heavy_units_launch_apart(gpu):
    # more units than CUs, four of them (the largest uneven instances, n ~ 200,
    # d = 128) far heavier than the mean work per workgroup slot: those run in
    # a 1024-thread launch of their own beside the others; every unit still
    # equals the oracle and the one-launch-per-unit path
    from cpim_amd import _native as N
    heavy_idx = [39, 3722, 572, 1817]
    small = [workload.random_binary(12, 4, 20, 0.3, 4000 + i) for i in range(300)]
    models = [workload.uneven_instance(i) for i in heavy_idx] + small
    before = N.tune_get("unit_threads")
    N.tune_set("unit_threads", 512)
    b = Batch(models)
    try:
        res, _ = b.run(NODES, trace_capacity=2 * NODES)
        assert b.heavy_units() == len(heavy_idx)
        for i in list(range(len(heavy_idx))) + list(range(len(heavy_idx), len(models), 15)):
            r = oracle.Oracle(models[i]).mac(node_limit=NODES, trace_cap=2 * NODES)
            got = res[i]
            assert (got.num_positive, got.num_negative) == (r.num_positive, r.num_negative), i
            assert got.status == oracle_status(r, NODES), i
            assert got.trace == r.trace, i
        per, _ = b.run(NODES, persistent=False, trace_capacity=2 * NODES)
        assert b.heavy_units() == 0
        assert [(x.status, x.num_positive, x.num_negative, x.trace) for x in res] == \
               [(x.status, x.num_positive, x.num_negative, x.trace) for x in per]
    finally:
        b.close()
        N.tune_set("unit_threads", before)


def solve_alone(desc, node_limit):
    with GModelAdapter.Build(desc, GModelOptions(skip_non_binary=False)) as m:
        st, sol, _ = m.Solve(node_limit=node_limit)
    status = 1 if st.num_solutions else (0 if st.time_out else 2)
    return status, st.num_positive, st.num_negative, (list(sol) if st.num_solutions else None)


@pytest.mark.parametrize("nodes_per_step,max_slots", [(64, 256), (7, 5)])
def test_serving_mix_equals_alone(gpu, nodes_per_step, max_slots):
    limit = 3000
    reqs = [workload.serving_request(i) for i in range(16)]
    srv = Server(max_slots=max_slots, nodes_per_step=nodes_per_step)
    try:
        ids = [srv.submit(r, node_limit=limit) for r in reqs]
        steps = 0
        finished = []
        while srv.step() > 0:
            finished += srv.poll(cap=3)
            steps += 1
            assert steps < 100000
        while True:
            more = srv.poll(cap=3)
            if not more:
                break
            finished += more
        assert sorted(finished) == sorted(ids)
        for rid, desc in zip(ids, reqs):
            got = srv.result(rid)
            assert got is not None
            r = oracle.Oracle(desc, oracle.AC3).mac(node_limit=limit)
            want = (oracle_status(r, limit), r.num_positive, r.num_negative,
                    list(r.solution) if r.num_sol else None)
            have = (got["status"], got["num_positive"], got["num_negative"],
                    list(got["solution"]) if got["solution"] is not None else None)
            assert have == want, rid
            assert solve_alone(desc, limit) == want, rid
    finally:
        srv.close()


def test_serving_result_busy_until_done(gpu):
    srv = Server(max_slots=2, nodes_per_step=1)
    try:
        rid = srv.submit(workload.serving_request(3), node_limit=500)
        assert srv.result(rid) is None
        assert srv.poll() == []
        finished = []
        while srv.step() > 0:
            finished += srv.poll()
        finished += srv.poll()
        assert finished == [rid]
        assert srv.result(rid) is not None
        assert srv.poll() == []
    finally:
        srv.close()


def test_batch_with_a_unit_in_global_scratch(gpu):
    # one unit too large for LDS (3,476 variables) beside small ones: every
    # unit's root fixpoint and first nodes equal the oracle's
    from test_gpu_edges import padded_pigeonhole
    models = [workload.uneven_instance(7), padded_pigeonhole(), workload.uneven_instance(9)]
    b = Batch(models)
    try:
        res, _ = b.run(NODES, trace_capacity=2 * NODES)
        for i, (m, got) in enumerate(zip(models, res)):
            ok, root, _ = oracle.Oracle(m).root_ac()
            r = oracle.Oracle(m).mac(node_limit=NODES, trace_cap=2 * NODES)
            assert got.root_consistent == ok, i
            if ok:
                assert (b.root_domains(i) == root).all(), i
            assert (got.num_positive, got.num_negative) == (r.num_positive, r.num_negative), i
            assert got.trace == r.trace, i
    finally:
        b.close()


def drain(srv, ids, watch=None):
    """Step until every request is done; watch(srv) runs after each step."""
    finished, steps = [], 0
    while srv.step() > 0:
        finished += srv.poll()
        if watch:
            watch(srv)
        steps += 1
        assert steps < 100000
    finished += srv.poll()
    assert sorted(finished) == sorted(ids)


def test_serving_hbm_budget_smaller_than_the_mix(gpu):
    # the budget holds about two of these requests at once: the rest wait in
    # the queue, the bytes charged never exceed the budget, and every result
    # still equals the oracle's and the request solved alone
    limit = 3000
    reqs = [workload.serving_request(i) for i in range(10)]
    probe = Server(max_slots=4, nodes_per_step=1)
    rid = probe.submit(reqs[0], node_limit=1)
    drain(probe, [rid])
    one = probe.result(rid)["footprint_bytes"]
    probe.close()
    budget = int(2.5 * one)
    srv = Server(max_slots=64, hbm_budget_bytes=budget, nodes_per_step=16)
    seen = {"queued": 0, "resident": 0}

    def watch(s):
        mem = s.memory()
        assert mem["resident_bytes"] <= budget and mem["peak_bytes"] <= budget
        seen["queued"] = max(seen["queued"], mem["queued"])
        seen["resident"] = max(seen["resident"], mem["resident"])

    try:
        ids = [srv.submit(r, node_limit=limit) for r in reqs]
        drain(srv, ids, watch)
        assert seen["queued"] > 0 and 1 <= seen["resident"] <= 3
        for rid, desc in zip(ids, reqs):
            got = srv.result(rid)
            r = oracle.Oracle(desc, oracle.AC3).mac(node_limit=limit)
            want = (oracle_status(r, limit), r.num_positive, r.num_negative,
                    list(r.solution) if r.num_sol else None)
            have = (got["status"], got["num_positive"], got["num_negative"],
                    list(got["solution"]) if got["solution"] is not None else None)
            assert have == want, rid
            assert got["footprint_bytes"] <= budget
            srv.release(rid)
            with pytest.raises(Exception, match="unknown request id"):
                srv.result(rid)
        assert srv.memory()["resident_bytes"] == 0
    finally:
        srv.close()


def test_serving_c3_shaped_requests_alone_and_in_the_mix(gpu):
    # requests of the benchmark SAC network's shape (n=500, d=256, e=20000,
    # t=0.5) plus four arity-3 tables of 500 tuples over 16 values, under a
    # budget of 64 nodes: in a mix whose budget holds two at once, each result
    # equals the request served alone and the oracle's (AC3 with bitset
    # supports for the binary constraints, tuple search for the tables)
    from concurrent.futures import ThreadPoolExecutor
    limit = 64
    reqs = [workload.serving_request_c3(i) for i in range(4)]

    def oracle_run(desc):
        r = oracle.Oracle(desc, oracle.AC3MIX).mac(node_limit=limit)
        return (oracle_status(r, limit), r.num_positive, r.num_negative)

    with ThreadPoolExecutor(max_workers=4) as ex:
        wants = list(ex.map(oracle_run, reqs))
    alone = []
    for desc in reqs:
        s1 = Server(max_slots=1, nodes_per_step=32)
        try:
            rid = s1.submit(desc, node_limit=limit)
            drain(s1, [rid])
            g = s1.result(rid)
            alone.append((g["status"], g["num_positive"], g["num_negative"]))
            one = g["footprint_bytes"]
        finally:
            s1.close()
    srv = Server(max_slots=8, hbm_budget_bytes=int(2.5 * one), nodes_per_step=16)
    try:
        ids = [srv.submit(r, node_limit=limit) for r in reqs]
        drain(srv, ids)
        for rid, want, a in zip(ids, wants, alone):
            g = srv.result(rid)
            assert (g["status"], g["num_positive"], g["num_negative"]) == want == a, rid
        assert srv.memory()["peak_bytes"] <= int(2.5 * one)
    finally:
        srv.close()
