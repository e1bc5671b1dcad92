"""The autotuner on the engine (needs an MI355X).

Every candidate launch shape must give the same results (the tuner raises
TuneMismatch otherwise), the kept choice must be one of them, and results
under the kept choice must equal the oracle's.  A second request for the
same bucket is served from the cache without measuring.
"""
import numpy as np
import pytest

import oracle
from cpim_amd import Batch, GModelAdapter, autotune, workload
from cpim_amd import _native as N
from cpim_amd.model import domains_to_sets

pytestmark = pytest.mark.gpu

NODES = 64


def test_tune_sac_every_shape_agrees(gpu, tmp_path):
    path = str(tmp_path / "tune.json")
    desc = workload.random_binary(120, 64, 1500, 0.75, 31, preremove=0.1)  # 34 of the 64 sampled tests fail
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    tests = np.array([(v, a) for v, vals in enumerate(domains_to_sets(root, desc.dom_size)) for a in vals],
                     np.int32)
    pick = np.linspace(0, len(tests) - 1, 64).astype(np.int64)
    want, _ = o.singleton_tests(tests[pick], nthreads=8)
    assert 0 < want.sum() < len(pick)
    with GModelAdapter.Build(desc) as m:
        m.EnforceGAC(False)
        entry = autotune.tune_sac(m, tests, path=path)
        ran = [t for t in entry["timings"] if "seconds" in t]
        assert len(ran) >= 5 and entry["choice"] in autotune.SAC_SPACE
        before = {k: N.tune_get(k) for k in autotune.SAC_KNOBS}
        choice = autotune.tuned_sac(m, tests, path=path)  # from the cache
        assert choice == entry["choice"]
        try:
            fail, _ = m.SingletonTests(tests, mode="sliced")
            assert (fail[pick] == want).all()
        finally:
            autotune.apply(before)


def test_tune_batch_every_shape_agrees(gpu, tmp_path, monkeypatch):
    path = str(tmp_path / "tune.json")
    idx = list(range(0, 4096, 211))
    models = [workload.uneven_instance(i) for i in idx]
    b = Batch(models)
    try:
        entry = autotune.tune_batch(b, NODES, path=path, trace_capacity=2 * NODES)
        assert entry["choice"] in autotune.BATCH_SPACE
        assert len([t for t in entry["timings"] if "seconds" in t]) >= 4
        # a second request for the bucket does not measure
        monkeypatch.setattr(autotune, "_search", lambda *a, **k: pytest.fail("measured again"))
        before = {k: N.tune_get(k) for k in autotune.BATCH_KNOBS}
        assert autotune.tuned_batch(b, NODES, path=path) == entry["choice"]
        try:
            res, _ = b.run(NODES, trace_capacity=2 * NODES)
        finally:
            autotune.apply(before)
        for i in (0, len(models) // 2, len(models) - 1):
            r = oracle.Oracle(models[i]).mac(node_limit=NODES, trace_cap=2 * NODES)
            assert (res[i].num_positive, res[i].num_negative) == (r.num_positive, r.num_negative)
            assert res[i].trace == r.trace
    finally:
        b.close()
