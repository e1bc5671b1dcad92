"""Autotuner host logic (CPU only): buckets, cache file, choice search.

The timed search runs against a stand-in workload here (its results and
times are scripted); on the GPU it runs on the engine (tests/test_gpu_autotune.py).
"""
import json

import numpy as np
import pytest

from cpim_amd import _native as N
from cpim_amd import autotune


def test_buckets_round_sizes_up_to_powers_of_two():
    assert autotune.sac_bucket(500, 4, 20000, 128000) == "sac/W4/n512/e32768/t131072"
    assert autotune.sac_bucket(1, 1, 0, 1) == "sac/W1/n1/e1/t1"
    assert autotune.batch_bucket(2, 128, 9788, 4096) == "batch/W2/d128/e16384/u4096"


def test_cache_round_trip(tmp_path):
    p = str(tmp_path / "sub" / "tune.json")
    assert autotune.load_cache(p) == {}
    autotune.store("sac/W1/n8/e8/t8", {"choice": {"sac_slice_masks": 2}}, p)
    autotune.store("batch/W1/d8/e8/u8", {"choice": {"unit_threads": 512}}, p)
    data = json.load(open(p))
    assert set(data) == {"sac/W1/n8/e8/t8", "batch/W1/d8/e8/u8"}
    (tmp_path / "bad.json").write_text("not json")
    assert autotune.load_cache(str(tmp_path / "bad.json")) == {}


def test_search_keeps_the_fastest_and_restores_knobs(tmp_path, monkeypatch):
    p = str(tmp_path / "tune.json")
    N.tune_set("unit_threads", 0)
    # scripted workload: 512 threads is fastest, 1024 cannot run
    cost = {256: 3.0, 512: 1.0, 1024: None}
    clock = [0.0]
    monkeypatch.setattr(autotune.time, "perf_counter", lambda: clock[0])

    def run():
        t = N.tune_get("unit_threads")
        if cost[t] is None:
            raise N.CpimError(N.CPIM_ERR_INVALID_ARGUMENT, "does not fit")
        clock[0] += cost[t]
        return "same result"

    space = [{"unit_threads": t, "mac_lds_net": 0} for t in (256, 512, 1024)]
    e = autotune._search("batch/x", space, autotune.BATCH_KNOBS, run, lambda a, b: a == b, 2, True, p)
    assert e["choice"]["unit_threads"] == 512
    assert any("error" in t for t in e["timings"])
    assert N.tune_get("unit_threads") == 0  # knobs restored
    assert autotune.load_cache(p)["batch/x"]["choice"]["unit_threads"] == 512


def test_search_rejects_differing_results(tmp_path):
    def run():
        return N.tune_get("unit_threads")  # a "result" that depends on the shape

    space = [{"unit_threads": t, "mac_lds_net": 0} for t in (256, 512)]
    with pytest.raises(autotune.TuneMismatch):
        autotune._search("batch/y", space, autotune.BATCH_KNOBS, run, lambda a, b: a == b, 1, False,
                         str(tmp_path / "t.json"))
    assert N.tune_get("unit_threads") == 0
