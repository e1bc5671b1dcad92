"""Measured choice of the engine's launch shapes, cached per workload bucket.

The engine's built-in choices are analytic (masks per workgroup from the test
count, nibble tables from the LDS budget, waves per MAC unit from the unit
count).  This module times the alternatives on the caller's own workload,
checks that every alternative gives the same results (the choices change how
the work is laid out, never what is computed), keeps the fastest, and caches
it in a JSON file keyed by the workload's bucket, so that a later workload of
the same bucket gets the choice without being measured again.

Buckets:
  sac/W{words}/n{vars}/e{constraints}/t{tests}   bit-sliced singleton-test passes
  batch/W{words}/d{largest domain}/e{largest constraint count}/u{units}
                                                 MAC unit batches (uneven instances)
with every size rounded up to a power of two.

A choice is a set of tuning knobs (cpim_tune_set); apply() sets them.  The
cache file is $CPIM_TUNE_CACHE, else ~/.cache/cpim_amd/tune.json.
"""
from __future__ import annotations

import json
import os
import time
from typing import Callable, Dict, List, Optional, Sequence

import numpy as np

from . import _native as N

Choice = Dict[str, int]

# Candidates.  SAC: kernel 0 is the shared-stage kernel (masks 1, 2 or 4 per
# value; nibble tables on or off; with them, none, 4 or 6 byte tables in
# front), kernel 1 the one-stage-per-wave kernel.
SAC_SPACE: List[Choice] = (
    [{"sac_slice_kernel": 0, "sac_slice_masks": m, "sac_slice_nibbles": nb, "sac_slice_bytes": by}
     for m in (1, 2, 4) for nb, by in ((1, 1), (1, 4), (1, 6), (2, 1))] +
    [{"sac_slice_kernel": 1, "sac_slice_masks": 0, "sac_slice_nibbles": 0, "sac_slice_bytes": 0}])
# MAC unit batches: waves per unit and the network image in LDS.
BATCH_SPACE: List[Choice] = [{"unit_threads": t, "mac_lds_net": lds}
                             for t in (256, 512, 1024) for lds in (0, 1)]
SAC_KNOBS = ("sac_slice_kernel", "sac_slice_masks", "sac_slice_nibbles", "sac_slice_bytes")
BATCH_KNOBS = ("unit_threads", "mac_lds_net")


def cache_path() -> str:
    return os.environ.get("CPIM_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "cpim_amd", "tune.json")


def _pow2(x: int) -> int:
    p = 1
    while p < max(1, int(x)):
        p *= 2
    return p


def sac_bucket(num_vars: int, words: int, num_cons: int, num_tests: int) -> str:
    return f"sac/W{words}/n{_pow2(num_vars)}/e{_pow2(num_cons)}/t{_pow2(num_tests)}"


def batch_bucket(words: int, max_dom: int, max_cons: int, units: int) -> str:
    return f"batch/W{words}/d{_pow2(max_dom)}/e{_pow2(max_cons)}/u{_pow2(units)}"


def load_cache(path: Optional[str] = None) -> Dict[str, dict]:
    path = path or cache_path()
    try:
        with open(path) as f:
            data = json.load(f)
        return data if isinstance(data, dict) else {}
    except (OSError, ValueError):
        return {}


def store(bucket: str, entry: dict, path: Optional[str] = None) -> None:
    path = path or cache_path()
    data = load_cache(path)
    data[bucket] = entry
    os.makedirs(os.path.dirname(path) or ".", exist_ok=True)
    tmp = path + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def current(knobs: Sequence[str]) -> Choice:
    return {k: N.tune_get(k) for k in knobs}


def apply(choice: Choice) -> None:
    for k, v in choice.items():
        N.tune_set(k, int(v))


class TuneMismatch(RuntimeError):
    """Two candidate launch shapes gave different results: an engine defect."""


def _measure(run: Callable[[], object], reps: int) -> (float, object):
    out = run()  # warm-up (and the result compared across candidates)
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        run()
        best = min(best, time.perf_counter() - t)
    return best, out


def _search(bucket: str, space: List[Choice], knobs: Sequence[str], run: Callable[[], object],
            same: Callable[[object, object], bool], reps: int, use_cache: bool,
            path: Optional[str]) -> dict:
    saved = current(knobs)
    timings = []
    first = None
    try:
        for cand in space:
            apply({k: cand.get(k, 0) for k in knobs})
            try:
                secs, out = _measure(run, reps)
            except N.CpimError as e:  # a shape this workload cannot take (LDS)
                timings.append({"choice": cand, "error": str(e)})
                continue
            if first is None:
                first = out
            elif not same(first, out):
                raise TuneMismatch(f"{bucket}: results differ under {cand}")
            timings.append({"choice": cand, "seconds": secs})
    finally:
        apply(saved)
    ok = [t for t in timings if "seconds" in t]
    if not ok:
        raise RuntimeError(f"{bucket}: no candidate ran")
    best = min(ok, key=lambda t: t["seconds"])
    entry = {"choice": best["choice"], "seconds": best["seconds"], "timings": timings,
             "measured_at": time.time()}
    if use_cache:
        store(bucket, entry, path)
    return entry


def tune_sac(model, tests: np.ndarray, reps: int = 2, space: Optional[List[Choice]] = None,
             use_cache: bool = True, path: Optional[str] = None) -> dict:
    """Time every candidate on model.SingletonTests(tests, mode="sliced") and
    keep the fastest for the model's bucket.  Every candidate must give the
    same outcomes."""
    bucket = sac_bucket(model.num_vars, model.dom_words, model.num_constraints, len(tests))

    def run():
        fail, _ = model.SingletonTests(tests, mode="sliced")
        return fail

    entry = _search(bucket, space or SAC_SPACE, SAC_KNOBS, run,
                    lambda a, b: np.array_equal(a, b), reps, use_cache, path)
    entry["bucket"] = bucket
    return entry


def tuned_sac(model, tests: np.ndarray, path: Optional[str] = None) -> Choice:
    """The cached choice of the model's bucket, measured first when missing;
    applied before returning."""
    bucket = sac_bucket(model.num_vars, model.dom_words, model.num_constraints, len(tests))
    entry = load_cache(path).get(bucket) or tune_sac(model, tests, path=path)
    apply(entry["choice"])
    return entry["choice"]


def batch_key(models) -> str:
    """The tuning bucket of a batch of models."""
    words = max(m.words for m in models)
    return batch_bucket(words, max(m.max_dom for m in models), max(m.num_cons for m in models), len(models))


def tune_batch(batch, node_limit: int, reps: int = 1, space: Optional[List[Choice]] = None,
               use_cache: bool = True, path: Optional[str] = None, trace_capacity: int = 0) -> dict:
    """Time every candidate on batch.run(node_limit) and keep the fastest for
    the batch's bucket.  Every candidate must give the same outcome per unit
    (status, node counts and, with trace_capacity, the node sequence)."""
    bucket = batch_key(batch.models)

    def run():
        res, _ = batch.run(node_limit, trace_capacity=trace_capacity)
        return [(r.root_consistent, r.status, r.num_positive, r.num_negative, tuple(r.trace)) for r in res]

    entry = _search(bucket, space or BATCH_SPACE, BATCH_KNOBS, run, lambda a, b: a == b, reps,
                    use_cache, path)
    entry["bucket"] = bucket
    return entry


def tuned_batch(batch, node_limit: int, path: Optional[str] = None) -> Choice:
    bucket = batch_key(batch.models)
    entry = load_cache(path).get(bucket) or tune_batch(batch, node_limit, path=path)
    apply(entry["choice"])
    return entry["choice"]
