"""Write tests/golden/cascade/c3_cascade.npz: oracle outcomes of sampled
singleton tests of the C3-shaped cascade network (n=500, d=256, e=20000,
t=0.86, workload.batched_sac_cascade()).

Each of these tests costs the oracle about half a second of CPU (long
propagation chains), so the outcomes of a seeded random sample of 3,000 tests
are computed once here and committed; the GPU test compares the device's
complete 128,000-test pass with them, and re-checks a smaller part live.

Stored: the root fixpoint's domain hash, the sample's (variable, value) pairs,
and fail[i] = 1 iff singleton test i wipes out under the oracle's AC3bit.

Run from the repository root:  python tests/golden/cascade/make_cascade.py
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, REPO)

import oracle  # noqa: E402
from cpim_amd import workload  # noqa: E402
from cpim_amd.model import domains_to_sets  # noqa: E402

SAMPLE = 3000
SAMPLE_SEED = 2026


def main():
    desc = workload.batched_sac_cascade()
    o = oracle.Oracle(desc)
    ok, root, _ = o.root_ac()
    assert ok
    tests = np.array([(v, a) for v, vals in enumerate(domains_to_sets(root, desc.dom_size))
                      for a in vals], np.int32)
    rng = np.random.default_rng(SAMPLE_SEED)
    pick = np.sort(rng.choice(len(tests), SAMPLE, replace=False))
    sample = tests[pick]
    t0 = time.time()
    fail, _ = o.singleton_tests(sample, nthreads=os.cpu_count() or 1)
    print(f"{SAMPLE} tests in {time.time() - t0:.0f} s, {int(fail.sum())} fail", flush=True)
    np.savez_compressed(os.path.join(HERE, "c3_cascade.npz"), tests=sample, fail=fail,
                        root_hash=np.uint64(oracle.dom_hash(root)), num_tests=len(tests))


if __name__ == "__main__":
    main()
