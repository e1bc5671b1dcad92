"""The allocation-free fixpoint entry point against the one it varies (needs an MI355X).

cpim_gmodel_enforce_gac_noalloc runs the persistent fixpoint with its control
block and frontiers in a file-scope device array instead of allocated
memory.  Called through the C-ABI, it must give what cpim_gmodel_enforce_gac
in CPIM_GAC_PERSISTENT mode gives, bit for bit: the same domains, wipe-out
flag, convergence and deletions, on two calls in a row, the second one on a
smaller network (the array keeps the first call's contents, which the entry
point must clear).
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from cpim_amd import GModelAdapter, workload
from cpim_amd import _native as N

pytestmark = pytest.mark.gpu


def noalloc(m, var):
    st = N.CGacStats()
    N.check(N.gpu_lib().cpim_gmodel_enforce_gac_noalloc(m.handle, var, C.byref(st)))
    return st


def existing(m, var):
    st = N.CGacStats()
    N.check(N.gpu_lib().cpim_gmodel_enforce_gac(m.handle, var, N.CPIM_GAC_PERSISTENT, C.byref(st)))
    return st


def same(a, b):
    return (a.deletions, a.inconsistent, a.converged) == (b.deletions, b.inconsistent, b.converged)


def test_noalloc_matches_persistent_entry_twice(gpu):
    big = workload.single_large_cascade()             # 40,000 constraints, 25 rounds
    small = workload.random_binary(300, 32, 2000, 0.8, 11, preremove=0.2)  # 547 deletions
    ok_s, want_s, _ = oracle.Oracle(small).root_ac()
    for desc in (big, small):                          # second call: the smaller network
        with GModelAdapter.Build(desc) as a, GModelAdapter.Build(desc) as b:
            sa = existing(a, -1)
            sb = noalloc(b, -1)
            assert same(sa, sb), (sa.deletions, sb.deletions)
            assert sb.converged or sb.inconsistent
            assert (a.ReadDomains(0) == b.ReadDomains(0)).all()
            if desc is big:
                assert sb.deletions == 15369 and sb.iterations >= 3
            else:
                assert sb.inconsistent == (not ok_s)
                if ok_s:
                    assert (b.ReadDomains(0) == want_s).all()


def test_noalloc_seeded_and_wipe_out(gpu):
    # a one-variable seed after a removal, then a network that wipes out
    desc = workload.random_binary(200, 16, 1500, 0.7, 5)  # 238 deletions at the root
    with GModelAdapter.Build(desc) as a, GModelAdapter.Build(desc) as b:
        existing(a, -1)
        noalloc(b, -1)
        dom = a.ReadDomains(0)
        assert (dom == b.ReadDomains(0)).all()
        var = int(np.argmax([bin(int(w)).count("1") for w in dom[:, 0]]))
        cut = dom.copy()
        cut[var, 0] &= cut[var, 0] - np.uint64(1)       # drop the smallest value
        cut[var, 0] &= cut[var, 0] - np.uint64(1)
        a.WriteDomains(0, cut)
        b.WriteDomains(0, cut)
        sa, sb = existing(a, var), noalloc(b, var)
        assert same(sa, sb)
        assert (a.ReadDomains(0) == b.ReadDomains(0)).all()
    wipe = workload.random_binary(2000, 64, 40000, 0.894, workload.instance_seed(2, 1), preremove=0.10)
    with GModelAdapter.Build(wipe) as a, GModelAdapter.Build(wipe) as b:
        sa, sb = existing(a, -1), noalloc(b, -1)
        assert sa.inconsistent and sb.inconsistent and not sb.converged
