"""The legacy CModel entry points against the oracle (needs an MI355X).

CModel (include/cuSAC.cuh:422-447) is served by the same engine in both
facades: enforceGAC() is the root fixpoint, enforceGAC(var, type) propagates
a decision on var, enforceSAC() is the SAC closure (SAC1, src/SAC1.cpp:57-103;
the reference's body is empty), solve(ms) is the device MAC, and
CreateNewLevel / BackLevel return the new level as the reference does.  The
Python facade is called directly; the C++ one (include/cpim/gmodel.hpp)
through `cpim_solve --cmodel`, which prints the domain hash of each level it
leaves.
"""
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
from cpim_amd import CModel, workload
from cpim_amd.model import domains_to_sets

pytestmark = pytest.mark.gpu

CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cpim_amd", "cpim_solve")
CASES = [(30, 16, 120, 0.40, 1), (40, 8, 200, 0.35, 2), (25, 70, 90, 0.5, 3), (20, 12, 60, 0.55, 4)]


def expected(desc):
    """Root closure, the closure after the first decision the drivers make (the
    lowest variable with two or more values, its smallest value), SAC1 and MAC."""
    ok, root, _ = oracle.Oracle(desc).root_ac()
    out = {"ok": ok, "root": root}
    if ok:
        sets = domains_to_sets(root, desc.dom_size)
        var = next((v for v, s in enumerate(sets) if len(s) >= 2), None)
        out["var"] = var
        if var is not None:
            a = sets[var][0]
            cut = root.copy()
            cut[var, :] = 0
            cut[var, a // 64] = np.uint64(1 << (a % 64))
            d2 = type(desc)(desc.dom_size, desc.con_arity, desc.con_scope, desc.bin_rel, init_dom=cut)
            ok2, dom2, _ = oracle.Oracle(d2).root_ac()
            out["decision"] = (a, ok2, dom2)
    out["sac"] = oracle.Oracle(desc).sac1()
    out["mac"] = oracle.Oracle(desc).mac()
    return out


@pytest.mark.parametrize("case", range(len(CASES)))
def test_python_cmodel(gpu, case):
    n, d, e, p, seed = CASES[case]
    desc = workload.random_binary(n, d, e, p, seed)
    want = expected(desc)
    cm = CModel(desc)
    assert cm.enforceGAC() == want["ok"]
    if want["ok"]:
        assert (cm.g.ReadDomains(0) == want["root"]).all()
        if want.get("var") is not None:
            a, ok2, dom2 = want["decision"]
            assert cm.CreateNewLevel() == 1
            assert cm.g.AssignValue(want["var"], a, 1)
            assert cm.enforceGAC(want["var"], 1) == ok2
            if ok2:
                assert (cm.g.ReadDomains(1) == dom2).all()
            assert cm.BackLevel() == 0
        sok, sdom, _ = want["sac"]
        assert cm.enforceSAC() == sok
        if sok:
            assert (cm.g.ReadDomains(0) == sdom).all()
    fresh = CModel(desc)
    st = fresh.solve(0)
    mac = want["mac"]
    assert (st.num_sol, st.num_positive, st.num_negative) == (mac.num_sol, mac.num_positive, mac.num_negative)
    if mac.num_sol:
        assert list(fresh.solution) == list(mac.solution)


def cli(*args):
    if not os.path.exists(CLI):
        pytest.fail("cpim_solve not built")
    r = subprocess.run([CLI] + [str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("case", range(len(CASES)))
def test_cpp_cmodel(gpu, case):
    n, d, e, p, seed = CASES[case]
    desc = workload.random_binary(n, d, e, p, seed)
    want = expected(desc)
    out = cli("--n", n, "--d", d, "--e", e, "--p", p, "--seed", seed, "--cmodel")
    m = re.search(r"CModel enforceGAC: consistent=(\w+) hash=(\d+)", out)
    assert (m.group(1) == "true") == want["ok"]
    if want["ok"]:
        assert int(m.group(2)) == oracle.dom_hash(want["root"])
        if want.get("var") is not None:
            a, ok2, dom2 = want["decision"]
            m = re.search(r"CModel decision: level=(\d+) var=(\d+) value=(\d+) consistent=(\w+) hash=(\d+)", out)
            assert (int(m.group(1)), int(m.group(2)), int(m.group(3))) == (1, want["var"], a)
            assert (m.group(4) == "true") == ok2
            if ok2:
                assert int(m.group(5)) == oracle.dom_hash(dom2)
        assert "CModel BackLevel: 0" in out
        sok, sdom, _ = want["sac"]
        m = re.search(r"CModel enforceSAC: consistent=(\w+) hash=(\d+)", out)
        assert (m.group(1) == "true") == sok
        if sok:
            assert int(m.group(2)) == oracle.dom_hash(sdom)
    mac = want["mac"]
    m = re.search(r"CModel solve: num_sol=(\d+) positives=(\d+) negatives=(\d+)", out)
    assert tuple(int(x) for x in m.groups()) == (mac.num_sol, mac.num_positive, mac.num_negative)
    if mac.num_sol:
        m = re.search(r"CModel solution: ([\d ]+)", out)
        assert [int(x) for x in m.group(1).split()] == list(mac.solution)


def test_cpp_find_all(gpu):
    n, d, e, p, seed = 9, 3, 12, 0.3, 904
    desc = workload.random_binary(n, d, e, p, seed)
    want = oracle.mac_all(oracle.Oracle(desc), 10**6, sol_cap=4096)
    out = cli("--n", n, "--d", d, "--e", e, "--p", p, "--seed", seed, "--find-all", 1000000)
    sols = [[int(x) for x in line.split()[1:]] for line in out.splitlines() if line.startswith("solution:")]
    m = re.search(r"solutions found: (\d+) kept: (\d+)", out)
    assert int(m.group(1)) == want.num_sol == len(sols)
    assert np.array_equal(np.array(sols), want.solutions)
    mm = re.search(r"MAC stats: .* positives=(\d+), negatives=(\d+)", out)
    assert (int(mm.group(1)), int(mm.group(2))) == (want.num_positive, want.num_negative)
