"""XCSP files through the device solver (needs an MI355X): the reference's
own instances, read by the product reader, give the reference's known
answers and the oracle's node counts, and the CLI prints the reference's
'MAC stats:' line."""
import os
import re

import pytest

import oracle
from cpim_amd import GModelAdapter, solve, xcsp

pytestmark = pytest.mark.gpu

XML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "xcsp")


def run_cli(capsys, *args):
    code = solve.main(list(args))
    return code, capsys.readouterr().out


def test_cli_known_answers(gpu, capsys):
    code, out = run_cli(capsys, os.path.join(XML, "test.xml"))
    assert code == 0
    assert "3 variables, 2 constraints" in out
    assert "MAC solution (original values): 0 0 1" in out
    assert re.search(r"MAC stats: time=\d+ ms, positives=3, negatives=0, nodes=3, timeout=false", out)
    code, out = run_cli(capsys, os.path.join(XML, "queens-4_ext.xml"))
    # the known answer [1, 3, 0, 2] in canonical indices; the file's domain is 1..4
    assert code == 0 and "MAC solution (canonical indices): 1 3 0 2" in out
    assert "MAC solution (original values): 2 4 1 3" in out
    code, out = run_cli(capsys, os.path.join(XML, "composed-25-1-2-2_ext.xml"))
    assert code == 1 and "did not find a solution" in out


@pytest.mark.parametrize("fname", sorted(f for f in os.listdir(XML) if f.endswith(".xml")))
def test_xcsp_files_match_oracle(gpu, fname):
    inst = xcsp.read(os.path.join(XML, fname))
    r = oracle.Oracle(inst.desc).mac(node_limit=20000)
    with GModelAdapter.Build(inst.desc) as m:
        st, sol, _ = m.Solve(node_limit=20000)
    assert (st.num_solutions, st.num_positive, st.num_negative) == \
        (r.num_sol, r.num_positive, r.num_negative)
    if r.num_sol:
        assert list(sol) == list(r.solution)


def test_cli_sac_and_gac_only(gpu, capsys):
    code, out = run_cli(capsys, os.path.join(XML, "queens-4_ext.xml"), "--gac-only")
    assert code == 0 and "GAC: consistent" in out
    code, out = run_cli(capsys, os.path.join(XML, "queens-4_ext.xml"), "--sac")
    assert "SAC: consistent" in out and "MAC stats:" in out
    code, _ = run_cli(capsys, os.path.join(XML, "missing.xml"))
    assert code == 2


def test_tier_runner_over_the_fixtures(gpu, tmp_path, capsys):
    # batch_test_v2.py counterpart: every file through the CLI in a child
    # process, solutions verified against the files, results in CSV
    import csv

    from cpim_amd import tier
    out = tmp_path / "tier.csv"
    code = tier.main([XML, "--timeout", "3", "--verify", "--stats", "--export-csv", str(out)])
    text = capsys.readouterr().out
    assert code == 0, text
    rows = {r["file"]: r for r in csv.DictReader(open(out))}
    assert sorted(rows) == sorted(f for f in os.listdir(XML) if f.endswith(".xml"))
    t = rows["test.xml"]
    assert (t["status"], t["branches"], t["backtracks"], t["verified"]) == ("SAT", "3", "0", "OK")
    assert (t["num_vars"], t["num_cons"]) == ("3", "2")
    assert (rows["queens-4_ext.xml"]["status"], rows["queens-4_ext.xml"]["verified"]) == ("SAT", "OK")
    assert rows["composed-25-1-2-2_ext.xml"]["status"] == "UNSAT"
    for r in rows.values():
        assert r["status"] in ("SAT", "UNSAT", "TIMEOUT"), r
        assert r["verified"] == ("OK" if r["status"] == "SAT" else "N/A"), r
    assert "Verified: OK=" in text


@pytest.mark.parametrize("varh,code", [("dom/deg", 1), ("dom/wdeg", 2)])
def test_cli_variable_orderings(gpu, capsys, varh, code):
    path = os.path.join(XML, "queens-12_ext.xml")
    inst = xcsp.read(path)
    with oracle.var_heuristic(code):
        r = oracle.Oracle(inst.desc).mac(node_limit=20000)
    rc, out = run_cli(capsys, path, "--varh", varh, "--node-limit", "20000")
    m = re.search(r"positives=(\d+), negatives=(\d+)", out)
    assert m and (int(m.group(1)), int(m.group(2))) == (r.num_positive, r.num_negative)
    assert (rc == 0) == bool(r.num_sol)


CPP_CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cpim_amd", "cpim_solve")


@pytest.mark.parametrize("varh,code", [("dom", 0), ("domdeg", 1), ("domwdeg", 2)])
@pytest.mark.parametrize("limit", [0, 3000])
def test_cpp_cli_matches_oracle(gpu, varh, code, limit):
    # the C++ facade (include/cpim/gmodel.hpp) through cpim_solve: same
    # generated instance as workload.random_binary, node counts of the oracle
    import subprocess
    from cpim_amd import workload
    if not os.path.exists(CPP_CLI):
        pytest.skip("cpim_solve not built")
    args = ["--n", "30", "--d", "16", "--e", "120", "--p", "0.4", "--seed", "11", "--varh", varh]
    if limit:
        args += ["--node_limit", str(limit)]
    r = subprocess.run([CPP_CLI] + args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    m = re.search(r"positives=(\d+), negatives=(\d+)", r.stdout)
    desc = workload.random_binary(30, 16, 120, 0.4, 11)
    with oracle.var_heuristic(code):
        o = oracle.Oracle(desc).mac(node_limit=limit)
    assert m and (int(m.group(1)), int(m.group(2))) == (o.num_positive, o.num_negative), r.stdout
    if o.num_sol:
        sol = re.search(r"MAC solution \(canonical indices\):((?: \d+)*)", r.stdout)
        assert sol and [int(x) for x in sol.group(1).split()] == list(o.solution)


def test_cpp_cli_sac_and_dway(gpu):
    import subprocess
    from cpim_amd import workload
    if not os.path.exists(CPP_CLI):
        pytest.skip("cpim_solve not built")
    base = ["--n", "30", "--d", "16", "--e", "120", "--p", "0.4", "--seed", "11"]
    r = subprocess.run([CPP_CLI] + base + ["--sac"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    sok, sdom, _ = oracle.Oracle(workload.random_binary(30, 16, 120, 0.4, 11)).sac1()
    assert ("consistent=true" if sok else "consistent=false") in r.stdout
    r = subprocess.run([CPP_CLI] + base + ["--dway"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and "MAC stats:" in r.stdout, r.stdout + r.stderr
    r = subprocess.run([CPP_CLI] + base + ["--dway", "--varh", "domdeg"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
