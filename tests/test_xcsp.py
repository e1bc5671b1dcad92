"""XCSP 2.x reader (CPU): the reference's fixture files read by the product
reader equal the golden models, and the normalization rules of
src/model/model_normalizer.cpp hold on hand-made instances."""
import os
import shutil

import numpy as np
import pytest

from conftest import load_golden
from cpim_amd import tier, xcsp

HERE = os.path.dirname(os.path.abspath(__file__))
XML = os.path.join(HERE, "golden", "xcsp")
FILES = sorted(f for f in os.listdir(XML) if f.endswith(".xml"))


def as_sets(desc):
    return [set(map(tuple, np.asarray(t).reshape(-1, len(s)).tolist()))
            for t, s in zip(desc.tuple_lists(), desc.scopes())]


@pytest.mark.parametrize("fname", FILES)
def test_fixture_matches_golden_model(fname):
    inst = xcsp.read(os.path.join(XML, fname))
    gdesc, _ = load_golden(fname[:-4])
    assert list(inst.desc.dom_size) == list(gdesc.dom_size)
    assert [list(s) for s in inst.desc.scopes()] == [list(s) for s in gdesc.scopes()]
    assert as_sets(inst.desc) == as_sets(gdesc)


def test_test_xml_known_values():
    inst = xcsp.read(os.path.join(XML, "test.xml"))
    assert inst.var_names == ["V0", "V1", "V2"]
    assert inst.var_values == [[0, 1, 2]] * 3
    assert as_sets(inst.desc) == [{(0, 0), (1, 1), (2, 2)}, {(0, 1), (0, 2), (1, 2)}]


DOC = """<instance><presentation name="h"/>
<domains><domain name="D0">7 3,5 3</domain><domain name="D1">-1..1</domain></domains>
<variables><variable name="X" domain="D0"/><variable name="Y" domain="D1"/><variable name="Z" domain="D1"/></variables>
<relations>
 <relation name="R0" arity="2" semantics="conflicts">3 -1|7 1|3 -1</relation>
 <relation name="R1" arity="3" semantics="supports">5 0 1|7,1,-1</relation>
</relations>
<constraints>
 <constraint name="C0" arity="2" scope="X Y" reference="R0"/>
 <constraint name="C1" arity="3" scope="X Y Z" reference="R1"/>
 <constraint name="C2" arity="2" scope="Y Z" reference="P0"/>
</constraints></instance>"""


def test_normalization_rules():
    with pytest.raises(xcsp.XcspError, match="not an extension relation"):
        xcsp.read(DOC)
    inst = xcsp.read(DOC, skip_unsupported=True)
    assert inst.skipped == ["C2"]
    # enumerated domain sorted and deduplicated, range expanded
    assert inst.var_values == [[3, 5, 7], [-1, 0, 1], [-1, 0, 1]]
    c0, c1 = inst.desc.tuple_lists()
    # conflicts -> complement over canonical values, lexicographic order
    forbidden = {(0, 0), (2, 2)}
    want = [(a, b) for a in range(3) for b in range(3) if (a, b) not in forbidden]
    assert [tuple(t) for t in c0.reshape(-1, 2).tolist()] == want
    assert [tuple(t) for t in c1.reshape(-1, 3).tolist()] == [(1, 1, 2), (2, 2, 0)]
    assert inst.original_solution([2, 0, 1]) == [7, -1, 0]


@pytest.mark.parametrize("bad, msg", [
    (DOC.replace("5 0 1|", "5 0 9|"), "not present in original domain"),
    (DOC.replace("5 0 1|", "5 0|"), "Tuple size mismatch"),
    (DOC.replace('scope="X Y" reference="R0"', 'scope="X Q" reference="R0"'), "Variable not found"),
    (DOC.replace(">7 3,5 3<", "><"), "Domain values are empty"),
])
def test_errors(bad, msg):
    with pytest.raises(xcsp.XcspError, match=msg):
        xcsp.read(bad, skip_unsupported=True)


def test_expanded_conflicts_limit():
    with pytest.raises(xcsp.XcspError, match="exceeds limit"):
        xcsp.read(DOC, skip_unsupported=True, max_expanded_tuples=8)


# ---------------------------------------------------------------------------
# Tier runner (samples/batch_test_v2.py counterpart): host-side parts
# ---------------------------------------------------------------------------
def test_tier_expand_paths(tmp_path):
    from cpim_amd import tier
    lst = tmp_path / "tier0.txt"
    lst.write_text("# smoke tier\n%s\n\n%s\n" % (os.path.join(XML, "test.xml"), "queens-4_ext.xml"))
    files = tier.expand_paths([XML, "@" + str(lst)])
    names = [os.path.basename(f) for f in files]
    assert names == sorted(f for f in os.listdir(XML) if f.endswith(".xml")) + ["queens-4_ext.xml"]
    # the list's relative entry resolves against the list's folder
    assert files[-1] == os.path.join(str(tmp_path), "queens-4_ext.xml")


def test_tier_parses_solver_output():
    from cpim_amd import tier
    out = ("Model q: 4 variables, 6 constraints\n"
           "MAC solution (canonical indices): 1 3 0 2\n"
           "MAC solution (original values): 2 4 1 3\n"
           "MAC stats: time=12 ms, positives=5, negatives=1, nodes=6, timeout=false\n")
    f = tier.parse_solver_output(out, 0)
    assert (f["status"], f["time_ms"], f["branches"], f["backtracks"], f["num_vars"], f["num_cons"]) == \
        ("SAT", 12, 5, 1, 4, 6)
    assert f["solution"] == [1, 3, 0, 2]
    f = tier.parse_solver_output("MAC search did not find a solution.\nMAC stats: time=3 ms, "
                                 "positives=7, negatives=7, nodes=14, timeout=false\n", 1)
    assert (f["status"], f["branches"], f["solution"]) == ("UNSAT", 7, None)
    f = tier.parse_solver_output("MAC search timed out before finding a solution.\n", 1)
    assert f["status"] == "TIMEOUT"
    assert tier.parse_solver_output("error: no such file\n", 2)["status"].startswith("ERROR:no such")


def test_tier_check_solution():
    from cpim_amd import tier
    inst = xcsp.read(os.path.join(XML, "queens-4_ext.xml"))
    assert tier.check_solution(inst, [1, 3, 0, 2])
    assert not tier.check_solution(inst, [0, 1, 2, 3])
    assert not tier.check_solution(inst, [1, 3, 0])
    assert not tier.check_solution(inst, [1, 3, 0, 9])


def test_cli_rejects_dway_with_other_orderings(capsys):
    from cpim_amd import solve
    code = solve.main([os.path.join(XML, "test.xml"), "--dway", "--varh", "dom/wdeg"])
    assert code == 2 and "d-way" in capsys.readouterr().err


def test_tier_reports_unreadable_files(tmp_path):
    # the child process fails on input before any device work
    from cpim_amd import tier
    bad = tmp_path / "broken.xml"
    bad.write_text("<instance><presentation name='x'/><domains nbDomains='1'>")
    r = tier.run_one(str(bad), timeout_s=20)
    assert r.status.startswith("ERROR"), r
    r = tier.run_one(str(tmp_path / "missing.xml"), timeout_s=20)
    assert r.status.startswith("ERROR"), r
    assert tier.main([str(bad), "--quiet"]) == 1


# ---------------------------------------------------------------------------
# Bench manifests (BMPath.xml, src/model/libxml2_parser.cpp:88-116, 165-383)
# ---------------------------------------------------------------------------
GOLDEN_XCSP = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "xcsp")


def manifest(entries):
    body = "".join(f'<BMFile id="{i}"{(" type=%s" % repr(t)) if t else ""}>{p}</BMFile>'
                   for i, (p, t) in enumerate(entries))
    return f'<?xml version="1.0"?><BenchMark><BMFiles nbBMFiles="{len(entries)}">{body}</BMFiles></BenchMark>'


def test_manifest_in_the_reference_layout(tmp_path):
    # samples/bench/BMPath.xml names ../samples/bench/<file>: relative to the
    # manifest's grandparent
    bench = tmp_path / "samples" / "bench"
    bench.mkdir(parents=True)
    shutil.copy(os.path.join(GOLDEN_XCSP, "queens-12_ext.xml"), bench / "queens-12_ext.xml")
    m = bench / "BMPath.xml"
    m.write_text(manifest([("../samples/bench/queens-12_ext.xml", None)]))
    entries = xcsp.load_bench_manifest(str(m))
    assert len(entries) == 1 and entries[0].kind == "file"
    assert entries[0].files[0].path == str((bench / "queens-12_ext.xml").resolve())
    inst = xcsp.read(str(m))
    direct = xcsp.read(str(bench / "queens-12_ext.xml"))
    assert inst.num_vars == direct.num_vars == 12 and inst.num_cons == direct.num_cons


def test_manifest_directory_entries_and_tier_expansion(tmp_path):
    d = tmp_path / "suite"
    (d / "sub").mkdir(parents=True)
    shutil.copy(os.path.join(GOLDEN_XCSP, "queens-4_ext.xml"), d / "b.xml")
    shutil.copy(os.path.join(GOLDEN_XCSP, "test.xml"), d / "sub" / "a.xml")
    shutil.copy(os.path.join(GOLDEN_XCSP, "queens-12_ext.xml"), tmp_path / "q12.xml")
    m = tmp_path / "BMPath.xml"
    m.write_text(manifest([("suite", "dir"), ("q12.xml", None), ("  ", None)]))
    entries = xcsp.load_bench_manifest(str(m))
    assert [e.kind for e in entries] == ["dir", "file"]
    assert [os.path.basename(f.path) for f in entries[0].files] == ["b.xml", "a.xml"]
    assert all(f.format for e in entries for f in e.files)
    # the parser reads the first file of the first entry
    assert xcsp.read(str(m)).num_vars == 4
    assert [os.path.basename(p) for p in tier.expand_paths([str(m)])] == ["b.xml", "a.xml", "q12.xml"]


@pytest.mark.parametrize("case", ["no_section", "empty", "missing", "type_mismatch"])
def test_manifest_errors(tmp_path, case):
    shutil.copy(os.path.join(GOLDEN_XCSP, "test.xml"), tmp_path / "t.xml")
    m = tmp_path / "BMPath.xml"
    if case == "no_section":
        m.write_text('<BenchMark><BMFile id="0">t.xml</BMFile></BenchMark>')
        with pytest.raises(xcsp.XcspError, match="missing <BMFiles>"):
            xcsp.load_bench_manifest(str(m))
    elif case == "empty":
        m.write_text(manifest([("", None)]))
        with pytest.raises(xcsp.XcspError, match="did not contain usable"):
            xcsp.load_bench_manifest(str(m))
    elif case == "missing":
        m.write_text(manifest([("nowhere.xml", None)]))
        with pytest.raises(FileNotFoundError, match="Bench path not found"):
            xcsp.read(str(m))
    else:
        m.write_text(manifest([("t.xml", "dir")]))
        with pytest.raises(xcsp.XcspError, match="expected directory but resolved to file"):
            xcsp.load_bench_manifest(str(m))
