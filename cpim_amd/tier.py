"""Batch runner over XCSP instance files: the counterpart of the reference's
samples/batch_test_v2.py (run_cpim :62-165, run_tier :233-270, print_summary
:273-334, export_csv :337-351) for the device solver.

    python -m cpim_amd.tier PATH... [--timeout S] [--varh dom|dom/deg|dom/wdeg]
                            [--verify] [--stats] [--export-csv FILE] [--quiet]

PATH is an XCSP file, a directory (every *.xml below it, sorted), or @LIST, a
text file with one path per line (relative to the list's folder; blank lines
and lines starting with '#' are skipped) -- the form in which the reference's
tier lists (tests/tier_definitions.py) can be handed over.

Each instance runs as `python -m cpim_amd.solve FILE --time-limit` in a child
process with a wall-clock limit, exactly as the reference runs
cpim_test_parser, and the same patterns read its output ("MAC stats: time=..
ms, positives=.., negatives=..", "N variables", "M constraints").  The
reference compares against OR-Tools, which is not part of this package; the
check offered instead is --verify: a found solution is tested against every
constraint of the file (the reference's verify_search role).  Exit status 0
when no instance ended in ERROR or a failed verification, 1 otherwise.
"""
from __future__ import annotations

import argparse
import csv
import os
import re
import subprocess
import sys
import time
from dataclasses import asdict, dataclass
from typing import Iterable, List, Optional, Sequence

from . import xcsp


@dataclass
class TierResult:
    file: str
    num_vars: int
    num_cons: int
    status: str          # SAT, UNSAT, TIMEOUT, ERROR:...
    time_ms: int
    branches: int        # num_positive
    backtracks: int      # num_negative
    verified: str        # OK, FAIL, N/A

    def __str__(self) -> str:
        run = f"{self.status}({self.time_ms}ms,{self.branches}b,{self.backtracks}f)"
        return f"{self.file:<50} {run:<32} {self.verified}"


def expand_paths(paths: Sequence[str]) -> List[str]:
    """Files, directories (recursive *.xml), @list files and bench manifests
    (BMPath.xml: every file of every entry, src/model/libxml2_parser.cpp:165-283),
    in order, without repeats."""
    out: List[str] = []
    seen = set()

    def add(p: str) -> None:
        p = os.path.abspath(p)
        if p not in seen:
            seen.add(p)
            out.append(p)

    for p in paths:
        if p.startswith("@"):
            lst = p[1:]
            base = os.path.dirname(os.path.abspath(lst))
            with open(lst) as f:
                for line in f:
                    line = line.strip()
                    if line and not line.startswith("#"):
                        add(line if os.path.isabs(line) else os.path.join(base, line))
        elif os.path.isdir(p):
            found = []
            for root, _, files in os.walk(p):
                found += [os.path.join(root, f) for f in files if f.endswith(".xml")]
            for f in sorted(found):
                add(f)
        elif _is_manifest_file(p):
            for e in xcsp.load_bench_manifest(p):
                for f in e.files:
                    add(f.path)
        else:
            add(p)
    return out


def _is_manifest_file(path: str) -> bool:
    if not path.endswith(".xml") or not os.path.isfile(path):
        return False
    try:
        import xml.etree.ElementTree as ET
        return xcsp.is_manifest(ET.parse(path).getroot())
    except Exception:
        return False


def parse_solver_output(output: str, returncode: Optional[int]) -> dict:
    """Fields of one solver run, read as batch_test_v2.run_cpim reads them."""
    def num(pattern: str) -> int:
        m = re.search(pattern, output)
        return int(m.group(1)) if m else 0

    if "did not find a solution" in output or "UNSAT" in output:
        status = "UNSAT"
    elif "MAC solution" in output:
        status = "SAT"
    elif "timed out" in output or returncode == 124:
        status = "TIMEOUT"
    elif returncode == 2 or "error:" in output:
        line = next((l for l in output.splitlines() if l.startswith("error:")), "error")
        status = "ERROR:" + line[len("error:"):].strip()[:40]
    elif "Traceback" in output:
        lines = [l for l in output.strip().splitlines() if l.strip()]
        status = "ERROR:" + (lines[-1].strip()[:40] if lines else "exception")
    else:
        status = "UNKNOWN"
    sol = None
    m = re.search(r"MAC solution \(canonical indices\):((?: -?\d+)*)", output)
    if m:
        sol = [int(x) for x in m.group(1).split()]
    return {
        "status": status,
        "time_ms": num(r"time=(\d+)\s*ms"),
        "branches": num(r"positives=(\d+)"),
        "backtracks": num(r"negatives=(\d+)"),
        "num_vars": num(r"(\d+)\s+variables"),
        "num_cons": num(r"(\d+)\s+constraints"),
        "solution": sol,
    }


def check_solution(inst: "xcsp.XcspInstance", sol: Sequence[int]) -> bool:
    """Every constraint of the instance holds under sol (canonical values)."""
    desc = inst.desc
    if len(sol) != desc.num_vars:
        return False
    for v, a in enumerate(sol):
        if not 0 <= a < int(desc.dom_size[v]):
            return False
    for scope, table in zip(desc.scopes(), desc.tuple_lists()):
        t = tuple(int(sol[v]) for v in scope)
        if not any(tuple(int(x) for x in row) == t for row in table):
            return False
    return True


def run_one(path: str, timeout_s: int, varh: str = "dom", verify: bool = False,
            skip_unsupported: bool = False) -> TierResult:
    cmd = [sys.executable, "-m", "cpim_amd.solve", path, "--time-limit", str(timeout_s * 1000),
           "--varh", varh]
    if skip_unsupported:
        cmd.append("--skip-unsupported")
    env = dict(os.environ)
    pkg_root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    env["PYTHONPATH"] = pkg_root + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
    t0 = time.monotonic()
    try:
        r = subprocess.run(cmd, capture_output=True, text=True, timeout=timeout_s + 30, env=env)
        fields = parse_solver_output(r.stdout + r.stderr, r.returncode)
    except subprocess.TimeoutExpired:
        fields = {"status": "TIMEOUT", "time_ms": int((time.monotonic() - t0) * 1000), "branches": 0,
                  "backtracks": 0, "num_vars": 0, "num_cons": 0, "solution": None}
    verified = "N/A"
    if verify and fields["status"] == "SAT" and fields["solution"] is not None:
        try:
            inst = xcsp.read(path, skip_unsupported=skip_unsupported)
            verified = "OK" if check_solution(inst, fields["solution"]) else "FAIL"
        except (xcsp.XcspError, OSError):
            verified = "FAIL"
    return TierResult(os.path.basename(path), fields["num_vars"], fields["num_cons"], fields["status"],
                      fields["time_ms"], fields["branches"], fields["backtracks"], verified)


def print_summary(results: List[TierResult], show_stats: bool = False) -> None:
    total = len(results)
    print("=" * 100)
    print("Summary:")
    print(f"  Total: {total}")
    for s in ("SAT", "UNSAT", "TIMEOUT"):
        print(f"  {s}: {sum(1 for r in results if r.status == s)}")
    errors = [r for r in results if r.status.startswith("ERROR") or r.status == "UNKNOWN"]
    print(f"  Errors: {len(errors)}")
    for r in errors:
        print(f"    - {r.file}: {r.status}")
    ok = sum(1 for r in results if r.verified == "OK")
    bad = [r for r in results if r.verified == "FAIL"]
    print(f"  Verified: OK={ok}, FAIL={len(bad)}")
    for r in bad:
        print(f"    - {r.file}")
    if show_stats:
        done = [r for r in results if r.status in ("SAT", "UNSAT")]
        if done:
            print("  Performance Statistics:")
            print(f"    avg time: {sum(r.time_ms for r in done) / len(done):.1f}ms")
            print(f"    avg branches: {sum(r.branches for r in done) / len(done):.1f}")
            print(f"    avg backtracks: {sum(r.backtracks for r in done) / len(done):.1f}")
    print("=" * 100)


def export_csv(results: Iterable[TierResult], path: str) -> None:
    rows = [asdict(r) for r in results]
    with open(path, "w", newline="") as f:
        w = csv.DictWriter(f, fieldnames=list(TierResult.__dataclass_fields__))
        w.writeheader()
        w.writerows(rows)
    print(f"Results exported to: {path}")


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cpim_amd.tier", description=__doc__.split("\n")[0])
    ap.add_argument("paths", nargs="+", help="XCSP files, folders, or @list files")
    ap.add_argument("--timeout", type=int, default=30, help="seconds per instance (default 30)")
    ap.add_argument("--varh", choices=["dom", "dom/deg", "dom/wdeg"], default="dom")
    ap.add_argument("--verify", action="store_true", help="check every found solution against the file")
    ap.add_argument("--stats", action="store_true", help="print average time and node counts")
    ap.add_argument("--export-csv", type=str, help="write the results to this CSV file")
    ap.add_argument("--quiet", action="store_true", help="print the summary only")
    ap.add_argument("--skip-unsupported", action="store_true",
                    help="ignore constraints that reference no extension relation")
    a = ap.parse_args(argv)
    files = expand_paths(a.paths)
    if not a.quiet:
        print("=" * 100)
        print(f"cpim_amd tier run ({len(files)} instances, timeout={a.timeout}s, varh={a.varh})")
        print("=" * 100)
        print(f"{'File':<50} {'Device MAC':<32} Verified")
        print("-" * 100)
    results = []
    for i, f in enumerate(files, 1):
        r = run_one(f, a.timeout, a.varh, a.verify, a.skip_unsupported)
        results.append(r)
        if not a.quiet:
            print(f"[{i}/{len(files)}] {r}", flush=True)
    print_summary(results, a.stats)
    if a.export_csv:
        export_csv(results, a.export_csv)
    failed = any(r.status.startswith("ERROR") or r.status == "UNKNOWN" or r.verified == "FAIL"
                 for r in results)
    return 1 if failed else 0


if __name__ == "__main__":
    sys.exit(main())
