"""Command-line solver over the device engine, for XCSP 2.x instance files.

    python -m cpim_amd.solve FILE.xml [--time-limit MS] [--node-limit N]
                             [--varh dom|dom/deg|dom/wdeg] [--sac] [--dway] [--gac-only]

Prints, like the reference's cpim_test_parser (samples/main_new_parser.cpp:
200-243), the model size, the solution in canonical indices and in original
values, and the line

    MAC stats: time=T ms, positives=P, negatives=N, nodes=P+N, timeout=false

that the reference's batch tooling parses (samples/batch_test_v2.py:101-117).
The reference prints its never-incremented `nodes` counter (0) there; this
tool prints the number of search nodes.  Exit status: 0 solved, 1 no solution
or limit reached, 2 bad input.
"""
from __future__ import annotations

import argparse
import sys
import time

from . import xcsp
from .gmodel import GModelAdapter, GModelOptions, GModelSolver
from .model import domains_to_sets


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m cpim_amd.solve", description=__doc__.split("\n")[0])
    ap.add_argument("file")
    ap.add_argument("--time-limit", type=int, default=0, help="ms, 0 = none")
    ap.add_argument("--node-limit", type=int, default=0, help="search nodes, 0 = none")
    ap.add_argument("--varh", choices=["dom", "dom/deg", "dom/wdeg"], default="dom",
                    help="variable ordering of the binary search (MAC.cpp:353-411)")
    ap.add_argument("--sac", action="store_true", help="enforce SAC before the search")
    ap.add_argument("--dway", action="store_true", help="d-way search of the reference GModelSolver")
    ap.add_argument("--gac-only", action="store_true", help="print the root GAC domains and stop")
    ap.add_argument("--skip-unsupported", action="store_true",
                    help="ignore constraints that reference no extension relation")
    a = ap.parse_args(argv)
    if a.dway and a.varh != "dom":
        print("error: the d-way search orders variables by smallest domain only", file=sys.stderr)
        return 2
    try:
        inst = xcsp.read(a.file, skip_unsupported=a.skip_unsupported)
    except (xcsp.XcspError, OSError) as e:
        print(f"error: {e}", file=sys.stderr)
        return 2
    print(f"Model {inst.name or a.file}: {inst.num_vars} variables, {inst.num_cons} constraints")
    for s in inst.skipped:
        print(f"warning: skipped {s} (no extension relation)", file=sys.stderr)
    with GModelAdapter.Build(inst.desc, GModelOptions(skip_non_binary=False)) as m:
        if a.gac_only:
            st = m.EnforceGAC(False)
            print(f"GAC: {'inconsistent' if st.inconsistent else 'consistent'}, deletions={st.deletions}")
            if not st.inconsistent:
                for v, vals in enumerate(domains_to_sets(m.ReadDomains(0), inst.desc.dom_size)):
                    print(f"  {inst.var_names[v]}: {[inst.var_values[v][x] for x in vals]}")
            return 0 if not st.inconsistent else 1
        t0 = time.perf_counter()
        if a.sac:
            sac = m.EnforceSAC()
            print(f"SAC: {'consistent' if sac.consistent else 'inconsistent'}, removed={sac.removed}, "
                  f"passes={sac.passes}")
            if not sac.consistent:
                print("MAC search did not find a solution.")
                print(f"MAC stats: time={int((time.perf_counter() - t0) * 1000)} ms, positives=0, "
                      f"negatives=0, nodes=0, timeout=false")
                return 1
        solver = GModelSolver(m, branching="dway" if a.dway else "binary", heuristic=a.varh)
        st = solver.Solve(time_limit=a.time_limit, node_limit=a.node_limit)
        ms = int((time.perf_counter() - t0) * 1000)
        sol = solver.GetSolution()
        if sol:
            print("MAC solution (canonical indices): " + " ".join(map(str, sol)))
            print("MAC solution (original values): " + " ".join(map(str, inst.original_solution(sol))))
        elif st.time_out:
            print("MAC search timed out before finding a solution.")
        else:
            print("MAC search did not find a solution.")
        print(f"MAC stats: time={ms} ms, positives={st.num_positive}, negatives={st.num_negative}, "
              f"nodes={st.num_positive + st.num_negative}, timeout={'true' if st.time_out else 'false'}")
        return 0 if sol else 1


if __name__ == "__main__":
    sys.exit(main())
