"""Write tests/golden/*.npz from the reference's fixture instances.

Reads the XCSP 2.x files under /root/reference/samples/bench (data only),
normalizes them the way the reference does (tests/golden/xcsp2.py), and stores
for each instance:
  * the normalized network (domain sizes, scopes, dense binary relations);
  * the oracle's root AC3bit fixpoint, SAC1 closure and MAC run (dom-min /
    min-value, node limit for the instances MAC cannot finish quickly) with
    its first nodes;
  * brute-force confirmations: definitional GAC and SAC closures, and the
    MAC run of the independent closure-per-node driver.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import glob
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

import oracle  # noqa: E402
import xcsp2  # noqa: E402
from cpim_amd.model import dense_from_tuples, from_tuples  # noqa: E402

BENCH = "/root/reference/samples/bench"
TRACE_CAP = 512
# MAC node limits for instances whose full search takes minutes on the CPU.
NODE_LIMIT = {"haystacks-11_ext": 3000, "rand-23-23-253-131": 3000}


def fixture_files():
    files = ["test.xml", "queens-4_ext.xml", "queens-12_ext.xml", "haystacks-11_ext.xml"]
    files = [os.path.join(BENCH, f) for f in files]
    files += sorted(glob.glob(os.path.join(BENCH, "composed-25-1-2", "*.xml")))
    files += sorted(glob.glob(os.path.join(BENCH, "rand-2-23", "*.xml")))
    return files


def node_limit_for(name: str) -> int:
    for k, v in NODE_LIMIT.items():
        if name.startswith(k):
            return v
    return 0


def main():
    for path in fixture_files():
        name = os.path.basename(path).replace(".xml", "")
        ds, scopes, tuples = xcsp2.read(path)
        desc = dense_from_tuples(from_tuples(ds, scopes, tuples, name=name))
        root_ok, root_dom, _ = oracle.Oracle(desc).root_ac()
        sac_ok, sac_dom, _ = oracle.Oracle(desc).sac1()
        limit = node_limit_for(name)
        mac = oracle.Oracle(desc).mac(node_limit=limit, trace_cap=TRACE_CAP)
        # brute-force confirmations
        bf_ok, bf_dom = oracle.bf_gac(desc)
        assert bf_ok == root_ok and (not root_ok or (bf_dom == root_dom).all()), name
        bs_ok, bs_dom = oracle.bf_sac(desc)
        assert bs_ok == sac_ok and (not sac_ok or (bs_dom == sac_dom).all()), name
        bmac = oracle.bf_mac(desc, node_limit=limit, trace_cap=TRACE_CAP)
        assert (bmac.num_sol, bmac.num_positive, bmac.num_negative, bmac.time_out) == \
            (mac.num_sol, mac.num_positive, mac.num_negative, mac.time_out), name
        assert bmac.trace == mac.trace, name
        trace = np.array([[p, v, a, c, h] for (p, v, a, c, h) in mac.trace], dtype=np.uint64) \
            .reshape(-1, 5)
        out = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(
            out,
            dom_size=desc.dom_size, con_arity=desc.con_arity, con_scope=desc.con_scope,
            bin_rel=desc.bin_rel,
            root_ok=np.int32(root_ok), root_dom=root_dom,
            sac_ok=np.int32(sac_ok), sac_dom=sac_dom,
            node_limit=np.int64(limit),
            mac_num_sol=np.int64(mac.num_sol), mac_pos=np.int64(mac.num_positive),
            mac_neg=np.int64(mac.num_negative), mac_timeout=np.int32(mac.time_out),
            mac_solution=(mac.solution if mac.solution is not None
                          else np.full(desc.num_vars, -1, np.int32)),
            mac_trace=trace)
        print(f"{name}: n={desc.num_vars} e={desc.num_cons} d={desc.max_dom} root={root_ok} "
              f"sac={sac_ok} mac sol={mac.num_sol} pos={mac.num_positive} neg={mac.num_negative} "
              f"timeout={mac.time_out}")


if __name__ == "__main__":
    main()
