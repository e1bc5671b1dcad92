#!/usr/bin/env python3
"""Benchmark: batched SAC singleton-test fixpoints per second on one MI355X.

Workload (synthetic, seeded): random binary CSP n=500, d=256 (4 x 64-bit
words), e=20000 distinct pairs, every tuple forbidden with p=0.5.  After the
root GAC fixpoint, one step = one SAC pass = every singleton test (x = a) of
the current domains (128,000 tests), each a full GAC fixpoint on its own
domain copy, all in one persistent launch.  The plain path (one host-driven
fixpoint per test, one launch + sync per round) and the CPU oracle are timed
on bounded samples of the same tests in the same run.

Prints one JSON line.  roofline.achieved counts the schedule-independent
algorithmic bytes of each test's seed revisions (8 W (4 deg(x) + n) per test);
roofline.traffic is the measured HBM bytes per launch from the committed
rocprofv3 counter summary under profiles/, reported only when that summary
was taken from the kernel sources in the tree.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "GAC fixpoints/sec batched vs one-per-launch on MI355X; HBM GB/s; p99 solve ms"
HBM_PEAK_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--plain-sample", type=int, default=400,
                    help="singleton tests timed on the plain one-per-launch path")
    ap.add_argument("--cpu-sample", type=int, default=600,
                    help="singleton tests timed on the CPU oracle (about 10-15 s)")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="threads of the all-cores CPU baseline (the GPU host's CPU share is 16)")
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--sac-mode", default="sliced", choices=["sliced", "workgroup"],
                    help="pass kernel: 64 bit-sliced tests per workgroup, or one test per workgroup")
    ap.add_argument("--skip-extras", action="store_true",
                    help="skip the single-fixpoint, uneven-batch and serving measurements")
    ap.add_argument("--batch-instances", type=int, default=4096,
                    help="instances of the 4096-instance uneven batch timed in the extras")
    ap.add_argument("--per-instance-sample", type=int, default=256,
                    help="uneven instances also timed one launch each (0 = skip that comparison)")
    ap.add_argument("--serve-requests", type=int, default=2048)
    ap.add_argument("--serve-rate", type=float, default=1000.0,
                    help="Poisson arrival rate of serving requests (requests/s)")
    ap.add_argument("--serve-c3-requests", type=int, default=96,
                    help="requests of the benchmark SAC network's shape (with tables) served in the extras")
    ap.add_argument("--serve-c3-rate", type=float, default=8.0,
                    help="Poisson arrival rate of those requests (requests/s)")
    ap.add_argument("--only", choices=["single_fixpoint", "cascade_fixpoint", "uneven_batch", "serving"],
                    default=None,
                    help="run only this measurement of the extras and print its JSON (for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a device copy of MIB MiB (read MIB + write MIB) and exit: the "
                         "reference kernel for scaling rocprof FETCH_SIZE / WRITE_SIZE counters")
    return ap.parse_args()


# committed rocprofv3 summaries of the pass kernel, per SAC mode
PROFILES = {"sliced": os.path.join(REPO, "profiles", "k_sliced_wide_batched_sac.txt")}


KERNEL_SOURCES = ("device.h", "kernels.h")  # every __global__ kernel and the device code it runs


def kernel_source_fingerprint() -> str:
    """sha256 over the engine's device sources (cpim_amd/csrc/device.h and
    kernels.h, in that order): a committed profile taken from other device
    code is stale.  Host-side changes (launch code, serving, bindings) leave
    it unchanged."""
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        h.update(name.encode())
        with open(os.path.join(REPO, "cpim_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()


def profiled_traffic(mode: str):
    """(HBM bytes per launch of the pass kernel, note), from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy-kernel
    calibration).  The bytes are None when the summary was taken from other
    kernel sources than the ones in the tree (its kernel_sources_sha256 line)."""
    path = PROFILES.get(mode)
    if path is None:
        return None, "no counter profile for this pass kernel"
    traffic, stamp = None, None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch:"):
                    traffic = float(line.split(":", 1)[1])
                elif line.startswith("kernel_sources_sha256:"):
                    stamp = line.split(":", 1)[1].strip()
    except OSError:
        return None, f"{os.path.basename(path)} missing"
    now = kernel_source_fingerprint()
    if stamp != now:
        return None, (f"{os.path.relpath(path, REPO)} was taken from other kernel sources "
                      f"({(stamp or 'unstamped')[:12]} vs {now[:12]}); traffic not reported")
    return traffic, f"from {os.path.relpath(path, REPO)} (kernel sources {now[:12]})"


def copy_bandwidth(mib: int, reps: int) -> float:
    """Best streaming-copy bandwidth (GB/s, read + write) of k_stream_copy."""
    import ctypes as C
    from cpim_amd import _native as N
    gbps = C.c_double(0)
    N.check(N.gpu_lib().cpim_measure_copy_bandwidth(0, mib << 20, reps, C.byref(gbps)))
    return gbps.value


def calibrate_copy(mib: int) -> None:
    bw = copy_bandwidth(mib, 1)
    print(json.dumps({"calibration_copy_bytes": mib << 20, "read_bytes": mib << 20,
                      "write_bytes": mib << 20, "gbps": bw}))


def fixpoint_bytes(desc) -> float:
    """Algorithmic bytes of a root fixpoint: both directions of every
    constraint, 8 W (min(|D(s)|, |D(t)|) + 3) each, at the start domains."""
    sizes = np.array([sum(bin(int(w)).count("1") for w in row) for row in desc.initial_domains()])
    sc = desc.con_scope.reshape(-1, 2)
    return float(2 * 8 * desc.words * np.sum(np.minimum(sizes[sc[:, 0]], sizes[sc[:, 1]]) + 3))


def single_fixpoint_one(desc, label: str, log) -> dict:
    """One root GAC fixpoint of desc in every mode: median host wall time
    (launches, syncs and readback included) and median device kernel time."""
    from cpim_amd import GModelAdapter
    from cpim_amd import _native as N
    modes = (("plain", N.CPIM_GAC_PLAIN), ("persistent", N.CPIM_GAC_PERSISTENT),
             ("async", N.CPIM_GAC_ASYNC), ("block", N.CPIM_GAC_BLOCK))
    alg_bytes = fixpoint_bytes(desc)
    out = {"workload": label, "algorithmic_bytes": alg_bytes}
    with GModelAdapter.Build(desc) as m:
        init = m.ReadDomains(0)
        for name, code in modes:
            wall, kern = [], []
            for rep in range(5 if name != "block" else 3):
                m.WriteDomains(0, init)
                t = time.perf_counter()
                st = m._gac(-1, code)
                dt = time.perf_counter() - t
                if rep:
                    wall.append(dt)
                    kern.append(st.kernel_ms)
            out[f"{name}_ms"] = 1000.0 * float(np.median(wall))
            out[f"{name}_kernel_ms"] = float(np.median(kern))
            out[f"{name}_revisions"] = int(st.revisions)
            out[f"{name}_rounds"] = int(st.iterations)
        out["deletions"] = int(st.deletions)
        out["inconsistent"] = bool(st.inconsistent)
    best = min(("plain", "persistent", "async", "block"), key=lambda k: out[f"{k}_kernel_ms"])
    ach = alg_bytes / (out[f"{best}_kernel_ms"] * 1e-3) / 1e9
    out["roofline"] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                       "frac": ach / HBM_PEAK_GBS,
                       "note": f"device kernel time of the fastest mode ({best})"}
    log(f"single fixpoint [{label[:40]}]: " + ", ".join(
        f"{k} {out[k + '_ms']:.3f} ms (kernel {out[k + '_kernel_ms']:.3f})"
        for k in ("plain", "persistent", "async", "block")) + f", deletions {out['deletions']}")
    return out


def cascade_fixpoint(log) -> dict:
    """The cascade instance of the single-fixpoint shape (t=0.892, 25 rounds,
    13% of the values deleted)."""
    from cpim_amd import workload
    return single_fixpoint_one(
        workload.single_large_cascade(), "root GAC fixpoint, random binary CSP n=2000 d=64 "
        "e=40000 tightness 0.892, 10% of values pre-removed (multi-round cascade)", log)


def single_fixpoint(log) -> dict:
    """Root GAC fixpoints of n=2000 d=64 e=40000 (10% of values pre-removed):
    the benchmark instance (t=0.25, one round, nothing deleted) and the
    cascade instance of the same shape (t=0.892, 25 rounds, 13% deleted)."""
    from cpim_amd import workload
    return {
        "benchmark": single_fixpoint_one(
            workload.single_large(), "root GAC fixpoint, random binary CSP n=2000 d=64 e=40000 "
            "tightness 0.25, 10% of values pre-removed", log),
        "cascade": cascade_fixpoint(log),
    }


def uneven_batch(count: int, log, per_instance_sample: int = 256) -> dict:
    """Root fixpoint + first 64 MAC nodes of `count` instances of the uneven
    batch in one launch (units heaviest first, the unit shape chosen by the
    autotuner on this batch before the timed run), against one launch per
    instance timed on a spread sample of the same instances."""
    from concurrent.futures import ThreadPoolExecutor
    from cpim_amd import Batch, autotune, workload
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        models = list(ex.map(workload.uneven_instance, range(count)))
    log(f"uneven batch: {count} instances generated in {time.perf_counter() - t:.1f} s")
    t = time.perf_counter()
    b = Batch(models)
    log(f"uneven batch: uploaded in {time.perf_counter() - t:.1f} s")
    try:
        tt = time.perf_counter()
        choice = autotune.tuned_batch(b, 64)
        tune_s = time.perf_counter() - tt
        b.run(64)  # warm-up
        t = time.perf_counter()
        res, kms = b.run(64, persistent=True)
        one = time.perf_counter() - t
        heavy = b.heavy_units()
    finally:
        b.close()
    nodes = sum(r.num_positive + r.num_negative for r in res)
    out = {"workload": f"{count} uneven instances (n~U[20,200], d in {{8..128}}, density~U[0.1,0.5], "
                       "phase-transition tightness): root fixpoint + first 64 MAC nodes each",
           "instances": count, "nodes": int(nodes),
           "one_launch_ms": 1000.0 * one, "kernel_ms": kms,
           "fixpoints_per_s": (count + nodes) / one,
           "heavy_units": heavy,
           "tuned_choice": choice, "tuning_s": tune_s,
           "tuning_timings_s": [(t["choice"], t.get("seconds")) for t in
                                autotune.load_cache().get(autotune.batch_key(models), {}).get("timings", [])]}
    if per_instance_sample > 0:
        # one launch per instance, on a spread sample
        idx = np.linspace(0, count - 1, min(per_instance_sample, count)).astype(np.int64)
        sub = [models[i] for i in idx]
        bs = Batch(sub)
        try:
            bs.run(64, persistent=False)
            t = time.perf_counter()
            rs, _ = bs.run(64, persistent=False)
            per = time.perf_counter() - t
        finally:
            bs.close()
        for i, r in zip(idx, rs):
            a = res[i]
            assert (a.status, a.num_positive, a.num_negative) == (r.status, r.num_positive, r.num_negative), \
                "one launch and one launch per instance disagree"
        sub_fix = len(sub) + sum(r.num_positive + r.num_negative for r in rs)
        out.update({"per_instance_sample": len(sub), "per_instance_fixpoints_per_s": sub_fix / per,
                    "speedup_vs_per_instance": ((count + nodes) / one) / (sub_fix / per)})
    log(f"uneven batch: one launch {out['one_launch_ms']:.1f} ms ({out['fixpoints_per_s']:,.0f} fixpoints/s, "
        f"{heavy} heavy units apart), per instance "
        f"{out.get('per_instance_fixpoints_per_s', float('nan')):,.0f} fixpoints/s, tuned {choice}")
    return out


def serve_poisson(reqs, count: int, rate: float, node_limit: int, nodes_per_step: int, log,
                  label: str) -> dict:
    """Poisson arrivals of `count` requests at `rate`/s into one server,
    cycling through the prepared distinct requests (each submission is a
    fresh request: its own upload, arena and search).  Reports solves/s over
    the whole run, the sustained completion rate over the last 80% of the
    arrival window (steady state: it matches the offered rate when the
    server keeps up), the largest backlog, and latency p50 / p99 from arrival
    to result."""
    from cpim_amd import Server
    rng = np.random.default_rng(0x5EED2026)
    arrive = np.cumsum(rng.exponential(1.0 / rate, count))
    srv = Server(max_slots=256, nodes_per_step=nodes_per_step)
    lat, done_t, limited, solved, unsat, backlog = [], [], 0, 0, 0, 0
    try:
        pending = {}
        nxt = 0
        t0 = time.perf_counter()
        while nxt < count or pending:
            now = time.perf_counter() - t0
            while nxt < count and arrive[nxt] <= now:
                pending[srv.submit(reqs[nxt % len(reqs)], node_limit=node_limit)] = arrive[nxt]
                nxt += 1
            backlog = max(backlog, len(pending))
            if not pending:
                time.sleep(max(0.0, arrive[nxt] - now))
                continue
            srv.step()
            done_at = time.perf_counter() - t0
            for rid in srv.poll():
                r = srv.result(rid)
                srv.release(rid)
                lat.append(done_at - pending.pop(rid))
                done_t.append(done_at)
                limited += r["status"] == 0
                solved += r["status"] == 1
                unsat += r["status"] == 2
        wall = time.perf_counter() - t0
        peak = srv.memory()["peak_bytes"]
    finally:
        srv.close()
    lat_ms = 1000.0 * np.array(lat)
    t_end = float(arrive[-1])
    t_beg = 0.2 * t_end
    done_t = np.array(done_t)
    sustained = float(((done_t >= t_beg) & (done_t <= t_end)).sum()) / max(1e-9, t_end - t_beg)
    out = {"solves_per_s": count / wall, "offered_per_s": rate, "sustained_per_s": sustained,
           "keeps_up": bool(sustained >= 0.9 * rate), "max_backlog": int(backlog),
           "p50_ms": float(np.percentile(lat_ms, 50)), "p99_ms": float(np.percentile(lat_ms, 99)),
           "requests": int(count), "distinct_requests": len(reqs), "arrival_window_s": t_end,
           "solved": int(solved), "unsat": int(unsat), "node_limited": int(limited),
           "peak_device_bytes": int(peak)}
    log(f"serving [{label}]: {count} requests, offered {rate:g}/s, sustained {sustained:.1f}/s "
        f"({'keeps up' if out['keeps_up'] else 'falls behind'}, backlog <= {backlog}), "
        f"{out['solves_per_s']:.1f} solves/s overall, p50 {out['p50_ms']:.1f} ms, p99 {out['p99_ms']:.1f} ms")
    return out


C3_DISTINCT = 16  # distinct C3-shaped requests generated (164 MB of relations each)


def serving(count: int, rate: float, c3_count: int, c3_rate: float, log) -> dict:
    """Poisson arrivals of MAC solve requests: small binary CSPs with arity-3
    tables, and requests of the benchmark SAC network's shape with tables
    under a node budget."""
    from concurrent.futures import ThreadPoolExecutor
    from cpim_amd import workload
    out = {}
    reqs = [workload.serving_request(i) for i in range(count)]
    small = serve_poisson(reqs, count, rate, 5000, 64, log, "small")
    small["workload"] = (f"{count} requests, Poisson arrivals at {rate:g}/s; each a MAC solve (node limit "
                         "5000) of n~U[20,40] d=16 binary CSP + 4 arity-3 tables of 500 tuples")
    out["small"] = small
    if c3_count > 0:
        distinct = min(c3_count, C3_DISTINCT)
        t = time.perf_counter()
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            reqs = list(ex.map(workload.serving_request_c3, range(distinct)))
        log(f"serving [c3]: {distinct} distinct requests generated in {time.perf_counter() - t:.1f} s")
        c3 = serve_poisson(reqs, c3_count, c3_rate, 64, 16, log, "c3")
        c3["workload"] = (f"{c3_count} requests ({distinct} distinct), Poisson arrivals at {c3_rate:g}/s; each "
                          "a MAC search with a budget of 64 nodes on a random binary CSP n=500 d=256 e=20000 "
                          "tightness 0.5 + 4 arity-3 tables of 500 tuples over 16 values")
        out["c3_shaped"] = c3
    return out


def main():
    args = parse()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (asked for --gpus {args.gpus}); "
              "singleton tests would be split across GPUs with one OR all-reduce of removal masks per SAC pass; that path is not built", file=sys.stderr)
        sys.exit(2)
    from cpim_amd import GModelAdapter, workload
    from cpim_amd.model import domains_to_sets

    log = lambda *a: print("[bench]", *a, file=sys.stderr, flush=True)  # noqa: E731
    if args.only:
        fn = {"single_fixpoint": lambda: single_fixpoint(log),
              "cascade_fixpoint": lambda: cascade_fixpoint(log),
              "uneven_batch": lambda: uneven_batch(args.batch_instances, log, args.per_instance_sample),
              "serving": lambda: serving(args.serve_requests, args.serve_rate,
                                         args.serve_c3_requests, args.serve_c3_rate, log)}[args.only]
        print(json.dumps({args.only: fn()}), flush=True)
        return
    t0 = time.time()
    desc = workload.batched_sac()
    n, W = desc.num_vars, desc.words
    log(f"workload generated n={n} d={desc.max_dom} e={desc.num_cons} in {time.time()-t0:.1f}s")
    m = GModelAdapter.Build(desc)
    log(f"model on device: {m.device_bytes/2**20:.0f} MiB")
    root = m.EnforceGAC_Auto()
    if root.inconsistent:
        raise SystemExit("root GAC wiped out; no singleton tests to run")
    dom = m.ReadDomains(0)
    tests = np.array([(v, a) for v, vals in enumerate(domains_to_sets(dom, desc.dom_size))
                      for a in vals], np.int32)
    ntests = len(tests)
    # algorithmic bytes of one singleton test (x = a): the seed revisions
    # t <- x read one support row and three domain word-sets each, plus the
    # domain copy: 8 W (4 deg(x) + n)
    deg = np.bincount(desc.con_scope, minlength=n)
    bytes_per_pass = float(np.sum(8.0 * W * (4.0 * deg[tests[:, 0]] + n)))

    for _ in range(args.warmup):
        m.SingletonTests(tests, mode=args.sac_mode)
    kernel_ms = []
    t_start = time.perf_counter()
    for _ in range(args.steps):
        fail, st = m.SingletonTests(tests, mode=args.sac_mode)
        kernel_ms.append(st.kernel_ms)
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1000.0 * elapsed / args.steps
    value = ntests * args.steps / elapsed
    avg_kernel_ms = float(np.mean(kernel_ms))
    achieved = bytes_per_pass / (avg_kernel_ms * 1e-3) / 1e9
    log(f"batched: {value:,.0f} fixpoints/s, {ms_per_step:.1f} ms/step, kernel {avg_kernel_ms:.1f} ms,"
        f" {int(fail.sum())} failed tests")

    plain = None
    if not args.skip_plain:
        k = min(args.plain_sample, ntests)
        idx = np.linspace(0, ntests - 1, k).astype(np.int64)
        m.SingletonTests(tests[idx[:8]], batched=False)
        tp = time.perf_counter()
        pf, _ = m.SingletonTests(tests[idx], batched=False)
        plain = k / (time.perf_counter() - tp)
        assert (pf == fail[idx]).all(), "plain and batched singleton tests disagree"
        log(f"plain one-per-launch: {plain:,.0f} fixpoints/s on {k} tests")

    cpu = None
    cpu_all = None
    if not args.skip_cpu:
        import oracle
        oracle.use_native()  # -O3 -march=native for this host, as the reference is built
        o = oracle.Oracle(desc)
        o.root_ac()
        k = min(args.cpu_sample, ntests)
        idx = np.linspace(0, ntests - 1, k).astype(np.int64)
        tc = time.perf_counter()
        cf, _ = o.singleton_tests(tests[idx], nthreads=1)
        cdt = time.perf_counter() - tc
        assert (cf == fail[idx]).all(), "device and CPU oracle disagree"
        cpu = {"value": k / cdt, "unit": "fixpoints/s", "cores": 1, "kind": "port",
               "sample": f"{k} singleton tests of the same pass (AC3bit restatement built "
                         f"-O3 -march=native, {cdt:.1f} s single-threaded)"}
        log(f"cpu oracle: {cpu['value']:.1f} fixpoints/s")
        # the same restatement on every thread of the host share
        th = max(1, min(args.cpu_threads, os.cpu_count() or 1))
        km = min(ntests, k * th)
        idx = np.linspace(0, ntests - 1, km).astype(np.int64)
        tc = time.perf_counter()
        cf, _ = o.singleton_tests(tests[idx], nthreads=th)
        cdt = time.perf_counter() - tc
        assert (cf == fail[idx]).all(), "device and CPU oracle disagree"
        cpu_all = {"value": km / cdt, "unit": "fixpoints/s", "cores": th, "kind": "port",
                   "sample": f"{km} singleton tests of the same pass, {th} threads, {cdt:.1f} s"}
        log(f"cpu oracle, {th} threads: {cpu_all['value']:.1f} fixpoints/s")

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "fixpoints/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic (seeded random binary CSP)",
        "config": {
            "workload": "batched SAC1 singleton tests, random binary CSP n=500 d=256 e=20000 "
                        "tightness 0.5, one SAC pass per step",
            "singleton_tests_per_step": ntests,
            "sac_kernel": "k_sliced_wide<4, nibble tables> (256 tests per workgroup)"
                          if args.sac_mode == "sliced" else "k_singleton_tests",
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "plain_fixpoints_per_s": plain,
        "speedup_vs_plain": (value / plain) if plain else None,
        "kernel_ms_per_step": avg_kernel_ms,
        "stream_copy_gbps": copy_bandwidth(1024, 5),
        # sliced: arc revisions of a 64-test group count once
        "revisions_per_s": st.revisions / (avg_kernel_ms * 1e-3),
    }
    if not args.skip_extras:
        for key, fn in (("single_fixpoint", lambda: single_fixpoint(log)),
                        ("uneven_batch", lambda: uneven_batch(args.batch_instances, log,
                                                              args.per_instance_sample)),
                        ("serving", lambda: serving(args.serve_requests, args.serve_rate,
                                                    args.serve_c3_requests, args.serve_c3_rate, log))):
            try:
                out[key] = fn()
            except Exception as exc:  # reported in the line, never hidden
                log(f"{key} failed: {exc!r}")
                out[key] = {"error": repr(exc)}
    traffic, note = profiled_traffic(args.sac_mode)
    out["roofline"]["traffic"] = traffic
    out["roofline"]["traffic_note"] = note
    if traffic:
        out["measured_dram_gbps"] = traffic / (avg_kernel_ms * 1e-3) / 1e9
    m.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
