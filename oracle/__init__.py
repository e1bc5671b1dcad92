"""Python view of the CPU oracle (liboracle.so).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this package, and only as the checker or the CPU baseline; the product
(cpim_amd) never does.  See oracle.h for the reference files each function
restates and for the brute-force pins.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from cpim_amd import _native as N

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

AC3BIT = 0
AC3 = 1
AC3MIX = 2


class OrcStats(C.Structure):
    _fields_ = [
        ("num_positive", C.c_int64),
        ("num_negative", C.c_int64),
        ("num_sol", C.c_int64),
        ("revisions", C.c_int64),
        ("deletions", C.c_int64),
        ("fixpoints", C.c_int64),
        ("time_out", C.c_int32),
        ("consistent", C.c_int32),
    ]


def build(force: bool = False) -> str:
    """Compile liboracle.so with g++ (portable x86-64-v2 code: the library
    travels to hosts whose CPU may differ from this one)."""
    src = os.path.join(_HERE, "oracle.cpp")
    if force or not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(
            os.path.getmtime(src), os.path.getmtime(os.path.join(_HERE, "oracle.h"))):
        cmd = ["g++", "-O3", "-march=x86-64-v2", "-std=c++17", "-fPIC", "-shared", "-Wall",
               "-o", LIB_PATH + ".tmp", src, "-lpthread"]
        subprocess.run(cmd, check=True)
        os.replace(LIB_PATH + ".tmp", LIB_PATH)
    return LIB_PATH


NATIVE_LIB_PATH = os.path.join(_HERE, "liboracle_native.so")


def build_native() -> str:
    """Compile liboracle_native.so with -O3 -march=native for the host it runs
    on: the timed CPU baseline, compiled for that host's CPU.  The portable
    liboracle.so stays the checker."""
    src = os.path.join(_HERE, "oracle.cpp")
    cmd = ["g++", "-O3", "-march=native", "-std=c++17", "-fPIC", "-shared", "-Wall",
           "-o", NATIVE_LIB_PATH + ".tmp", src, "-lpthread"]
    subprocess.run(cmd, check=True)
    os.replace(NATIVE_LIB_PATH + ".tmp", NATIVE_LIB_PATH)
    return NATIVE_LIB_PATH


def use_native() -> None:
    """Route every later Oracle through the -march=native build (bench's CPU
    baseline leg only); must be called before the first oracle call."""
    global _lib_path, _lib
    if _lib is not None and _lib_path != NATIVE_LIB_PATH:
        raise RuntimeError("the oracle library is already loaded")
    build_native()
    _lib_path = NATIVE_LIB_PATH


_lib = None
_lib_path = LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if _lib_path == LIB_PATH and not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(_lib_path)
        vp = C.c_void_p
        L.orc_create.restype = vp
        L.orc_create.argtypes = [C.POINTER(N.CModelDesc), C.c_int]
        L.orc_destroy.argtypes = [vp]
        L.orc_num_vars.argtypes = [vp]
        L.orc_dom_words.argtypes = [vp]
        L.orc_weights.argtypes = [vp, C.POINTER(C.c_float)]
        L.orc_root_ac.argtypes = [vp, N.p_u64, C.POINTER(OrcStats)]
        L.orc_sac1.argtypes = [vp, N.p_u64, C.POINTER(OrcStats)]
        L.orc_singleton_tests.argtypes = [vp, N.p_i32, C.c_int64, N.p_u8, C.c_int, C.POINTER(OrcStats)]
        L.orc_mac.argtypes = [vp, C.c_int64, C.c_int64, C.POINTER(OrcStats), N.p_i32,
                              C.POINTER(N.CNodeRecord), C.c_int32, N.p_i32]
        L.orc_mac_all.argtypes = [vp, C.c_int64, C.c_int64, C.c_int64, C.POINTER(OrcStats), N.p_i32,
                                  C.c_int64, C.POINTER(N.CNodeRecord), C.c_int32, N.p_i32]
        L.orc_bf_mac_all.argtypes = [C.POINTER(N.CModelDesc), C.c_int64, C.c_int64, C.POINTER(OrcStats),
                                     N.p_i32, C.c_int64, C.POINTER(N.CNodeRecord), C.c_int32, N.p_i32]
        L.orc_dom_hash.restype = C.c_uint64
        L.orc_dom_hash.argtypes = [N.p_u64, C.c_int, C.c_int]
        L.orc_bf_gac.argtypes = [C.POINTER(N.CModelDesc), N.p_u64]
        L.orc_bf_sac.argtypes = [C.POINTER(N.CModelDesc), N.p_u64]
        L.orc_bf_count_solutions.restype = C.c_int64
        L.orc_bf_count_solutions.argtypes = [C.POINTER(N.CModelDesc), C.c_int64, N.p_i32]
        L.orc_bf_mac.argtypes = [C.POINTER(N.CModelDesc), C.c_int64, C.POINTER(OrcStats), N.p_i32,
                                 C.POINTER(N.CNodeRecord), C.c_int32, N.p_i32]
        L.orc_set_fault.argtypes = [C.c_int]
        L.orc_set_var_heuristic.argtypes = [C.c_int]
        _lib = L
    return _lib


def _ptr(a, t):
    return a.ctypes.data_as(C.POINTER(t))


@dataclass
class MacResult:
    num_sol: int
    num_positive: int
    num_negative: int
    time_out: bool
    solution: Optional[np.ndarray]
    trace: List[Tuple[int, int, int, int, int]]
    stats: OrcStats


class Oracle:
    """One restated Network (+ AC3bit or generic AC3) for a model."""

    def __init__(self, desc, algo: int = AC3BIT):
        self.desc = desc
        self._c = desc.c_desc()
        self.h = lib().orc_create(C.byref(self._c), algo)
        self.n = desc.num_vars
        self.W = desc.words

    def close(self):
        if self.h:
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def root_ac(self) -> Tuple[bool, np.ndarray, OrcStats]:
        out = np.zeros((self.n, self.W), np.uint64)
        st = OrcStats()
        ok = lib().orc_root_ac(self.h, _ptr(out, C.c_uint64), C.byref(st))
        return bool(ok), out, st

    def sac1(self) -> Tuple[bool, np.ndarray, OrcStats]:
        out = np.zeros((self.n, self.W), np.uint64)
        st = OrcStats()
        ok = lib().orc_sac1(self.h, _ptr(out, C.c_uint64), C.byref(st))
        return bool(ok), out, st

    def singleton_tests(self, tests: np.ndarray, nthreads: int = 1) -> Tuple[np.ndarray, OrcStats]:
        tests = np.ascontiguousarray(tests, np.int32)
        fail = np.zeros(len(tests), np.uint8)
        st = OrcStats()
        lib().orc_singleton_tests(self.h, _ptr(tests, C.c_int32), len(tests), _ptr(fail, C.c_uint8),
                                  nthreads, C.byref(st))
        return fail, st

    def weights(self) -> np.ndarray:
        """Constraint weights (raised by AC3 on each wipe-out)."""
        k = lib().orc_weights(self.h, None)
        w = np.zeros(max(1, k), np.float32)
        lib().orc_weights(self.h, w.ctypes.data_as(C.POINTER(C.c_float)))
        return w[:k]

    def mac(self, time_limit_ms: int = 0, node_limit: int = 0, trace_cap: int = 0) -> MacResult:
        sol = np.full(self.n, -1, np.int32)
        tr = (N.CNodeRecord * max(1, trace_cap))()
        tl = C.c_int32(0)
        st = OrcStats()
        lib().orc_mac(self.h, time_limit_ms, node_limit, C.byref(st), _ptr(sol, C.c_int32), tr,
                      trace_cap, C.byref(tl))
        trace = [(r.positive, r.var, r.value, r.consistent, r.dom_hash)
                 for r in tr[:min(tl.value, trace_cap)]]
        return MacResult(int(st.num_sol), int(st.num_positive), int(st.num_negative),
                         bool(st.time_out), sol if st.num_sol else None, trace, st)


@dataclass
class MacAllResult:
    num_sol: int
    num_positive: int
    num_negative: int
    time_out: bool
    solutions: np.ndarray  # [min(num_sol, sol_cap), n]
    trace: List[Tuple[int, int, int, int, int]]


def _mac_all_result(st, sols, n, cap, tr, tl, trace_cap):
    k = min(int(st.num_sol), cap)
    trace = [(r.positive, r.var, r.value, r.consistent, r.dom_hash)
             for r in tr[:min(tl.value, trace_cap)]]
    return MacAllResult(int(st.num_sol), int(st.num_positive), int(st.num_negative),
                        bool(st.time_out), sols[:k * n].reshape(k, n), trace)


def mac_all(o: "Oracle", max_solutions: int, node_limit: int = 0, sol_cap: int = 1024,
            trace_cap: int = 0) -> MacAllResult:
    """MAC that counts every solution up to max_solutions (orc_mac_all)."""
    sols = np.full(max(1, sol_cap) * o.n, -1, np.int32)
    tr = (N.CNodeRecord * max(1, trace_cap))()
    tl = C.c_int32(0)
    st = OrcStats()
    lib().orc_mac_all(o.h, 0, node_limit, max_solutions, C.byref(st), _ptr(sols, C.c_int32), sol_cap,
                      tr, trace_cap, C.byref(tl))
    return _mac_all_result(st, sols, o.n, sol_cap, tr, tl, trace_cap)


def bf_mac_all(desc, max_solutions: int, node_limit: int = 0, sol_cap: int = 1024,
               trace_cap: int = 0) -> MacAllResult:
    """The brute-force closure MAC with the same enumeration (orc_bf_mac_all)."""
    c = desc.c_desc()
    sols = np.full(max(1, sol_cap) * desc.num_vars, -1, np.int32)
    tr = (N.CNodeRecord * max(1, trace_cap))()
    tl = C.c_int32(0)
    st = OrcStats()
    rc = lib().orc_bf_mac_all(C.byref(c), node_limit, max_solutions, C.byref(st), _ptr(sols, C.c_int32),
                              sol_cap, tr, trace_cap, C.byref(tl))
    if rc < 0:
        raise ValueError("the brute-force MAC has no dom/wdeg form")
    return _mac_all_result(st, sols, desc.num_vars, sol_cap, tr, tl, trace_cap)


def dom_hash(dom: np.ndarray) -> int:
    dom = np.ascontiguousarray(dom, np.uint64)
    return int(lib().orc_dom_hash(_ptr(dom, C.c_uint64), dom.shape[0], dom.shape[1]))


def bf_gac(desc) -> Tuple[bool, np.ndarray]:
    c = desc.c_desc()
    out = np.zeros((desc.num_vars, desc.words), np.uint64)
    ok = lib().orc_bf_gac(C.byref(c), _ptr(out, C.c_uint64))
    return bool(ok), out


def bf_sac(desc) -> Tuple[bool, np.ndarray]:
    c = desc.c_desc()
    out = np.zeros((desc.num_vars, desc.words), np.uint64)
    ok = lib().orc_bf_sac(C.byref(c), _ptr(out, C.c_uint64))
    return bool(ok), out


def bf_count_solutions(desc, limit: int = 0) -> Tuple[int, np.ndarray]:
    c = desc.c_desc()
    first = np.full(desc.num_vars, -1, np.int32)
    cnt = lib().orc_bf_count_solutions(C.byref(c), limit, _ptr(first, C.c_int32))
    return int(cnt), first


def bf_mac(desc, node_limit: int = 0, trace_cap: int = 0) -> MacResult:
    c = desc.c_desc()
    sol = np.full(desc.num_vars, -1, np.int32)
    tr = (N.CNodeRecord * max(1, trace_cap))()
    tl = C.c_int32(0)
    st = OrcStats()
    rc = lib().orc_bf_mac(C.byref(c), node_limit, C.byref(st), _ptr(sol, C.c_int32), tr,
                          trace_cap, C.byref(tl))
    if rc < 0:
        raise ValueError("the brute-force MAC has no dom/wdeg form")
    trace = [(r.positive, r.var, r.value, r.consistent, r.dom_hash)
             for r in tr[:min(tl.value, trace_cap)]]
    return MacResult(int(st.num_sol), int(st.num_positive), int(st.num_negative),
                     bool(st.time_out), sol if st.num_sol else None, trace, st)


class var_heuristic:
    """Context manager selecting the oracle's variable ordering
    (0 = dom-min, 1 = dom/deg, 2 = dom/wdeg) for orc_mac and orc_bf_mac
    (which has no dom/wdeg form)."""

    def __init__(self, varh: int):
        self.varh = varh

    def __enter__(self):
        lib().orc_set_var_heuristic(self.varh)
        return self

    def __exit__(self, *exc):
        lib().orc_set_var_heuristic(0)


class fault:
    """Context manager planting a mistake in the oracle (for pin tests)."""

    def __init__(self, kind: int):
        self.kind = kind

    def __enter__(self):
        lib().orc_set_fault(self.kind)
        return self

    def __exit__(self, *exc):
        lib().orc_set_fault(0)
        return False
