"""Python mirror of the reference's GPU entry points, over the C-ABI.

    GModelAdapter.Build(model, options) -> GModel   src/model/gmodel_adapter.cu:44-312
    GModel.EnforceGAC / EnforceGAC_Persistent      src/GModel.cu:1110-1391
    GModel level API                               src/GModel.cu:1397-1701
    GModelSolver(model).Solve(time_limit_ms)       src/GModelSolver.cu:31-189
    CModel enforceGAC / enforceSAC / solve         src/cuSAC.cu:157-413

Names, argument meanings, return conventions and exceptions follow the
reference: level-API misuse raises RuntimeError with the reference's message,
EnforceGAC never raises on inconsistency (it reports it in GacStats), the
solver's times are in seconds.  Every call runs on the GPU through
libcpim_gpu.so; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _native as N
from .model import ModelDesc


_SAC_MODES = {"auto": N.CPIM_SAC_BATCHED, "plain": N.CPIM_SAC_PLAIN,
              "workgroup": N.CPIM_SAC_WORKGROUP, "sliced": N.CPIM_SAC_SLICED}


_VAR_HEURISTICS = {"dom": N.CPIM_VARH_DOM_MIN, "dom/deg": N.CPIM_VARH_DOM_DEG,
                   "dom/wdeg": N.CPIM_VARH_DOM_WDEG}


def _var_heuristic(name: str) -> int:
    if name not in _VAR_HEURISTICS:
        raise ValueError("heuristic must be one of %s" % sorted(_VAR_HEURISTICS))
    return _VAR_HEURISTICS[name]


def _sac_mode(batched: bool, mode: Optional[str]) -> int:
    if mode is not None:
        if mode not in _SAC_MODES:
            raise ValueError(f"unknown SAC mode {mode!r} (one of {sorted(_SAC_MODES)})")
        return _SAC_MODES[mode]
    return N.CPIM_SAC_BATCHED if batched else N.CPIM_SAC_PLAIN


@dataclass
class GModelOptions:
    device_id: int = 0
    enable_prefetch: bool = False
    skip_non_binary: bool = True
    skip_conflicts: bool = False

    def c(self) -> N.CGModelOptions:
        o = N.CGModelOptions()
        o.device_id = self.device_id
        o.enable_prefetch = int(self.enable_prefetch)
        o.skip_non_binary = int(self.skip_non_binary)
        o.skip_conflicts = int(self.skip_conflicts)
        return o


@dataclass
class GacStats:
    iterations: int = 0
    deletions: int = 0
    inconsistent: bool = False
    converged: bool = False
    revisions: int = 0
    kernel_ms: float = 0.0  # device time of the fixpoint's kernels


@dataclass
class GpuSearchStatistics:
    num_positive: int = 0
    num_negative: int = 0
    num_solutions: int = 0
    gac_iterations: int = 0
    gac_deletions: int = 0
    time_out: bool = False
    unsolvable: bool = False
    solve_time: float = 0.0
    gac_time: float = 0.0


@dataclass
class SacStats:
    consistent: bool = True
    passes: int = 0
    singleton_tests: int = 0
    removed: int = 0
    revisions: int = 0
    kernel_ms: float = 0.0


def _i32():
    return C.c_int32(0)


class GModel:
    """Device-resident model with its domain level stack (owns the handle)."""

    def __init__(self, handle, desc: ModelDesc):
        self._h = C.c_void_p(handle)
        self.desc = desc
        info = N.CGModelInfo()
        N.check(N.gpu_lib().cpim_gmodel_get_info(self._h, C.byref(info)))
        self.num_vars = info.num_vars
        self.num_constraints = info.num_constraints
        self.num_active = info.num_active
        self.max_dom_size = info.max_dom_size
        self.dom_words = info.dom_words
        self.bit_dom_int_size = info.bit_dom_int_size
        self.max_depth = info.max_depth
        self.subscription_size = info.subscription_size
        self.device_bytes = info.device_bytes

    # -- lifetime --------------------------------------------------------------
    def close(self):
        if self._h is not None and self._h.value:
            N.gpu_lib().cpim_gmodel_destroy(self._h)
        self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        if self._h is None:
            raise RuntimeError("GModel has been closed")
        return self._h

    # -- propagation -------------------------------------------------------------
    def _gac(self, assigned_var: int, mode: int) -> GacStats:
        st = N.CGacStats()
        N.check(N.gpu_lib().cpim_gmodel_enforce_gac(self.handle, assigned_var, mode, C.byref(st)))
        return GacStats(st.iterations, st.deletions, bool(st.inconsistent), bool(st.converged),
                        st.revisions, st.kernel_ms)

    def EnforceGAC(self, verbose: bool = True, assigned_var: int = -1) -> GacStats:
        """Host-driven rounds (one launch + sync per round), like the reference."""
        st = self._gac(assigned_var, N.CPIM_GAC_PLAIN)
        if verbose:
            print(f"[GAC-Bitmap] iterations={st.iterations} deletions={st.deletions} "
                  f"inconsistent={'true' if st.inconsistent else 'false'}")
        return st

    def EnforceGAC_Persistent(self, verbose: bool = True, assigned_var: int = -1) -> GacStats:
        """All rounds in one launch of co-resident workgroups, a grid barrier
        between rounds (src/GModel.cu:1222-1391)."""
        st = self._gac(assigned_var, N.CPIM_GAC_PERSISTENT)
        if verbose:
            print(f"[GAC-Persistent] iterations={st.iterations} deletions={st.deletions} "
                  f"inconsistent={'true' if st.inconsistent else 'false'} "
                  f"converged={'true' if st.converged else 'false'}")
        return st

    def EnforceGAC_Legacy(self, verbose: bool = True) -> GacStats:
        """Deprecated in the reference (include/GModel.cuh:199-201: the
        event-list fixpoint of CsCheckMainKernel, one launch per round from
        every constraint).  Its closure is the host-driven rounds' one."""
        st = self._gac(-1, N.CPIM_GAC_PLAIN)
        if verbose:
            print(f"[GAC-Legacy] iterations={st.iterations} deletions={st.deletions} "
                  f"inconsistent={'true' if st.inconsistent else 'false'}")
        return st

    def EnforceGAC_Async(self, assigned_var: int = -1) -> GacStats:
        """Whole GPU, one launch, asynchronous work queue."""
        return self._gac(assigned_var, N.CPIM_GAC_ASYNC)

    def EnforceGAC_Block(self, assigned_var: int = -1) -> GacStats:
        """One workgroup, domains in LDS (one launch)."""
        return self._gac(assigned_var, N.CPIM_GAC_BLOCK)

    def EnforceGAC_Auto(self, assigned_var: int = -1) -> GacStats:
        return self._gac(assigned_var, N.CPIM_GAC_AUTO)

    # -- levels --------------------------------------------------------------
    def CreateNewLevel(self) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_create_new_level(self.handle, C.byref(v)))
        return v.value

    def BackToLevel(self, level: int) -> None:
        N.check(N.gpu_lib().cpim_gmodel_back_to_level(self.handle, level))

    def GetCurrentLevel(self) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_current_level(self.handle, C.byref(v)))
        return v.value

    def AssignValue(self, var: int, value: int, level: int) -> bool:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_assign_value(self.handle, var, value, level, C.byref(v)))
        return bool(v.value)

    def RemoveValue(self, var: int, value: int, level: int) -> bool:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_remove_value(self.handle, var, value, level, C.byref(v)))
        return bool(v.value)

    def GetDomainSize(self, var: int, level: int) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_domain_size(self.handle, var, level, C.byref(v)))
        return v.value

    def GetMinDomainVar(self, level: int) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_min_domain_var(self.handle, level, C.byref(v)))
        return v.value

    def IsAssigned(self, var: int, level: int) -> bool:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_is_assigned(self.handle, var, level, C.byref(v)))
        return bool(v.value)

    def IsFullyAssigned(self, level: int) -> bool:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_is_fully_assigned(self.handle, level, C.byref(v)))
        return bool(v.value)

    def GetAssignedValue(self, var: int, level: int) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_assigned_value(self.handle, var, level, C.byref(v)))
        return v.value

    def GetFirstValue(self, var: int, level: int) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_first_value(self.handle, var, level, C.byref(v)))
        return v.value

    def GetNextValue(self, var: int, value: int, level: int) -> int:
        v = _i32()
        N.check(N.gpu_lib().cpim_gmodel_next_value(self.handle, var, value, level, C.byref(v)))
        return v.value

    def Weights(self, reset: bool = False) -> np.ndarray:
        """dom/wdeg constraint weights (one per propagated constraint, in
        constraint order).  They start at 1 with the model and persist across
        solves, as Tabular::weight does in the reference's network
        (Network.cpp:303, AC3.cpp:122-126); reset=True sets them back to 1
        after reading them."""
        lib = N.gpu_lib()
        k = C.c_int64(0)
        N.check(lib.cpim_gmodel_weights(self._h, None, 0, C.byref(k), 0))
        w = np.zeros(max(1, k.value), np.float32)
        N.check(lib.cpim_gmodel_weights(self._h, w.ctypes.data_as(C.POINTER(C.c_float)), k.value,
                                        C.byref(k), 1 if reset else 0))
        return w[:k.value]

    def ReadDomains(self, level: Optional[int] = None) -> np.ndarray:
        """Host copy of one level ([num_vars, dom_words] uint64)."""
        if level is None:
            level = self.GetCurrentLevel()
        out = np.zeros((self.num_vars, self.dom_words), np.uint64)
        N.check(N.gpu_lib().cpim_gmodel_read_domains(self.handle, level,
                                                     out.ctypes.data_as(N.p_u64)))
        return out

    def WriteDomains(self, level: int, dom: np.ndarray) -> None:
        dom = np.ascontiguousarray(dom, np.uint64)
        N.check(N.gpu_lib().cpim_gmodel_write_domains(self.handle, level, dom.ctypes.data_as(N.p_u64)))

    # -- SAC / singleton tests ---------------------------------------------------
    def EnforceSAC(self, batched: bool = True, mode: Optional[str] = None) -> SacStats:
        """SAC closure.  mode: "auto" (default when batched), "sliced" (64 tests per
        workgroup, binary networks), "workgroup" (one test per workgroup) or "plain"."""
        st = N.CSacStats()
        N.check(N.gpu_lib().cpim_gmodel_enforce_sac(self.handle, _sac_mode(batched, mode),
                                                    C.byref(st)))
        return SacStats(bool(st.consistent), st.passes, st.singleton_tests, st.removed, st.revisions,
                        st.kernel_ms)

    def SingletonTests(self, tests: np.ndarray, batched: bool = True, mode: Optional[str] = None):
        """fail[i] = 1 iff GAC with D(x_i) = {a_i} wipes out (modes as in EnforceSAC)."""
        tests = np.ascontiguousarray(tests, np.int32).reshape(-1, 2)
        fail = np.zeros(len(tests), np.uint8)
        st = N.CSacStats()
        N.check(N.gpu_lib().cpim_gmodel_singleton_tests(
            self.handle, tests.ctypes.data_as(N.p_i32), len(tests), _sac_mode(batched, mode),
            fail.ctypes.data_as(N.p_u8), C.byref(st)))
        return fail, SacStats(bool(st.consistent), st.passes, st.singleton_tests, st.removed,
                              st.revisions, st.kernel_ms)

    # -- device search -------------------------------------------------------------
    def Solve(self, time_limit_ms: int = 0, node_limit: int = 0, trace_capacity: int = 0,
              heuristic: str = "dom"):
        """Device MAC (binary branching, CPU MAC semantics) from level 0.

        heuristic: "dom" (smallest domain, MAC.cpp:356-364), "dom/deg"
        (domain size over neighbour count, MAC.cpp:371-384) or "dom/wdeg"
        (domain size over failure-weighted degree, MAC.cpp:386-411; its
        fixpoints run in AC3's order on one wave)."""
        stats, sols, trace = self.SolveAll(1, 1, time_limit_ms, node_limit, trace_capacity, heuristic)
        return stats, (sols[0] if len(sols) else None), trace

    def SolveAll(self, max_solutions: int, solution_capacity: int, time_limit_ms: int = 0,
                 node_limit: int = 0, trace_capacity: int = 0, heuristic: str = "dom"):
        """Device MAC that counts up to max_solutions solutions (each found
        solution is then treated as a failed node) and keeps the first
        solution_capacity of them.  Returns (stats, solutions [k, n], trace)."""
        o = N.CSolveOpts()
        o.time_limit_ms = time_limit_ms
        o.node_limit = node_limit
        o.branching = N.CPIM_BRANCH_BINARY
        o.trace_capacity = trace_capacity
        o.var_heuristic = _var_heuristic(heuristic)
        o.max_solutions = max_solutions
        o.solution_capacity = max(1, solution_capacity)
        st = N.CSearchStats()
        sol = np.full(max(1, solution_capacity) * self.num_vars, -1, np.int32)
        tr = (N.CNodeRecord * max(1, trace_capacity))()
        tl = C.c_int32(0)
        N.check(N.gpu_lib().cpim_gmodel_solve(self.handle, C.byref(o), C.byref(st),
                                              sol.ctypes.data_as(N.p_i32), tr, C.byref(tl)))
        stats = GpuSearchStatistics(st.num_positive, st.num_negative, st.num_solutions,
                                    st.gac_iterations, st.gac_deletions, bool(st.time_out),
                                    bool(st.unsolvable), st.solve_time, st.gac_time)
        trace = [(r.positive, r.var, r.value, r.consistent, r.dom_hash)
                 for r in tr[:min(tl.value, trace_capacity)]]
        k = min(int(st.num_solutions), max(1, solution_capacity))
        return stats, sol[:k * self.num_vars].reshape(k, self.num_vars), trace


class GModelAdapter:
    @staticmethod
    def Build(model: ModelDesc, options: Optional[GModelOptions] = None) -> GModel:
        options = options or GModelOptions()
        h = C.c_void_p()
        cd = model.c_desc()
        co = options.c()
        N.check(N.gpu_lib().cpim_gmodel_create(C.byref(cd), C.byref(co), C.byref(h)))
        return GModel(h.value, model)


class GModelSolver:
    """GModelSolver (include/GModelSolver.h:32-70).  The default search is the
    CPU MAC's binary branching run on the device; branching="dway" replays the
    reference GModelSolver's d-way search (src/GModelSolver.cu:104-189) through
    the host-visible level API."""

    def __init__(self, model: GModel, verbose: bool = False, branching: str = "binary",
                 heuristic: str = "dom"):
        if model is None:
            raise RuntimeError("[GModelSolver] model is null!")
        if branching not in ("binary", "dway"):
            raise ValueError("branching must be 'binary' or 'dway'")
        _var_heuristic(heuristic)
        if branching == "dway" and heuristic != "dom":
            raise ValueError("the d-way search orders variables by smallest domain only")
        self.model_ = model
        self.verbose_ = verbose
        self.branching = branching
        self.heuristic = heuristic
        self.find_all_solutions_ = False
        self.max_solutions_ = 1
        self.solutions_: List[List[int]] = []
        self.trace: list = []

    def SetFindAllSolutions(self, find_all: bool) -> None:
        self.find_all_solutions_ = find_all
        if find_all:
            self.max_solutions_ = 1000000

    def SetMaxSolutions(self, max_solutions: int) -> None:
        self.max_solutions_ = max_solutions

    def GetSolution(self) -> List[int]:
        return list(self.solutions_[0]) if self.solutions_ else []

    def GetAllSolutions(self) -> List[List[int]]:
        return [list(s) for s in self.solutions_]

    def Solve(self, time_limit: int = 0, node_limit: int = 0,
              trace_capacity: int = 0) -> GpuSearchStatistics:
        """With SetFindAllSolutions(True) the binary search enumerates, up to
        SetMaxSolutions (1,000,000 by default then), as the reference's search
        stops (src/GModelSolver.cu:125-128): every solution is counted, then
        treated as a failed node, and GetAllSolutions() returns the first
        ones, as many as fit in SOLUTION_BYTES of device memory
        (num_solutions counts them all)."""
        self.solutions_ = []
        if self.branching == "binary":
            k = max(1, self.max_solutions_) if self.find_all_solutions_ else 1
            cap = max(1, min(k, self.SOLUTION_BYTES // (4 * max(1, self.model_.num_vars))))
            st, sols, self.trace = self.model_.SolveAll(k, cap, time_limit, node_limit, trace_capacity,
                                                        self.heuristic)
            self.solutions_ = [[int(x) for x in row] for row in sols]
            return st
        return self._solve_dway(time_limit)

    # device memory for the solutions kept by an enumerating search
    SOLUTION_BYTES = 256 << 20

    # d-way search of the reference GModelSolver, node for node.
    def _solve_dway(self, time_limit: int) -> GpuSearchStatistics:
        import time
        m = self.model_
        stats = GpuSearchStatistics()
        t0 = time.monotonic()
        g = m.EnforceGAC(False)
        stats.gac_iterations += g.iterations
        stats.gac_deletions += g.deletions
        if g.inconsistent:
            stats.unsolvable = True
            stats.solve_time = time.monotonic() - t0
            return stats
        if m.IsFullyAssigned(0):
            self._extract(0)
            stats.num_solutions = 1
            stats.solve_time = time.monotonic() - t0
            return stats

        def search(level: int) -> bool:
            if time_limit > 0 and (time.monotonic() - t0) * 1000 >= time_limit:
                stats.time_out = True
                return True
            var = m.GetMinDomainVar(level)
            if var == -1:
                if m.IsFullyAssigned(level):
                    self._extract(level)
                    return not self.find_all_solutions_ or len(self.solutions_) >= self.max_solutions_
                return False
            value = m.GetFirstValue(var, level)
            while value != -1:
                new_level = m.CreateNewLevel()
                stats.num_positive += 1
                m.AssignValue(var, value, new_level)
                gs = m.EnforceGAC(False)
                stats.gac_iterations += gs.iterations
                stats.gac_deletions += gs.deletions
                if gs.inconsistent:
                    m.BackToLevel(level)
                    stats.num_negative += 1
                    value = m.GetNextValue(var, value, level)
                    continue
                stop = search(new_level)
                m.BackToLevel(level)
                if stop:
                    return True
                value = m.GetNextValue(var, value, level)
            return False

        search(0)
        stats.solve_time = time.monotonic() - t0
        stats.num_solutions = len(self.solutions_)
        return stats

    def _extract(self, level: int) -> None:
        m = self.model_
        sol = []
        for v in range(m.num_vars):
            a = m.GetAssignedValue(v, level)
            if a == -1:
                raise RuntimeError(f"[GModelSolver::ExtractSolution] Variable {v} is not assigned "
                                   f"at level {level}")
            sol.append(a)
        self.solutions_.append(sol)


@dataclass
class SearchStatistics:
    """SearchStatistics (include/Solver.h:71-82); times in milliseconds."""
    num_sol: int = 0
    num_positive: int = 0
    num_negative: int = 0
    nodes: int = 0
    build_time: int = 0
    solve_time: int = 0
    total_time: int = 0
    time_out: bool = False
    pass_: bool = False


class CModel:
    """Legacy CModel entry points (include/cuSAC.cuh:422-447) served by the same
    engine: enforceGAC() is a full fixpoint, enforceGAC(var, type) propagates
    the decision on var (type 1 = assignment, 0 = removal), enforceSAC() is the
    SAC closure (the reference's body is empty), solve(ms) is the device MAC."""

    def __init__(self, model: ModelDesc, options: Optional[GModelOptions] = None):
        self.g = GModelAdapter.Build(model, options)
        self.kNumVars = self.g.num_vars
        self.kNumTabs = self.g.num_constraints
        self.kMaxDomSize = self.g.max_dom_size
        self.kBitDomIntSize = self.g.bit_dom_int_size
        self.statistics_ = SearchStatistics()

    def enforceGAC(self, var: int = -1, type: int = 1) -> bool:
        st = self.g.EnforceGAC_Auto(var if var >= 0 else -1)
        return not st.inconsistent

    def enforceSAC(self) -> bool:
        return self.g.EnforceSAC(batched=True).consistent

    def CreateNewLevel(self) -> int:
        return self.g.CreateNewLevel()

    def BackLevel(self) -> int:
        lvl = self.g.GetCurrentLevel() - 1
        self.g.BackToLevel(lvl)
        return lvl

    def solve(self, time_limits: float) -> SearchStatistics:
        st, sol, _ = self.g.Solve(int(time_limits))
        self.statistics_ = SearchStatistics(num_sol=st.num_solutions, num_positive=st.num_positive,
                                            num_negative=st.num_negative,
                                            solve_time=int(st.solve_time * 1000),
                                            time_out=st.time_out)
        self.solution = sol
        return self.statistics_


# ---------------------------------------------------------------------------
# Batches and serving
# ---------------------------------------------------------------------------
@dataclass
class BatchResult:
    root_consistent: bool
    status: int  # 0 node limit, 1 solution, 2 unsat
    num_positive: int
    num_negative: int
    trace: list = field(default_factory=list)


class Batch:
    """Many uneven instances: root GAC + MAC nodes for all of them in one launch."""

    def __init__(self, models: List[ModelDesc], options: Optional[GModelOptions] = None):
        self.models = models
        self._cdescs = (N.CModelDesc * max(1, len(models)))()
        for i, m in enumerate(models):
            self._cdescs[i] = m.c_desc()
        h = C.c_void_p()
        co = (options or GModelOptions()).c()
        N.check(N.gpu_lib().cpim_batch_create(self._cdescs, len(models), C.byref(co), C.byref(h)))
        self._h = h

    def close(self):
        if self._h is not None and self._h.value:
            N.gpu_lib().cpim_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, node_limit: int, persistent: bool = True, trace_capacity: int = 0):
        n = len(self.models)
        res = (N.CBatchResult * max(1, n))()
        tr = (N.CNodeRecord * max(1, n * trace_capacity))()
        ms = C.c_double(0)
        mode = N.CPIM_GAC_PERSISTENT if persistent else N.CPIM_GAC_PLAIN
        N.check(N.gpu_lib().cpim_batch_run(self._h, node_limit, mode, res, tr, trace_capacity,
                                           C.byref(ms)))
        out = []
        for i in range(n):
            r = res[i]
            tl = min(r.trace_len, trace_capacity)
            trace = [(x.positive, x.var, x.value, x.consistent, x.dom_hash)
                     for x in tr[i * trace_capacity:i * trace_capacity + tl]]
            out.append(BatchResult(bool(r.root_consistent), r.status, r.num_positive,
                                   r.num_negative, trace))
        return out, ms.value

    def heavy_units(self) -> int:
        """Units of the last run that ran in the separate 1024-thread launch
        for the batch's heaviest searches."""
        k = C.c_int32(0)
        N.check(N.gpu_lib().cpim_batch_heavy_units(self._h, C.byref(k)))
        return k.value

    def root_domains(self, index: int) -> np.ndarray:
        m = self.models[index]
        out = np.zeros((m.num_vars, m.words), np.uint64)
        N.check(N.gpu_lib().cpim_batch_read_root(self._h, index, out.ctypes.data_as(N.p_u64)))
        return out


class Server:
    """Continuous batching of MAC solve requests on one GPU."""

    def __init__(self, max_slots: int = 256, hbm_budget_bytes: int = 64 << 30,
                 nodes_per_step: int = 64, device_id: int = 0):
        o = N.CServerOpts()
        o.device_id = device_id
        o.max_slots = max_slots
        o.hbm_budget_bytes = hbm_budget_bytes
        o.nodes_per_step = nodes_per_step
        h = C.c_void_p()
        N.check(N.gpu_lib().cpim_server_create(C.byref(o), C.byref(h)))
        self._h = h
        self._n = {}

    def close(self):
        if self._h is not None and self._h.value:
            N.gpu_lib().cpim_server_destroy(self._h)
        self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, model: ModelDesc, node_limit: int = 0) -> int:
        cd = model.c_desc()
        rid = C.c_int64(0)
        N.check(N.gpu_lib().cpim_server_submit(self._h, C.byref(cd), node_limit, C.byref(rid)))
        self._n[rid.value] = model.num_vars
        return rid.value

    def step(self) -> int:
        a = C.c_int32(0)
        N.check(N.gpu_lib().cpim_server_step(self._h, C.byref(a)))
        return a.value

    def poll(self, cap: int = 4096) -> List[int]:
        """Ids of the requests finished since the last poll."""
        ids = (C.c_int64 * max(1, cap))()
        n = C.c_int32(0)
        N.check(N.gpu_lib().cpim_server_poll(self._h, ids, cap, C.byref(n)))
        return list(ids[:n.value])

    def result(self, rid: int):
        """(done, status, num_positive, num_negative, solution or None, steps)."""
        r = N.CRequestResult()
        sol = np.full(max(1, self._n.get(rid, 0)), -1, np.int32)  # unknown ids fail in the call
        code = N.gpu_lib().cpim_server_result(self._h, rid, C.byref(r), sol.ctypes.data_as(N.p_i32))
        if code == N.CPIM_ERR_BUSY:
            return None
        N.check(code)
        return dict(status=r.status, num_positive=r.num_positive, num_negative=r.num_negative,
                    solution=sol if r.status == 1 else None, steps=r.steps,
                    footprint_bytes=r.footprint_bytes)

    def release(self, rid: int) -> None:
        """Drop a finished request's result (the server keeps each one until then)."""
        N.check(N.gpu_lib().cpim_server_release(self._h, rid))
        self._n.pop(rid, None)

    def memory(self) -> dict:
        """Device bytes charged to resident requests (now and peak), queued and resident counts."""
        cur, peak = C.c_int64(0), C.c_int64(0)
        q, r = C.c_int32(0), C.c_int32(0)
        N.check(N.gpu_lib().cpim_server_memory(self._h, C.byref(cur), C.byref(peak), C.byref(q), C.byref(r)))
        return dict(resident_bytes=cur.value, peak_bytes=peak.value, queued=q.value, resident=r.value)
