"""ctypes view of the C-ABI in include/cpim_gpu.h.

The structures below mirror the header field for field.  The HIP library
(libcpim_gpu.so, built from cpim_amd/csrc/engine.hip) and the workload
library (libcpim_workload.so) are loaded lazily, from this directory only:
there is no fallback, so a missing build fails loudly.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
GPU_LIB_PATH = os.path.join(_HERE, "libcpim_gpu.so")
WORKLOAD_LIB_PATH = os.path.join(_HERE, "libcpim_workload.so")

CPIM_OK = 0
CPIM_ERR_INVALID_ARGUMENT = 1
CPIM_ERR_RUNTIME = 2
CPIM_ERR_HIP = 3
CPIM_ERR_OUT_OF_MEMORY = 4
CPIM_ERR_UNSUPPORTED = 5
CPIM_ERR_BUSY = 6

CPIM_GAC_PLAIN = 0
CPIM_GAC_PERSISTENT = 1
CPIM_GAC_BLOCK = 2
CPIM_GAC_AUTO = 3
CPIM_GAC_ASYNC = 4

CPIM_SAC_BATCHED = 0
CPIM_SAC_PLAIN = 1
CPIM_SAC_WORKGROUP = 2
CPIM_SAC_SLICED = 3

CPIM_BRANCH_BINARY = 0
CPIM_VARH_DOM_MIN = 0
CPIM_VARH_DOM_DEG = 1
CPIM_VARH_DOM_WDEG = 2

p_i32 = C.POINTER(C.c_int32)
p_i64 = C.POINTER(C.c_int64)
p_u64 = C.POINTER(C.c_uint64)
p_u8 = C.POINTER(C.c_uint8)


class CModelDesc(C.Structure):
    _fields_ = [
        ("num_vars", C.c_int32),
        ("dom_size", p_i32),
        ("num_cons", C.c_int32),
        ("con_arity", p_i32),
        ("con_scope", p_i32),
        ("bin_rel", p_u64),
        ("rel_words", C.c_int32),
        ("tup_off", p_i64),
        ("tuples", p_i32),
        ("init_dom", p_u64),
    ]


class CGModelOptions(C.Structure):
    _fields_ = [
        ("device_id", C.c_int32),
        ("enable_prefetch", C.c_int32),
        ("skip_non_binary", C.c_int32),
        ("skip_conflicts", C.c_int32),
    ]


class CDeviceInfo(C.Structure):
    _fields_ = [
        ("compute_units", C.c_int32),
        ("lds_bytes_per_block", C.c_int32),
        ("cooperative_launch", C.c_int32),
        ("wavefront_size", C.c_int32),
        ("global_mem_bytes", C.c_int64),
        ("arch", C.c_char * 64),
    ]


class CGModelInfo(C.Structure):
    _fields_ = [
        ("num_vars", C.c_int32),
        ("num_constraints", C.c_int32),
        ("num_active", C.c_int32),
        ("max_dom_size", C.c_int32),
        ("dom_words", C.c_int32),
        ("bit_dom_int_size", C.c_int32),
        ("max_depth", C.c_int32),
        ("subscription_size", C.c_int32),
        ("device_bytes", C.c_int64),
    ]


class CGacStats(C.Structure):
    _fields_ = [
        ("iterations", C.c_int32),
        ("deletions", C.c_int32),
        ("inconsistent", C.c_int32),
        ("converged", C.c_int32),
        ("revisions", C.c_int64),
        ("kernel_ms", C.c_double),
    ]


class CSacStats(C.Structure):
    _fields_ = [
        ("consistent", C.c_int32),
        ("passes", C.c_int32),
        ("singleton_tests", C.c_int64),
        ("removed", C.c_int64),
        ("revisions", C.c_int64),
        ("kernel_ms", C.c_double),
    ]


class CSolveOpts(C.Structure):
    _fields_ = [
        ("time_limit_ms", C.c_int64),
        ("node_limit", C.c_int64),
        ("branching", C.c_int32),
        ("trace_capacity", C.c_int32),
        ("var_heuristic", C.c_int32),
        ("reserved", C.c_int32),
        ("max_solutions", C.c_int64),
        ("solution_capacity", C.c_int64),
    ]


class CSearchStats(C.Structure):
    _fields_ = [
        ("num_solutions", C.c_int64),
        ("num_positive", C.c_int64),
        ("num_negative", C.c_int64),
        ("gac_iterations", C.c_int64),
        ("gac_deletions", C.c_int64),
        ("time_out", C.c_int32),
        ("unsolvable", C.c_int32),
        ("solve_time", C.c_double),
        ("gac_time", C.c_double),
    ]


class CNodeRecord(C.Structure):
    _fields_ = [
        ("positive", C.c_int32),
        ("var", C.c_int32),
        ("value", C.c_int32),
        ("consistent", C.c_int32),
        ("dom_hash", C.c_uint64),
    ]


class CBatchResult(C.Structure):
    _fields_ = [
        ("root_consistent", C.c_int32),
        ("status", C.c_int32),
        ("num_positive", C.c_int64),
        ("num_negative", C.c_int64),
        ("trace_len", C.c_int32),
        ("pad", C.c_int32),
    ]


class CServerOpts(C.Structure):
    _fields_ = [
        ("device_id", C.c_int32),
        ("max_slots", C.c_int32),
        ("hbm_budget_bytes", C.c_int64),
        ("nodes_per_step", C.c_int64),
    ]


class CRequestResult(C.Structure):
    _fields_ = [
        ("done", C.c_int32),
        ("status", C.c_int32),
        ("num_positive", C.c_int64),
        ("num_negative", C.c_int64),
        ("footprint_bytes", C.c_int64),
        ("steps", C.c_int32),
        ("pad", C.c_int32),
    ]


# Every function the header declares: name -> (restype, argtypes).
_VP = C.c_void_p
GPU_SIGNATURES = {
    "cpim_gmodel_options_default": (None, [C.POINTER(CGModelOptions)]),
    "cpim_last_error": (C.c_char_p, []),
    "cpim_version": (C.c_char_p, []),
    "cpim_get_device_info": (C.c_int, [C.c_int32, C.POINTER(CDeviceInfo)]),
    "cpim_measure_copy_bandwidth": (C.c_int, [C.c_int32, C.c_int64, C.c_int32, C.POINTER(C.c_double)]),
    "cpim_gmodel_create": (C.c_int, [C.POINTER(CModelDesc), C.POINTER(CGModelOptions), C.POINTER(_VP)]),
    "cpim_gmodel_destroy": (C.c_int, [_VP]),
    "cpim_gmodel_get_info": (C.c_int, [_VP, C.POINTER(CGModelInfo)]),
    "cpim_gmodel_enforce_gac": (C.c_int, [_VP, C.c_int32, C.c_int32, C.POINTER(CGacStats)]),
    "cpim_gmodel_enforce_gac_noalloc": (C.c_int, [_VP, C.c_int32, C.POINTER(CGacStats)]),
    "cpim_gmodel_create_new_level": (C.c_int, [_VP, p_i32]),
    "cpim_gmodel_back_to_level": (C.c_int, [_VP, C.c_int32]),
    "cpim_gmodel_current_level": (C.c_int, [_VP, p_i32]),
    "cpim_gmodel_assign_value": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_remove_value": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_domain_size": (C.c_int, [_VP, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_min_domain_var": (C.c_int, [_VP, C.c_int32, p_i32]),
    "cpim_gmodel_is_assigned": (C.c_int, [_VP, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_is_fully_assigned": (C.c_int, [_VP, C.c_int32, p_i32]),
    "cpim_gmodel_assigned_value": (C.c_int, [_VP, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_first_value": (C.c_int, [_VP, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_next_value": (C.c_int, [_VP, C.c_int32, C.c_int32, C.c_int32, p_i32]),
    "cpim_gmodel_read_domains": (C.c_int, [_VP, C.c_int32, p_u64]),
    "cpim_gmodel_write_domains": (C.c_int, [_VP, C.c_int32, p_u64]),
    "cpim_gmodel_enforce_sac": (C.c_int, [_VP, C.c_int32, C.POINTER(CSacStats)]),
    "cpim_gmodel_singleton_tests": (C.c_int, [_VP, p_i32, C.c_int64, C.c_int32, p_u8, C.POINTER(CSacStats)]),
    "cpim_gmodel_weights": (C.c_int, [_VP, C.POINTER(C.c_float), C.c_int64, p_i64, C.c_int32]),
    "cpim_gmodel_solve": (C.c_int, [_VP, C.POINTER(CSolveOpts), C.POINTER(CSearchStats), p_i32,
                                    C.POINTER(CNodeRecord), p_i32]),
    "cpim_batch_create": (C.c_int, [C.POINTER(CModelDesc), C.c_int32, C.POINTER(CGModelOptions), C.POINTER(_VP)]),
    "cpim_batch_destroy": (C.c_int, [_VP]),
    "cpim_batch_run": (C.c_int, [_VP, C.c_int64, C.c_int32, C.POINTER(CBatchResult), C.POINTER(CNodeRecord),
                                 C.c_int32, C.POINTER(C.c_double)]),
    "cpim_batch_read_root": (C.c_int, [_VP, C.c_int32, p_u64]),
    "cpim_batch_heavy_units": (C.c_int, [_VP, p_i32]),
    "cpim_server_create": (C.c_int, [C.POINTER(CServerOpts), C.POINTER(_VP)]),
    "cpim_server_destroy": (C.c_int, [_VP]),
    "cpim_server_submit": (C.c_int, [_VP, C.POINTER(CModelDesc), C.c_int64, p_i64]),
    "cpim_server_step": (C.c_int, [_VP, p_i32]),
    "cpim_tune_set": (C.c_int, [C.c_char_p, C.c_int64]),
    "cpim_tune_get": (C.c_int, [C.c_char_p, C.POINTER(C.c_int64)]),
    "cpim_server_poll": (C.c_int, [_VP, C.POINTER(C.c_int64), C.c_int32, C.POINTER(C.c_int32)]),
    "cpim_server_result": (C.c_int, [_VP, C.c_int64, C.POINTER(CRequestResult), p_i32]),
    "cpim_server_release": (C.c_int, [_VP, C.c_int64]),
    "cpim_server_memory": (C.c_int, [_VP, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32),
                                     C.POINTER(C.c_int32)]),
}

_gpu = None
_workload = None


def gpu_lib():
    """The HIP engine library.  Raises if it has not been built."""
    global _gpu
    if _gpu is None:
        if not os.path.exists(GPU_LIB_PATH):
            raise RuntimeError(
                f"{GPU_LIB_PATH} is missing: build it with cpim_amd.build.build_all() "
                "(there is no CPU fallback)")
        lib = C.CDLL(GPU_LIB_PATH)
        for name, (res, args) in GPU_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _gpu = lib
    return _gpu


def workload_lib():
    global _workload
    if _workload is None:
        if not os.path.exists(WORKLOAD_LIB_PATH):
            raise RuntimeError(f"{WORKLOAD_LIB_PATH} is missing: run cpim_amd.build.build_all()")
        lib = C.CDLL(WORKLOAD_LIB_PATH)
        lib.cpimw_splitmix64.restype = C.c_uint64
        lib.cpimw_splitmix64.argtypes = [C.c_uint64]
        lib.cpimw_instance_seed.restype = C.c_uint64
        lib.cpimw_instance_seed.argtypes = [C.c_uint64, C.c_uint64, C.c_uint64]
        lib.cpimw_random_pairs.restype = C.c_int
        lib.cpimw_random_pairs.argtypes = [C.c_uint64, C.c_int, C.c_int, p_i32]
        lib.cpimw_random_relations.restype = C.c_int
        lib.cpimw_random_relations.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, C.c_int,
                                               C.c_double, p_u64]
        lib.cpimw_domains_with_removals.restype = C.c_int
        lib.cpimw_domains_with_removals.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int,
                                                    C.c_double, p_u64]
        lib.cpimw_random_table.restype = C.c_int
        lib.cpimw_random_table.argtypes = [C.c_uint64, C.c_int, C.c_int, C.c_int, p_i32]
        lib.cpimw_uniform_int.restype = C.c_int64
        lib.cpimw_uniform_int.argtypes = [C.c_uint64, C.c_uint64, C.c_int64, C.c_int64]
        lib.cpimw_uniform_real.restype = C.c_double
        lib.cpimw_uniform_real.argtypes = [C.c_uint64, C.c_uint64, C.c_double, C.c_double]
        _workload = lib
    return _workload


class CpimError(RuntimeError):
    """A failed C-ABI call; code is the CPIM_ERR_* value."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class CpimInvalidArgument(CpimError, ValueError):
    """CPIM_ERR_INVALID_ARGUMENT (std::invalid_argument in the C++ facade)."""


def check(code: int) -> None:
    if code != CPIM_OK:
        msg = gpu_lib().cpim_last_error()
        text = msg.decode() if msg else f"cpim error {code}"
        if code == CPIM_ERR_INVALID_ARGUMENT:
            raise CpimInvalidArgument(code, text)
        raise CpimError(code, text)


def tune_set(name: str, value: int) -> None:
    """Set an engine tuning knob (0 = automatic); see include/cpim_gpu.h."""
    check(gpu_lib().cpim_tune_set(name.encode(), int(value)))


def tune_get(name: str) -> int:
    v = C.c_int64(0)
    check(gpu_lib().cpim_tune_get(name.encode(), C.byref(v)))
    return v.value
