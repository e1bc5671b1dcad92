"""C-ABI boundary and host logic (CPU only: no compute call reaches a GPU)."""
import ctypes as C
import hashlib
import os
import re

import numpy as np
import pytest

from cpim_amd import _native as N
from cpim_amd import workload
from cpim_amd.model import ModelDesc, from_tuples

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(REPO, "include", "cpim_gpu.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(cpim_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = C.CDLL(N.GPU_LIB_PATH)
    names = declared_functions()
    assert len(names) >= 30
    for name in names:
        assert hasattr(lib, name), name


def test_ctypes_signatures_cover_the_header():
    assert sorted(N.GPU_SIGNATURES) == declared_functions()


def test_struct_layouts_match_header_sizes():
    # 64-bit ABI sizes of the plain structs (alignment as in the header)
    assert C.sizeof(N.CModelDesc) == 80
    assert C.sizeof(N.CGacStats) == 32
    assert C.sizeof(N.CNodeRecord) == 24
    assert C.sizeof(N.CSearchStats) == 64
    assert C.sizeof(N.CBatchResult) == 32
    assert C.sizeof(N.CSolveOpts) == 48


def test_version_and_default_options():
    lib = N.gpu_lib()
    assert b"gfx950" in lib.cpim_version()
    o = N.CGModelOptions()
    lib.cpim_gmodel_options_default(C.byref(o))
    assert (o.device_id, o.enable_prefetch, o.skip_non_binary, o.skip_conflicts) == (0, 0, 1, 0)


def create_status(desc):
    lib = N.gpu_lib()
    h = C.c_void_p()
    cd = desc.c_desc()
    code = lib.cpim_gmodel_create(C.byref(cd), None, C.byref(h))
    return code, lib.cpim_last_error().decode()


@pytest.mark.parametrize("case", ["no_vars", "empty_domain", "bad_scope", "repeat_var",
                                  "short_rows", "bad_tuple", "big_domain"])
def test_invalid_models_rejected_before_device(case):
    good = workload.random_binary(6, 5, 8, 0.3, 1)
    if case == "no_vars":
        d = ModelDesc(np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros(0, np.int32))
    elif case == "empty_domain":
        ds = good.dom_size.copy()
        ds[2] = 0
        d = ModelDesc(ds, good.con_arity, good.con_scope, good.bin_rel)
    elif case == "bad_scope":
        sc = good.con_scope.copy()
        sc[3] = 99
        d = ModelDesc(good.dom_size, good.con_arity, sc, good.bin_rel)
    elif case == "repeat_var":
        sc = good.con_scope.copy()
        sc[1] = sc[0]
        d = ModelDesc(good.dom_size, good.con_arity, sc, good.bin_rel)
    elif case == "short_rows":
        d = workload.random_binary(6, 100, 8, 0.3, 1)
        d = ModelDesc(d.dom_size, d.con_arity, d.con_scope, d.bin_rel[:, :, :1].copy())
    elif case == "bad_tuple":
        d = from_tuples([3, 3], [(0, 1)], [np.array([[0, 5]], np.int32)])
    else:
        d = from_tuples([2000, 3], [(0, 1)], [np.array([[0, 1]], np.int32)])
    code, msg = create_status(d)
    assert code == N.CPIM_ERR_INVALID_ARGUMENT, (code, msg)
    assert msg


def test_null_arguments_are_errors_not_crashes():
    lib = N.gpu_lib()
    assert lib.cpim_gmodel_destroy(None) == N.CPIM_OK
    assert lib.cpim_gmodel_create(None, None, None) == N.CPIM_ERR_INVALID_ARGUMENT
    assert lib.cpim_gmodel_get_info(None, None) == N.CPIM_ERR_INVALID_ARGUMENT
    st = N.CGacStats()
    assert lib.cpim_gmodel_enforce_gac(None, -1, 0, C.byref(st)) == N.CPIM_ERR_INVALID_ARGUMENT
    assert lib.cpim_batch_destroy(None) == N.CPIM_OK
    assert lib.cpim_server_destroy(None) == N.CPIM_OK


# ---------------------------------------------------------------------------
# Workload generator
# ---------------------------------------------------------------------------
def test_generator_is_deterministic_and_distinct():
    a = workload.random_binary(40, 16, 200, 0.4, 12345)
    b = workload.random_binary(40, 16, 200, 0.4, 12345)
    c = workload.random_binary(40, 16, 200, 0.4, 12346)
    assert (a.bin_rel == b.bin_rel).all() and (a.con_scope == b.con_scope).all()
    assert not (a.bin_rel == c.bin_rel).all()
    pairs = a.con_scope.reshape(-1, 2)
    assert (pairs[:, 0] < pairs[:, 1]).all()
    assert len({tuple(p) for p in pairs.tolist()}) == len(pairs)
    # lexicographic order of the chosen pairs
    assert pairs.tolist() == sorted(pairs.tolist())


def test_generator_tightness_and_masks():
    d = workload.random_binary(20, 100, 60, 0.3, 7)
    bits = np.unpackbits(d.bin_rel.view(np.uint8), bitorder="little").reshape(60, 100, -1)
    assert bits[:, :, 100:].sum() == 0  # nothing beyond the domain
    allowed = bits[:, :, :100].mean()
    assert abs(allowed - 0.7) < 0.01


def test_generator_known_checksum():
    # pins the generator across compilers and hosts
    d = workload.random_binary(30, 16, 120, 0.40, workload.instance_seed(1, 0))
    h = hashlib.sha256(d.bin_rel.tobytes() + d.con_scope.tobytes()).hexdigest()[:16]
    assert h == "59fa02a3fdf52fa8"


def test_phase_transition_formula():
    # d^n (1 - p*)^e = 1 with e = rho n (n-1) / 2
    n, d, rho = 50, 10, 0.3
    p = workload.phase_transition_tightness(n, d, rho)
    e = rho * n * (n - 1) / 2
    assert abs(n * np.log(d) + e * np.log(1 - p)) < 1e-9


def test_random_table_distinct_sorted():
    t = workload.random_table(9, 3, 16, 500)
    assert t.shape == (500, 3) and t.min() >= 0 and t.max() < 16
    rows = [tuple(r) for r in t.tolist()]
    assert rows == sorted(set(rows))


def test_serving_request_shape():
    m = workload.serving_request(3)
    assert (m.con_arity == 3).sum() == 4
    lists = m.tuple_lists()
    assert all(len(lists[c]) == 500 for c in range(m.num_cons) if m.con_arity[c] == 3)


def test_tuning_knobs():
    from cpim_amd import _native as N
    for name in ("sac_slice_masks", "sac_slice_threads", "sac_slice_kernel", "sac_slice_nibbles",
                 "sac_slice_bytes", "unit_threads", "mac_lds_net"):
        old = N.tune_get(name)
        N.tune_set(name, 0)
        assert N.tune_get(name) == 0
        N.tune_set(name, old)
    N.tune_set("sac_slice_masks", 4)
    assert N.tune_get("sac_slice_masks") == 4
    N.tune_set("sac_slice_masks", 0)
    with pytest.raises(N.CpimInvalidArgument, match="out of range"):
        N.tune_set("sac_slice_masks", 3)
    with pytest.raises(N.CpimInvalidArgument, match="out of range"):
        N.tune_set("unit_threads", 100)
    with pytest.raises(N.CpimInvalidArgument, match="out of range"):
        N.tune_set("mac_lds_net", 2)
    with pytest.raises(N.CpimInvalidArgument, match="unknown tuning knob"):
        N.tune_set("no_such_knob", 1)


CLI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "cpim_amd", "cpim_solve")


def test_cpp_cli_reports_bad_models_as_invalid_argument():
    # the C++ facade maps a bad model to std::invalid_argument before any
    # device work
    import subprocess
    if not os.path.exists(CLI):
        pytest.skip("cpim_solve not built")
    r = subprocess.run([CLI, "--validate-only"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "invalid_argument: scope names an unknown variable" in r.stdout
    r = subprocess.run([CLI, "--varh", "wdeg"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "--varh" in r.stderr
