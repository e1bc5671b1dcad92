"""Compiled-kernel structure (CPU only: hipcc emits gfx950 assembly here)."""
import os
import subprocess

import pytest

from kernel_structure import (barrier_loops, barrier_loops_with_exec_exit, sleeping_loops,
                              unbounded_spin_loops)

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def engine_asm(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    out = tmp_path_factory.mktemp("asm") / "engine.s"
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O3", "-std=c++17", "--cuda-device-only", "-S",
                    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                    "-I", os.path.join(REPO, "include"), "-o", str(out),
                    os.path.join(REPO, "cpim_amd", "csrc", "engine.hip")],
                   check=True, capture_output=True)
    return out.read_text()


def test_no_barrier_loop_exits_through_exec_mask(engine_asm):
    bad = {k: v for k, v in barrier_loops_with_exec_exit(engine_asm).items() if v}
    assert not bad, f"barrier loops lowered to exec-mask loops: {bad}"


def test_every_workgroup_kernel_is_checked(engine_asm):
    # each kernel with a barrier in a loop is found, and so is that loop: the
    # persistent fixpoint's round loop (grid barrier) included
    loops = barrier_loops(engine_asm)
    for k in ("k_singleton_tests", "k_sliced_tests", "k_sliced_wide", "k_mac", "k_mac_units", "k_block_gac",
              "k_gac_persistent"):
        found = [n for n in loops if k in n]
        assert found, k
        assert all(loops[n] for n in found), f"{k}: no barrier loop recognised"


def test_every_spin_wait_is_bounded(engine_asm):
    # a loop that waits for other workgroups reads the wall clock, so that its
    # wait ends with an error flag when a partner never arrives
    bad = {k: v for k, v in unbounded_spin_loops(engine_asm).items() if v}
    assert not bad, f"spin loops without a wall-clock bound: {bad}"
    spins = sleeping_loops(engine_asm)
    for k in ("k_gac_persistent", "k_gac_async"):
        assert any(v for n, v in spins.items() if k in n), f"{k}: its wait loop was not found"


def test_checker_flags_an_unbounded_spin():
    # reduced shape of a grid barrier that waits for a generation change
    # without a clock read
    asm = "\n".join([
        "_Z4kernv:",
        ".LBB0_1:                               ; =>This Inner Loop Header: Depth=1",
        "\ts_sleep 1",
        "\tglobal_load_dword v5, v3, s[4:5] offset:32 sc1",
        "\ts_cbranch_execnz .LBB0_1",
        "; %bb.2:",
        "\ts_endpgm",
    ])
    assert unbounded_spin_loops(asm) == {"_Z4kernv": ["BB0_1"]}
    bounded = asm.replace("\ts_sleep 1", "\ts_memrealtime s[8:9]\n\ts_sleep 1")
    assert unbounded_spin_loops(bounded) == {"_Z4kernv": []}


def test_checker_flags_a_split_unit_loop():
    # reduced shape of a unit loop whose thread-0 lane leaves the barrier loop
    asm = "\n".join([
        "_Z4kernv:",
        ".LBB0_1:                               ; =>This Loop Header: Depth=1",
        "\ts_barrier",
        "\ts_andn2_b64 exec, exec, s[10:11]",
        "\ts_cbranch_execz .LBB0_9",
        ".LBB0_9:",
        "\ts_endpgm",
    ])
    assert barrier_loops_with_exec_exit(asm) == {"_Z4kernv": ["BB0_1"]}
