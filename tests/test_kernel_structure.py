"""Compiled-kernel structure (CPU only: hipcc emits gfx950 assembly here)."""
import os
import subprocess

import pytest

from kernel_structure import barrier_loops_with_exec_exit

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
    names = barrier_loops_with_exec_exit(engine_asm)
    for k in ("k_singleton_tests", "k_sliced_tests", "k_mac", "k_mac_units", "k_block_gac"):
        assert any(k in n for n in names), k


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
