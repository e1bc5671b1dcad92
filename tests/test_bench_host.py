"""Host logic of bench.py (CPU only): the profile stamp that decides whether
the committed counter profile may be reported as this build's traffic, and
the single-GPU contract of the command line."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)
import bench  # noqa: E402


def test_fingerprint_covers_the_device_sources_only():
    fp = bench.kernel_source_fingerprint()
    assert len(fp) == 64 and int(fp, 16) >= 0
    assert bench.KERNEL_SOURCES == ("device.h", "kernels.h")
    assert bench.kernel_source_fingerprint() == fp  # stable


def test_traffic_is_reported_only_from_a_profile_of_these_sources(tmp_path, monkeypatch):
    good = tmp_path / "good.txt"
    good.write_text("x\ntraffic_bytes_per_launch: 1.5e+10\n"
                    f"kernel_sources_sha256: {bench.kernel_source_fingerprint()}\n")
    stale = tmp_path / "stale.txt"
    stale.write_text("traffic_bytes_per_launch: 1.5e+10\nkernel_sources_sha256: " + "0" * 64 + "\n")
    unstamped = tmp_path / "unstamped.txt"
    unstamped.write_text("traffic_bytes_per_launch: 1.5e+10\n")
    monkeypatch.setattr(bench, "PROFILES", {"a": str(good), "b": str(stale), "c": str(unstamped),
                                            "d": str(tmp_path / "missing.txt")})
    t, note = bench.profiled_traffic("a")
    assert t == 1.5e10 and "kernel sources" in note
    for mode in ("b", "c"):
        t, note = bench.profiled_traffic(mode)
        assert t is None and "other kernel sources" in note
    t, note = bench.profiled_traffic("d")
    assert t is None and "missing" in note
    t, note = bench.profiled_traffic("workgroup")
    assert t is None


def test_committed_profiles_are_stamped():
    # every counter summary names the device sources it was taken with
    prof = os.path.join(REPO, "profiles")
    for name in ("k_sliced_wide_batched_sac.txt", "k_gac_cascade_c2.txt", "k_mac_units_uneven_c4.txt"):
        text = open(os.path.join(prof, name)).read()
        stamps = [l for l in text.splitlines() if l.startswith("kernel_sources_sha256:")]
        assert len(stamps) == 1, name
        assert len(stamps[0].split(":", 1)[1].strip()) == 64, name
        assert "rocprofv3" in text, name


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
