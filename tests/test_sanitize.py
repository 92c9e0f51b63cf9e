"""The library's threaded host code under the sanitizers (CPU only): make
sanitize builds taichi_amd/csrc/host_check.hip with the loader, generators,
host partition, format cache and device-layout planner, once with the
address + undefined-behaviour sanitizers and once with the thread
sanitizer, and runs both; any sanitizer report or failed check fails."""
import os
import shutil
import subprocess

import pytest

CSRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "taichi_amd", "csrc")


@pytest.mark.skipif(shutil.which("make") is None or not os.path.exists("/opt/rocm/llvm/bin/clang++"),
                    reason="needs make and ROCm's clang")
def test_host_code_under_sanitizers():
    jobs = str(min(8, os.cpu_count() or 4))
    r = subprocess.run(["make", "-s", "-C", CSRC, "-j", jobs, "sanitize"], capture_output=True, text=True,
                       timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "WARNING: ThreadSanitizer" not in out
    assert "runtime error" not in out  # UBSan
    assert out.count(": 0 failed checks") == 2, out[-4000:]
