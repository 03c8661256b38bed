"""The library's host code and the oracle's C under AddressSanitizer and UBSan, on the CPU:
`make -C pyvl_amd/csrc sanitize` builds every library source with the sanitizers on its host
side and runs tests/native/sanitize_driver.cpp (host topology builders against the oracle on
60 seeded meshes, plus the oracle's numerics); any sanitizer report fails the run."""

import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("g++") is None or not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="needs g++ and hipcc")
def test_host_code_under_sanitizers():
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(["make", "-s", "-j", jobs, "-C", str(ROOT / "pyvl_amd" / "csrc"), "sanitize"],
                       capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "no sanitizer report" in out
    assert "runtime error" not in out and "AddressSanitizer" not in out
