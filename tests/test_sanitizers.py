"""The host runtime's threaded C code under AddressSanitizer + UBSan and under
ThreadSanitizer, on the CPU: tests/sanitize/Makefile builds the tile cache,
task queue, host copy pool, grouped tables and GEMM server against a host
stand-in for the HIP layer (tests/sanitize/fake_hip.c) and runs a stress
driver (tests/sanitize/stress.c) that checks every result; a sanitizer
report fails the run."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_host_runtime_under_asan_ubsan_and_tsan(tmp_path):
    r = subprocess.run(["make", "-C", os.path.join(HERE, "sanitize"), f"OUT={tmp_path}"],
                       capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "WARNING: ThreadSanitizer" not in out
    assert "runtime error" not in out, out[-4000:]
    for part in ("cache ok", "queue ok", "host copy pool ok", "server ok"):
        assert out.count(part) == 2, (part, out[-4000:])  # ASan and TSan builds
