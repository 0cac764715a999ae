"""The library's host-only code (ingest.cpp, partition.cpp) built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on the CPU: RGG
generation at one and several ranks, vite files of both layouts (including a
truncated file), edge-list text, vertex partitions and the batch mix's host
generator (tests/integration/sanitize_host.cpp).  A sanitizer report fails
the test; no device code is built."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "xxtric_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_host_code_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "sanitize_host")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
           "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "integration", "sanitize_host.cpp"),
           os.path.join(CSRC, "ingest.cpp"), os.path.join(CSRC, "partition.cpp"), "-o", exe, "-lpthread"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "sanitized host checks passed" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr
