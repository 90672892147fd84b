"""Host-side AddressSanitizer run of the C ABI (CPU, no GPU needed).

tests/asan/build_and_run.sh compiles the host-heavy translation units with
the sanitizer on their host pass, links them with the ordinary kernel
objects and runs tests/asan/abi_driver.cpp, which drives every entry
point's argument-checking and error paths (including plan creation failing
without a device and its clean-up).  Any heap misuse, or a wrong status,
fails the run.
"""

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SCRIPT = os.path.join(ROOT, "tests", "asan", "build_and_run.sh")
OBJ = os.path.join(ROOT, "quantum-jax_amd", "_build")


def test_c_abi_host_code_under_address_sanitizer():
    if not (os.path.exists("/opt/rocm/bin/hipcc") or shutil.which("hipcc")):
        pytest.skip("hipcc is not available")
    if not os.path.isdir(OBJ) or not any(f.endswith(".o") for f in os.listdir(OBJ)):
        pytest.skip("the ordinary build's objects are missing (run __graft_entry__.build())")
    r = subprocess.run(["sh", SCRIPT], capture_output=True, text=True, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "ERROR: LeakSanitizer" not in out
    assert "ok: 0 failures" in out
