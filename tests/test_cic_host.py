"""Host check of the CIC wrap helpers (CPU, no GPU needed).

tests/host/cic_wrap_check.cpp compiles the device header's floored modulo
(quantum-stars.py:312 jnp.mod) and index wrap (:204-205) for the host and
compares their fast paths with the general forms bit for bit on edge values
(0, -0, L and its neighbours, 2L, -L, non-finite) and millions of random
positions.
"""

import os
import shutil
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "host", "cic_wrap_check.cpp")


def test_cic_wrap_fast_paths_equal_general_forms(tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc is not available")
    exe = str(tmp_path / "cic_wrap_check")
    inc = os.path.join(ROOT, "quantum-jax_amd", "csrc")
    r = subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "--offload-arch=gfx950",
                        "-Wno-unused-command-line-argument", "-I", inc, SRC, "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    assert " 0 mismatches" in r.stdout
