"""Host checks of device helper functions (CPU, no GPU needed).

The headers' helpers are compiled for the host and compared with general
or higher-precision forms:
  * tests/host/cic_wrap_check.cpp: the floored modulo of the star wrap
    (quantum-stars.py:312 jnp.mod) and the CIC index wrap (:204-205), fast
    paths against the general forms bit for bit, on edge values (0, -0, L
    and its neighbours, 2L, -L, non-finite) and millions of random positions;
  * tests/host/sincos_check.cpp: the small-angle sine and cosine of the kick
    phase factor (quantum-stars.py:301 exp(-i dt/2 m/hbar V), after FP64
    range reduction) against long-double sinl/cosl, within 1.1 ulp.
"""

import os
import shutil
import subprocess

import pytest

from conftest import ROOT


def _run(name, tmp_path):
    hipcc = "/opt/rocm/bin/hipcc" if os.path.exists("/opt/rocm/bin/hipcc") else shutil.which("hipcc")
    if not hipcc:
        pytest.skip("hipcc is not available")
    exe = str(tmp_path / name)
    inc = os.path.join(ROOT, "quantum-jax_amd", "csrc")
    src = os.path.join(ROOT, "tests", "host", name + ".cpp")
    r = subprocess.run([hipcc, "-std=c++17", "-O1", "-ffp-contract=off", "--offload-arch=gfx950",
                        "-Wno-unused-command-line-argument", "-I", inc, src, "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-4000:]
    return r.stdout


def test_cic_wrap_fast_paths_equal_general_forms(tmp_path):
    assert " 0 mismatches" in _run("cic_wrap_check", tmp_path)


def test_small_angle_sincos_within_an_ulp(tmp_path):
    assert "max" in _run("sincos_check", tmp_path)
