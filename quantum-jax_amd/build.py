"""Build recipe for libqjamd.so (gfx950 HIP kernels + the C ABI).

Compiles every ``csrc/*.hip`` translation unit with hipcc for gfx950 in
parallel and links them into ``libqjamd.so`` next to this file (in-tree, so
the library travels with the repository snapshot).  Usable as a module
(``build()``) or a script (``python quantum-jax_amd/build.py``).
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libqjamd.so")
ARCH = "gfx950"

COMMON = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    # packed FP32 forms are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
# per-file extra flags: the CIC index/weight sequence must round like
# separate IEEE operations (no FMA contraction)
EXTRA = {
    "qj_particles.hip": ["-ffp-contract=off"],
}
# translation units compiled several times with different defines (object
# suffix, flags): the spectral kernels build as four parallel parts
PARTS = {
    "qj_spectral.hip": [(f"_{k}", [f"-DQJ_PART={k}"]) for k in range(4)],
}


def _hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found: the gfx950 library cannot be built")


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def build(verbose: bool = False, jobs: int | None = None) -> str:
    hipcc = _hipcc()
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()

    units = [(f, suf, defs) for f in srcs for suf, defs in PARTS.get(f, [("", [])])]

    def compile_one(unit):
        f, suf, defs = unit
        out = os.path.join(OBJ, f[:-4] + suf + ".o")
        cmd = [hipcc, *COMMON, *EXTRA.get(f, []), *defs, "-I", CSRC,
               "-I", os.path.join(HERE, "..", "include"),
               "-c", os.path.join(CSRC, f), "-o", out]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {f}:\n{r.stderr[-6000:]}")
        if verbose and r.stderr.strip():
            print(r.stderr, file=sys.stderr)
        return out

    jobs = jobs or min(len(units), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(compile_one, units))
    tmp = LIB + ".tmp"
    cmd = [hipcc, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr[-6000:]}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
