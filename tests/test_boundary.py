"""The C ABI boundary on the CPU: the library exists, loads and exports every
symbol that include/qj_api.h declares; the ctypes mirrors match the header.
No compute call is made (there is no GPU here)."""

import ctypes
import os
import re

import pytest

from conftest import ROOT

HEADER = os.path.join(ROOT, "include", "qj_api.h")
LIB = os.path.join(ROOT, "quantum-jax_amd", "libqjamd.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qj_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_operator_surface():
    names = declared_functions()
    for want in ("qj_plan_create", "qj_get_potential", "qj_kick", "qj_drift",
                 "qj_cic_indices", "qj_bin_particles", "qj_interp", "qj_get_acceleration",
                 "qj_apply_grav_accel", "qj_solve_hydro", "qj_kdk_steps", "qj_last_error"):
        assert want in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(LIB):
        pytest.fail(f"{LIB} not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_signatures_cover_the_header():
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    assert set(_lib.SIGNATURES) == set(declared_functions())


def test_state_struct_layout_matches_header():
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    text = open(HEADER).read()
    body = text[text.index("typedef struct {\n  void* psi;"):text.index("} qj_state;")]
    fields = re.findall(r"\b([a-z_]+)(?:\[\d\])?;", re.sub(r"/\*.*?\*/", "", body, flags=re.S))
    assert fields == [f[0] for f in _lib.QjState._fields_]
    # 2 + 4 + 4 + 6 pointer-sized members, two ints, two doubles, one int
    # (padded to 8)
    assert ctypes.sizeof(_lib.QjState) == 192
    # and the C compiler's layout of the header agrees, member by member
    import shutil
    import subprocess
    import tempfile
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        return
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        with open(src, "w") as fh:
            fh.write('#include "qj_api.h"\n#include <stdio.h>\n#include <stddef.h>\n'
                     "int main(void) {\n"
                     '  printf("%zu\\n", sizeof(qj_state));\n' +
                     "".join(f'  printf("%zu\\n", offsetof(qj_state, {f}));\n' for f in fields) +
                     "  return 0;\n}\n")
        exe = os.path.join(d, "layout")
        subprocess.run([cc, "-I", os.path.dirname(HEADER), src, "-o", exe], check=True)
        got = [int(x) for x in subprocess.run([exe], capture_output=True, text=True,
                                              check=True).stdout.split()]
    assert got[0] == ctypes.sizeof(_lib.QjState)
    assert got[1:] == [getattr(_lib.QjState, f).offset for f in fields]


def test_params_struct_layout_matches_header():
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    text = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    body = text[text.index("typedef struct {\n  double dt;"):text.index("} qj_params;")]
    fields = re.findall(r"\b([A-Za-z_]+);", body)
    assert fields == [f[0] for f in _lib.QjParams._fields_]
    assert ctypes.sizeof(_lib.QjParams) == 56
    consts = dict(re.findall(r"#define (QJ_FLUX_[A-Z]+) (\d+)", text))
    assert {k[8:].lower(): int(v) for k, v in consts.items()} == _lib.FLUXES


def test_no_cpu_fallback_when_library_missing(monkeypatch):
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libqjamd.so")
    with pytest.raises(_lib.QjError):
        _lib.lib()


def test_package_ic_equals_oracle_ic():
    import importlib

    import numpy as np

    import oracle as O
    ic = importlib.import_module("quantum-jax_amd.ic")
    g = O.Grid(16, 8, 32, 1.0, 2.0, 0.5)
    a = ic.filtered_gaussian_psi(g.shape, 17, (g.Lx, g.Ly, g.Lz))
    b = O.filtered_gaussian_psi(g, 17)
    assert np.allclose(a, b, rtol=0, atol=1e-14)
    pa = ic.star_ic(100, 3, (1.0, 2.0, 0.5))
    pb = O.star_ic(g, 100, 3)
    assert np.array_equal(pa[0], pb[0]) and np.array_equal(pa[1], pb[1])
    ga = ic.gas_ic(g.shape, 5, (g.Lx, g.Ly, g.Lz))
    gb = O.gas_ic(g, 5)
    assert np.allclose(ga[0], gb[0], rtol=0, atol=1e-14)


@pytest.mark.parametrize("shape", [(48, 16, 16), (16, 4, 16), (16, 16, 1024), (0, 8, 8)])
def test_unsupported_plan_shapes_are_rejected(shape):
    """Argument checks of qj_plan_create return before any device work."""
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    core = importlib.import_module("quantum-jax_amd.core")
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.qj_plan_create(ctypes.byref(h), *shape, 1.0, 1.0, 1.0, _lib.QJ_C64) == -2
    assert "powers of two" in L.qj_last_error().decode()
    assert not h.value
    with pytest.raises(ValueError, match="powers of two"):
        core.Plan(*shape)


def test_bad_plan_arguments_are_rejected():
    import importlib
    _lib = importlib.import_module("quantum-jax_amd._lib")
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.qj_plan_create(ctypes.byref(h), 16, 16, 16, 1.0, 0.0, 1.0, _lib.QJ_C64) == -1
    assert "box lengths" in L.qj_last_error().decode()
    assert L.qj_plan_create(ctypes.byref(h), 16, 16, 16, 1.0, 1.0, 1.0, 7) == -1
    assert "dtype" in L.qj_last_error().decode()
    assert L.qj_plan_create(None, 16, 16, 16, 1.0, 1.0, 1.0, _lib.QJ_C64) == -1
