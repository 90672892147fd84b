"""Write the run-history fixture of tests/test_history.py and check it with
the HDF5 library.

The package's history writer (quantum-jax_amd/history.py, the layout of
save_run_to_hdf5 in quantum-fluid-accretion.py:1315-1398) writes
``history_run.h5`` from seeded inputs (``history_inputs``); the real HDF5
library, through h5py, then reads every dataset and attribute back, and what
it read is saved as ``history_h5py.npz``.  The CPU test rewrites the file
from the same inputs (it must come out byte for byte the same, so a change
of the writer needs this script run again) and reads it with the package's
own reader, comparing with what h5py read.

Run with an interpreter that has h5py (this image: /opt/conda/bin/python3.9,
h5py 3.3.0 on HDF5 1.10.6) from the repository root:

    /opt/conda/bin/python3.9 tests/golden/make_history_fixture.py
"""

import importlib.util
import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def load_writer():
    # the module alone (the package's __init__ needs torch)
    spec = importlib.util.spec_from_file_location(
        "qj_history", os.path.join(ROOT, "quantum-jax_amd", "history.py"))
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def history_inputs():
    """Seeded history, views and params of a short run (shapes as the
    reference's, sizes small)."""
    rng = np.random.default_rng(2024)
    n = 40
    H = {k: rng.random(n) for k in ("M_bh", "dMdt", "rho_inf", "r_B", "d_bh_peak", "vbh",
                                    "M_box_dm", "M_box_gas", "M_box_tot", "cap_ratio",
                                    "alpha_target")}
    H["time"] = np.linspace(13.0, 28.0, n)
    H["triax_npts"] = rng.integers(0, 5000, n)
    H["rc"] = np.linspace(0.05, 2.0, 12)
    H["prof_dm"] = [rng.random(12) for _ in range(4)]
    H["prof_gas"] = [rng.random(12) for _ in range(4)]
    H["mach_rc"] = np.linspace(0.05, 2.0, 12)
    H["mach_profile_last"] = rng.random(12)
    V = {"dm_xy": [rng.random((16, 16)) for _ in range(4)],
         "g_xy": [rng.random((16, 16)) for _ in range(4)],
         "bh_xy": [[float(i), float(2 * i)] for i in range(4)]}
    params = {"nx": 128, "Lx_kpc": 10.0, "dx_kpc": 10.0 / 128, "BH_ON": True, "BH_GRAV": True,
              "BH_FMAX": -1.0, "R_ACC_MULT_cells": 2.5, "LAMBDA_ISO": 1.1204222675845161,
              "ANN_LO": 1.1, "ANN_HI": 1.8, "cs_const": 60.0}
    return H, V, params


def main():
    w = load_writer()
    H, V, params = history_inputs()
    path = os.path.join(HERE, "history_run.h5")
    w.save_run_to_hdf5(path, H, V, 30, params, created="2026-01-01 00:00:00")
    out = {}
    with h5py.File(path, "r") as f:
        def visit(name, obj):
            if isinstance(obj, h5py.Dataset):
                assert obj.compression == "gzip" and obj.compression_opts == 4, name
                out[f"data:{name}"] = obj[...]
            for k, v in obj.attrs.items():
                out[f"attr:{name}:{k}"] = np.asarray(v)
        f.visititems(visit)
        for k, v in f.attrs.items():
            out[f"attr::{k}"] = np.asarray(v)
    np.savez(os.path.join(HERE, "history_h5py.npz"), **out)
    print(f"{path}: {len(out)} values read back by h5py {h5py.__version__} "
          f"(HDF5 {h5py.version.hdf5_version})")


if __name__ == "__main__":
    main()
