"""HDF5 run-history writer (save_run_to_hdf5 of quantum-fluid-accretion.py:1315-1398), CPU.

tests/golden/history_run.h5 was written by the package's writer and read
back by the HDF5 library itself (h5py 3.3.0, HDF5 1.10.6), which saved
what it read in history_h5py.npz (tests/golden/make_history_fixture.py).
Here the writer must reproduce the file byte for byte from the same inputs,
and the package's reader must read the same values as the HDF5 library.
"""

import importlib
import os

import numpy as np
import pytest

from conftest import ROOT

GOLDEN = os.path.join(ROOT, "tests", "golden")
hist = importlib.import_module("quantum-jax_amd.history")
h5 = importlib.import_module("quantum-jax_amd.hdf5")


def _inputs():
    # the fixture script imports h5py at the top; take its input builder only
    src = open(os.path.join(GOLDEN, "make_history_fixture.py")).read()
    ns = {"np": np}
    start = src.index("def history_inputs")
    end = src.index("def main")
    exec(src[start:end], ns)
    return ns["history_inputs"]()


def test_writer_reproduces_the_fixture(tmp_path):
    H, V, params = _inputs()
    out = tmp_path / "run.h5"
    hist.save_run_to_hdf5(str(out), H, V, 30, params, created="2026-01-01 00:00:00")
    assert out.read_bytes() == open(os.path.join(GOLDEN, "history_run.h5"), "rb").read()


def test_reader_agrees_with_the_hdf5_library():
    ref = np.load(os.path.join(GOLDEN, "history_h5py.npz"), allow_pickle=False)
    with h5.File(os.path.join(GOLDEN, "history_run.h5")) as f:
        names = {n for n, kind, _ in f.visit() if kind == "dataset"}
        assert names == {k[5:] for k in ref.files if k.startswith("data:")}
        for n in names:
            a = f.read(n)
            b = ref[f"data:{n}"]
            assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a, b), n
        for k in ref.files:
            if not k.startswith("attr:"):
                continue
            _, obj, name = k.split(":", 2)
            got = f.attrs[name] if obj == "" else f.attrs_of(obj)[name]
            assert np.array_equal(np.asarray(got), ref[k]), k


def test_layout_follows_the_reference():
    H, V, params = _inputs()
    with h5.File(os.path.join(GOLDEN, "history_run.h5")) as f:
        groups = {n for n, kind, _ in f.visit() if kind == "group"}
        assert groups == {"params", "history", "profiles", "mach", "views_xy"}
        assert f.read("profiles/prof_dm").dtype == np.float32
        assert f.read("views_xy/dm").shape == (4, 16, 16)
        assert np.array_equal(f.read("history/time"), H["time"])
        assert f.attrs["fps"] == 30


def test_writer_rejects_what_it_cannot_encode(tmp_path):
    w = hist.Writer()
    with pytest.raises(ValueError):
        w.create_dataset("x", np.array(["a", "b"]))
    w2 = hist.Writer()
    w2.create_dataset("empty", np.zeros((0,)))
    w2.create_dataset("scalar", np.float64(2.5))
    w2.save(str(tmp_path / "e.h5"))
    with h5.File(str(tmp_path / "e.h5")) as f:
        assert f.read("empty").shape == (0,)
        assert f.read("scalar") == 2.5
