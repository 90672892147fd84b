"""The HDF5 initial-condition reader against files written by the HDF5
library itself (tests/golden/make_hdf5_fixtures.py: h5py 3.3.0 on HDF5
1.10.6), in the storage variants of superblocks 0 and 3, version-1 and
version-2 object headers with continuation blocks, symbol-table and
link-message groups, compact, contiguous and chunked layouts with shuffle +
deflate, big-endian data and a user block.  The reference reads
f["psiRe"] + 1j f["psiIm"] and scales by 1e5 (quantum-cosmo.py:242-244);
its own data file is a Git LFS pointer here, so these fixtures stand in."""

import importlib
import os

import numpy as np
import pytest

h5 = importlib.import_module("quantum-jax_amd.hdf5")
qj = importlib.import_module("quantum-jax_amd")

GOLD = os.path.join(os.path.dirname(__file__), "golden")
EXPECTED = np.load(os.path.join(GOLD, "hdf5_expected.npz"))
FILES = ["ic_default.h5", "ic_chunked_gzip.h5", "ic_latest.h5", "ic_userblock.h5"]


def cases():
    for key in EXPECTED.files:
        fname, name = key.split(":", 1)
        yield pytest.param(fname, name, id=key)


@pytest.mark.parametrize("fname,name", list(cases()))
def test_datasets_equal_what_the_library_wrote(fname, name):
    want = EXPECTED[f"{fname}:{name}"]
    got = h5.read(os.path.join(GOLD, fname), name)
    assert got.shape == want.shape
    assert got.dtype == want.dtype.newbyteorder("=")
    assert np.array_equal(got, want)


def test_storage_variants_are_the_ones_named():
    with h5.File(os.path.join(GOLD, "ic_default.h5")) as f:
        assert f.superblock_version == 0
        f.visit()
        assert f.continuation_blocks > 0  # version-1 header continuation
    with h5.File(os.path.join(GOLD, "ic_latest.h5")) as f:
        assert f.superblock_version == 3
        f.visit()
        assert f.continuation_blocks > 0  # OCHK blocks of version-2 headers
    with h5.File(os.path.join(GOLD, "ic_userblock.h5")) as f:
        assert f.superblock_version == 0 and f._base == 512


@pytest.mark.parametrize("fname", FILES)
def test_load_psi_follows_the_reference(fname):
    path = os.path.join(GOLD, fname)
    re = EXPECTED[f"{fname}:psiRe"].astype(np.float64)
    im = EXPECTED[f"{fname}:psiIm"].astype(np.float64)
    psi = h5.load_psi(path)
    assert psi.dtype == np.complex128
    assert np.array_equal(psi, (re + 1j * im) * 1e5)
    # the reference forms psi in complex64: agreement to its rounding
    ref64 = (EXPECTED[f"{fname}:psiRe"].astype(np.complex64)
             + np.complex64(1j) * EXPECTED[f"{fname}:psiIm"].astype(np.complex64)) * np.float32(1e5)
    assert np.max(np.abs(psi - ref64)) <= 2e-7 * np.max(np.abs(psi))
    assert qj.ic.hdf5_psi is h5.load_psi


def test_listing_and_attributes():
    with h5.File(os.path.join(GOLD, "ic_default.h5")) as f:
        assert sorted(f.keys()) == ["meta", "psiIm", "psiRe"]
        listing = {p: (k, a) for p, k, a in f.visit()}
        assert [p for p, _, _ in f.visit()] == ["meta", "meta/nx", "psiIm", "psiRe"]
        assert listing["meta"][0] == "group" and listing["psiRe"][0] == "dataset"
        assert listing["meta"][1]["code"] == b"fdm-ic"
        notes = listing["psiRe"][1]
        assert sorted(notes) == [f"note{k:02d}" for k in range(12)]
        assert [float(notes[f"note{k:02d}"]) for k in range(12)] == [k / 4 for k in range(12)]
        assert float(f.attrs["redshift"]) == 127.0
        assert np.array_equal(f.attrs["box_kpc"], [1000.0, 1000.0, 1000.0])
        assert f.attrs_of("meta")["code"] == b"fdm-ic"
    with h5.File(os.path.join(GOLD, "ic_latest.h5")) as f:
        notes = f.attrs_of("psiIm")
        assert {k: int(v) for k, v in notes.items()} == {f"note{k}": -k for k in range(6)}


def test_errors_are_clear(tmp_path):
    lfs = tmp_path / "ic.hdf5"
    lfs.write_text("version https://git-lfs.github.com/spec/v1\noid sha256:" + "0" * 64
                   + "\nsize 268435456\n")
    with pytest.raises(h5.HDF5Error, match="LFS pointer"):
        h5.load_psi(lfs)
    empty = tmp_path / "empty.h5"
    empty.write_bytes(b"")
    with pytest.raises(h5.HDF5Error, match="empty"):
        h5.File(empty)
    junk = tmp_path / "junk.h5"
    junk.write_bytes(bytes(range(256)) * 8)
    with pytest.raises(h5.HDF5Error, match="no HDF5 signature"):
        h5.File(junk)
    src = open(os.path.join(GOLD, "ic_default.h5"), "rb").read()
    cut = tmp_path / "cut.h5"
    cut.write_bytes(src[:600])  # the superblock survives, the objects do not
    with pytest.raises(h5.HDF5Error):
        h5.read(cut, "psiRe")
    path = os.path.join(GOLD, "ic_default.h5")
    with pytest.raises(KeyError, match="psiX"):
        h5.read(path, "psiX")
    with pytest.raises(KeyError, match="meta"):
        h5.read(path, "meta/missing")
    with pytest.raises(h5.HDF5Error, match="group"):
        h5.read(path, "meta")
    with pytest.raises(h5.HDF5Error, match="differ in shape"):
        h5.load_psi(path, re="psiRe", im="meta/nx")


def test_corrupt_files_raise_only_reader_errors(tmp_path):
    """Truncations and byte overwrites of every fixture end in HDF5Error or
    KeyError (never IndexError, MemoryError or a crash)."""
    import random
    rng = random.Random(5)
    path = tmp_path / "x.h5"
    for fname in FILES:
        src = open(os.path.join(GOLD, fname), "rb").read()
        for t in range(60):
            b = bytearray(src)
            if t % 2:
                b = b[:rng.randrange(1, len(b))]
            else:
                for _ in range(rng.randrange(1, 8)):
                    b[rng.randrange(len(b))] = rng.randrange(256)
            path.write_bytes(bytes(b))
            try:
                with h5.File(path) as f:
                    f.visit()
                    f.read("psiRe")
                    f.read("psiIm")
                    f.attrs
            except (h5.HDF5Error, KeyError):
                pass
