"""Write the HDF5 initial-condition fixtures of tests/test_hdf5.py.

The reference reads its cosmological initial condition from an HDF5 file
holding the datasets ``psiRe`` and ``psiIm`` at the root
(quantum-cosmo.py:242-244; data/explore.py lists such a file).  Its data file
is not available, so these fixtures are small files of the same form, written
by the real HDF5 library in the storage variants a reader meets in practice.

Run with an interpreter that has h5py (this image: ``/opt/conda/bin/python3.9``,
h5py 3.3.0 on HDF5 1.10.6) from the repository root:

    /opt/conda/bin/python3.9 tests/golden/make_hdf5_fixtures.py

Every file's arrays are also saved, as written, in ``hdf5_expected.npz``.
"""

import os

import h5py
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def field(shape, seed, dtype):
    rng = np.random.default_rng(seed)
    return rng.standard_normal(shape).astype(dtype)


def main():
    expected = {}

    def keep(fname, name, a):
        expected[f"{fname}:{name}"] = np.asarray(a)

    # 1. h5py defaults (superblock 0, version-1 object headers, symbol-table
    #    groups, contiguous storage), as a NumPy float32 IC would be written
    f1 = "ic_default.h5"
    re, im = field((6, 5, 4), 1, np.float32), field((6, 5, 4), 2, np.float32)
    with h5py.File(os.path.join(HERE, f1), "w") as f:
        f["psiRe"] = re
        f["psiIm"] = im
        f.attrs["redshift"] = 127.0
        f.attrs["box_kpc"] = np.array([1000.0, 1000.0, 1000.0])
        g = f.create_group("meta")
        g["nx"] = np.array([6], dtype=np.int32)
        g.attrs["code"] = np.bytes_("fdm-ic")
    # attributes added to a closed dataset land in a continuation block
    with h5py.File(os.path.join(HERE, f1), "a") as f:
        for k in range(12):
            f["psiRe"].attrs[f"note{k:02d}"] = np.float32(k) / 4
    keep(f1, "psiRe", re)
    keep(f1, "psiIm", im)
    keep(f1, "meta/nx", np.array([6], dtype=np.int32))

    # 2. chunked storage with shuffle + deflate, float64 and a big-endian
    #    dataset, chunks that do not divide the shape
    f2 = "ic_chunked_gzip.h5"
    re, im = field((7, 5, 6), 3, np.float64), field((7, 5, 6), 4, np.float64)
    with h5py.File(os.path.join(HERE, f2), "w") as f:
        f.create_dataset("psiRe", data=re, chunks=(3, 2, 4), compression="gzip",
                         compression_opts=4, shuffle=True)
        f.create_dataset("psiIm", data=im.astype(">f8"), chunks=(4, 5, 3))
    keep(f2, "psiRe", re)
    keep(f2, "psiIm", im)

    # 3. the newest format (superblock 3, version-2 object headers, link
    #    messages; seven links, within the eight that stay in the group's
    #    header), and a compact dataset
    f3 = "ic_latest.h5"
    re, im = field((4, 4, 4), 5, np.float32), field((4, 4, 4), 6, np.float32)
    with h5py.File(os.path.join(HERE, f3), "w", libver="latest") as f:
        for k in range(4):
            f[f"aux{k}"] = np.arange(k + 1, dtype=np.int64)
        f["psiRe"] = re
        f["psiIm"] = im
        sp = h5py.h5s.create_simple((3,))
        plist = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        plist.set_layout(h5py.h5d.COMPACT)
        dset = h5py.h5d.create(f.id, b"small", h5py.h5t.IEEE_F64LE, sp, plist)
        small = np.array([1.5, -2.25, 3.0])
        dset.write(h5py.h5s.ALL, h5py.h5s.ALL, small)
    with h5py.File(os.path.join(HERE, f3), "a", libver="latest") as f:
        for k in range(6):
            f["psiIm"].attrs[f"note{k}"] = np.int16(-k)
    keep(f3, "psiRe", re)
    keep(f3, "psiIm", im)
    keep(f3, "small", small)
    for k in range(4):
        keep(f3, f"aux{k}", np.arange(k + 1, dtype=np.int64))

    # 4. a 512-byte user block ahead of the superblock (addresses relative to it)
    f4 = "ic_userblock.h5"
    re, im = field((2, 3, 4), 7, np.float32), field((2, 3, 4), 8, np.float32)
    with h5py.File(os.path.join(HERE, f4), "w", userblock_size=512) as f:
        f["psiRe"] = re
        f["psiIm"] = im
    keep(f4, "psiRe", re)
    keep(f4, "psiIm", im)

    np.savez(os.path.join(HERE, "hdf5_expected.npz"), **expected)


if __name__ == "__main__":
    main()
