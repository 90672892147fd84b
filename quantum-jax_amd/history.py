"""HDF5 run-history writer (save_run_to_hdf5 of quantum-fluid-accretion.py:1315-1449).

The reference stores a run's time series, radial profiles and projection
frames with h5py:

  /params/*            run parameters as attributes
  /history/<name>      1-D time series (time, M_bh, dMdt, rho_inf, ...)
  /profiles/rc, prof_dm, prof_gas   radial bins and float32 profiles
  /mach/r, profile_last
  /views_xy/dm, gas, bh_pixels      float32 projection frames
  attrs fps, created

every dataset gzip-compressed at level 4 (:1374-1376).  h5py is not part of
this package's environment, so this module writes the file format itself
(HDF5 File Format Specification, version 3) with NumPy and zlib: superblock
version 0, version-1 object headers, groups as symbol tables (version-1
B-tree, local heap and one symbol-table node), datasets chunked as a single
chunk through a version-1 B-tree with the deflate filter (level 4), or
contiguous when empty, and attributes of integer, float and fixed-length
string type.  Python bools are written as 8-bit integers (h5py writes an
enum of the same size).  Files are readable by the HDF5 library and by
``hdf5.File``.
"""

from __future__ import annotations

import os
import struct
import time
import zlib

import numpy as np

_UNDEF = 0xFFFFFFFFFFFFFFFF
_SIG = b"\x89HDF\r\n\x1a\n"
_LEAF_K = 32        # symbols per symbol-table node: 2 K
_GROUP_K = 16       # group B-tree internal node K
_CHUNK_K = 32       # chunk B-tree K (the version-0 superblock's implied value)
_HEAP_FREE_NULL = 1  # "no free block" in a local heap

# object-header message types
_DATASPACE, _DATATYPE, _FILL, _LAYOUT, _FILTERS, _ATTRIBUTE, _SYMTAB = (
    0x0001, 0x0003, 0x0005, 0x0008, 0x000B, 0x000C, 0x0011)


def _pad8(b: bytes) -> bytes:
    return b + b"\0" * (-len(b) % 8)


def _datatype(dt: np.dtype) -> bytes:
    """Datatype message body for a little-endian integer, float or
    fixed-length string."""
    dt = np.dtype(dt)
    if dt.kind == "f":
        if dt.itemsize == 4:
            sign, eloc, esize, msize, bias = 31, 23, 8, 23, 127
        elif dt.itemsize == 8:
            sign, eloc, esize, msize, bias = 63, 52, 11, 52, 1023
        else:
            raise ValueError(f"unsupported float size {dt.itemsize}")
        # class 1 version 1; little-endian, implied-1 mantissa normalisation
        head = struct.pack("<BBBBI", 0x11, 0x20, sign, 0, dt.itemsize)
        return head + struct.pack("<HHBBBBI", 0, 8 * dt.itemsize, eloc, esize, 0, msize, bias)
    if dt.kind in "iub":
        signed = 0x08 if dt.kind == "i" else 0
        head = struct.pack("<BBBBI", 0x10, signed, 0, 0, dt.itemsize)
        return head + struct.pack("<HH", 0, 8 * dt.itemsize)
    if dt.kind == "S":
        # class 3 version 1, null-terminated padding, ASCII
        return struct.pack("<BBBBI", 0x13, 0, 0, 0, dt.itemsize)
    raise ValueError(f"unsupported dtype {dt}")


def _dataspace(shape) -> bytes:
    shape = tuple(int(n) for n in shape)
    return struct.pack("<BBBBI", 1, len(shape), 0, 0, 0) + b"".join(
        struct.pack("<Q", n) for n in shape)


def _as_array(value):
    if isinstance(value, str):
        b = value.encode()
        return np.array(b, dtype=f"S{len(b) + 1}")
    a = np.asarray(value)
    if a.dtype == np.bool_:
        a = a.astype(np.uint8)
    if a.dtype.kind == "U":
        b = str(a.item()).encode()
        return np.array(b, dtype=f"S{len(b) + 1}")
    if a.dtype.kind not in "iufS":
        raise ValueError(f"unsupported value {value!r}")
    return a.astype(a.dtype.newbyteorder("<"), copy=False)


class _Node:
    def __init__(self):
        self.attrs: dict = {}


class _Group(_Node):
    def __init__(self):
        super().__init__()
        self.children: dict = {}


class _Dataset(_Node):
    def __init__(self, data: np.ndarray, level: int):
        super().__init__()
        self.data, self.level = data, level


class Writer:
    """Builds an HDF5 file in memory: groups, datasets and attributes.

    ``create_group("a/b")``, ``create_dataset("a/b/x", array)`` and
    ``attrs(path)[name] = value``; ``save(path)`` writes the file.
    """

    def __init__(self):
        self.root = _Group()

    def _node(self, path: str, make: bool = False) -> _Node:
        node = self.root
        for part in [p for p in path.split("/") if p]:
            if not isinstance(node, _Group):
                raise ValueError(f"{path}: {part} is inside a dataset")
            if part not in node.children:
                if not make:
                    raise KeyError(path)
                node.children[part] = _Group()
            node = node.children[part]
        return node

    def create_group(self, path: str) -> None:
        self._node(path, make=True)

    def create_dataset(self, path: str, data, level: int = 4) -> None:
        parent, _, name = path.rstrip("/").rpartition("/")
        g = self._node(parent, make=True)
        a = np.ascontiguousarray(_as_array(data))
        if a.dtype.kind == "S":
            raise ValueError("string datasets are not supported")
        g.children[name] = _Dataset(a, level)

    def attrs(self, path: str = "/") -> dict:
        return self._node(path).attrs

    # -- encoding ---------------------------------------------------------------
    def save(self, path: str) -> None:
        self.buf = bytearray(96)  # superblock, written last
        root_hdr, root_btree, root_heap = self._write_group(self.root)
        eof = len(self.buf)
        entry = struct.pack("<QQII", 0, root_hdr, 1, 0) + struct.pack("<QQ", root_btree, root_heap)
        sb = _SIG + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack(
            "<HHI", _LEAF_K, _GROUP_K, 0) + struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF) + entry
        assert len(sb) == 96
        self.buf[0:96] = sb
        d = os.path.dirname(os.path.abspath(path))
        os.makedirs(d, exist_ok=True)
        with open(path, "wb") as fh:
            fh.write(self.buf)
        del self.buf

    def _alloc(self, data: bytes) -> int:
        self.buf.extend(b"\0" * (-len(self.buf) % 8))
        addr = len(self.buf)
        self.buf.extend(data)
        return addr

    def _header(self, msgs) -> int:
        body = b""
        for mtype, data in msgs:
            data = _pad8(data)
            body += struct.pack("<HHBBBB", mtype, len(data), 0, 0, 0, 0) + data
        return self._alloc(struct.pack("<BBHII", 1, 0, len(msgs), 1, len(body)) + b"\0" * 4 + body)

    def _attr_msgs(self, node: _Node):
        out = []
        for name, value in node.attrs.items():
            a = _as_array(value)
            nm = name.encode() + b"\0"
            dt, ds = _datatype(a.dtype), _dataspace(a.shape)
            body = struct.pack("<BBHHH", 1, 0, len(nm), len(dt), len(ds)) + _pad8(nm) + \
                _pad8(dt) + _pad8(ds) + a.tobytes()
            out.append((_ATTRIBUTE, body))
        return out

    def _write_dataset(self, ds: _Dataset) -> int:
        a = ds.data
        msgs = [(_DATASPACE, _dataspace(a.shape)), (_DATATYPE, _datatype(a.dtype))]
        if a.size == 0:
            msgs += [(_FILL, bytes([2, 2, 2, 0])),
                     (_LAYOUT, struct.pack("<BBQQ", 3, 1, _UNDEF, 0))]
        elif a.ndim == 0:  # chunked storage needs rank >= 1: a scalar is stored as is
            addr = self._alloc(a.tobytes())
            msgs += [(_FILL, bytes([2, 2, 2, 0])),
                     (_LAYOUT, struct.pack("<BBQQ", 3, 1, addr, a.nbytes))]
        else:
            raw = zlib.compress(a.tobytes(), ds.level)
            chunk = self._alloc(raw)
            dims = list(a.shape)
            nd = len(dims) + 1

            def key(size, offs):
                return struct.pack("<II", size, 0) + b"".join(struct.pack("<Q", o) for o in offs)

            node = b"TREE" + struct.pack("<BBHQQ", 1, 0, 1, _UNDEF, _UNDEF) + \
                key(len(raw), [0] * nd) + struct.pack("<Q", chunk) + key(0, dims + [0])
            ksize = 8 + 8 * nd
            full = 24 + (2 * _CHUNK_K + 1) * ksize + 2 * _CHUNK_K * 8
            btree = self._alloc(node + b"\0" * (full - len(node)))
            layout = struct.pack("<BBBQ", 3, 2, nd, btree) + b"".join(
                struct.pack("<I", n) for n in dims + [a.dtype.itemsize])
            # deflate (filter 1), one client value: the level
            filt = struct.pack("<BBHI", 1, 1, 0, 0) + struct.pack("<HHHHII", 1, 0, 0, 1,
                                                                   ds.level, 0)
            msgs += [(_FILL, bytes([2, 3, 2, 0])), (_LAYOUT, layout), (_FILTERS, filt)]
        return self._header(msgs + self._attr_msgs(ds))

    def _write_group(self, g: _Group):
        entries = []
        for name in sorted(g.children, key=lambda s: s.encode()):
            child = g.children[name]
            if isinstance(child, _Group):
                hdr, bt, hp = self._write_group(child)
                entries.append((name, hdr, 1, bt, hp))
            else:
                entries.append((name, self._write_dataset(child), 0, 0, 0))
        if len(entries) > 2 * _LEAF_K:
            raise ValueError(f"a group holds at most {2 * _LEAF_K} objects here")
        # local heap: offset 0 is the empty name, then each name null-terminated
        heap_data = bytearray(8)
        offs = []
        for name, *_ in entries:
            offs.append(len(heap_data))
            heap_data.extend(_pad8(name.encode() + b"\0"))
        heap_data.extend(b"\0" * 8)  # room for the library's own additions
        seg = self._alloc(bytes(heap_data))
        heap = self._alloc(b"HEAP" + bytes([0, 0, 0, 0]) +
                           struct.pack("<QQQ", len(heap_data), _HEAP_FREE_NULL, seg))
        snod = b"SNOD" + struct.pack("<BBH", 1, 0, len(entries))
        for (name, hdr, cache, bt, hp), off in zip(entries, offs):
            snod += struct.pack("<QQII", off, hdr, cache, 0) + struct.pack("<QQ", bt, hp)
        snod_full = 8 + 2 * _LEAF_K * 40
        snod_addr = self._alloc(snod + b"\0" * (snod_full - len(snod))) if entries else _UNDEF
        if entries:
            node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, _UNDEF, _UNDEF) + \
                struct.pack("<QQQ", 0, snod_addr, offs[-1])
        else:  # an empty group: a B-tree with no children
            node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 0, _UNDEF, _UNDEF) + struct.pack("<Q", 0)
        full = 24 + (2 * _GROUP_K + 1) * 8 + 2 * _GROUP_K * 8
        btree = self._alloc(node + b"\0" * (full - len(node)))
        hdr = self._header([(_SYMTAB, struct.pack("<QQ", btree, heap))] + self._attr_msgs(g))
        return hdr, btree, heap


# names of the reference's history series, in its order (:1370-1398)
HISTORY_KEYS = ("time", "M_bh", "dMdt", "rho_inf", "r_B", "d_bh_peak", "vxbh", "vybh", "vzbh",
                "vbh", "M_box_dm", "M_box_gas", "M_box_tot", "M_soliton", "rc_fit", "rho0_fit",
                "dMdt_real", "cap_ratio", "r_acc", "triax_q", "triax_s", "triax_T", "triax_E",
                "triax_npts", "rho_center_cell", "rho_center_shell", "mach_at_rc",
                "alpha_target", "alpha_real")


def _stack(arrs, dtype=None):
    """_stack_list of quantum-fluid-accretion.py:1302-1313 (numeric only)."""
    if arrs is None or len(arrs) == 0:
        return None
    a = np.stack([np.asarray(x) for x in arrs], axis=0)
    return a.astype(dtype, copy=False) if dtype is not None else a


def save_run_to_hdf5(h5_path: str, H: dict, V: dict, fps: int, params: dict | None = None,
                     created: str | None = None) -> None:
    """Write a run history in the layout of quantum-fluid-accretion.py:1315-1398.

    H: history dict (series under HISTORY_KEYS, plus "rc", "prof_dm",
    "prof_gas", "mach_rc", "mach_profile_last"); V: views dict ("dm_xy",
    "g_xy", "bh_xy"); params: the /params attributes (the reference's run
    constants, :1337-1368); created: the root "created" attribute (default:
    the local time as the reference writes it)."""
    w = Writer()
    w.create_group("params")
    for k, v in (params or {}).items():
        w.attrs("params")[k] = v
    w.create_group("history")
    for k in HISTORY_KEYS:
        if H.get(k) is not None:
            w.create_dataset(f"history/{k}", np.asarray(H[k]))
    w.create_group("profiles")
    if H.get("rc") is not None:
        w.create_dataset("profiles/rc", np.asarray(H["rc"]))
    for k in ("prof_dm", "prof_gas"):
        a = _stack(H.get(k, []), np.float32)
        if a is not None:
            w.create_dataset(f"profiles/{k}", a)
    w.create_group("mach")
    mr, mp = np.asarray(H.get("mach_rc", [])), np.asarray(H.get("mach_profile_last", []))
    if mr.size:
        w.create_dataset("mach/r", mr)
    if mp.size:
        w.create_dataset("mach/profile_last", mp)
    w.create_group("views_xy")
    for k, name in (("dm_xy", "dm"), ("g_xy", "gas")):
        a = _stack(V.get(k, []), np.float32)
        if a is not None:
            w.create_dataset(f"views_xy/{name}", a)
    bh = np.asarray(V.get("bh_xy", []), dtype=np.float32)
    if bh.size:
        w.create_dataset("views_xy/bh_pixels", bh)
    w.attrs()["fps"] = np.int64(int(fps))
    w.attrs()["created"] = created if created is not None else time.strftime("%Y-%m-%d %H:%M:%S")
    w.save(h5_path)
