"""Reader for the HDF5 initial-condition files of the cosmological variant.

The reference loads its initial wavefunction with h5py from an HDF5 file
whose root holds the real and imaginary parts as the datasets ``psiRe`` and
``psiIm``, and multiplies it by 1e5 (quantum-cosmo.py:242-244);
data/explore.py:8-16 lists such a file's objects and attributes.  h5py is
not part of this package's environment, so this module reads the file
format itself (HDF5 File Format Specification, version 3), with NumPy and
zlib only, for the storage such files use:

- superblocks 0-3 at offset 0 or at the 512-byte multiples of a user block;
- version-1 and version-2 object headers, with continuation blocks;
- groups as symbol tables (version-1 B-tree, local heap, symbol table
  nodes) or as link messages in the group's header;
- dataspaces of rank 0-32; little- and big-endian integers and floats and
  fixed-length strings;
- compact, contiguous and chunked (version-1 B-tree index) layouts, with the
  shuffle, deflate and Fletcher-32 filters.

Anything else (dense link storage with fractal heaps, version-4 chunk
indexes, virtual datasets, variable-length types, other filters) raises
HDF5Error naming what it met.  Data are read through a memory map, so a
256^3 IC does not pass through a Python bytes copy.
"""

from __future__ import annotations

import mmap
import zlib

import numpy as np

_SIGNATURE = b"\x89HDF\r\n\x1a\n"

# object-header message types used here
_MSG_DATASPACE = 0x0001
_MSG_LINK_INFO = 0x0002
_MSG_DATATYPE = 0x0003
_MSG_LINK = 0x0006
_MSG_LAYOUT = 0x0008
_MSG_FILTERS = 0x000B
_MSG_ATTRIBUTE = 0x000C
_MSG_CONTINUATION = 0x0010
_MSG_SYMBOL_TABLE = 0x0011

# a dataspace claiming more than this is taken as corrupt
_MAX_DATASET_BYTES = 1 << 38

_FILTER_DEFLATE = 1
_FILTER_SHUFFLE = 2
_FILTER_FLETCHER32 = 3


class HDF5Error(ValueError):
    """The file is not HDF5, or uses a feature this reader does not cover."""


def _guarded(fn):
    """Reads that run past the end of a truncated or corrupt file, or into
    a corrupt deflate stream, raise HDF5Error rather than IndexError."""

    def wrapper(self, *args, **kwargs):
        try:
            return fn(self, *args, **kwargs)
        except HDF5Error:
            raise
        except (IndexError, OverflowError, ValueError, MemoryError, RecursionError,
                zlib.error) as e:
            raise HDF5Error(f"{self.path}: truncated or corrupt file ({e})") from None

    wrapper.__name__, wrapper.__doc__ = fn.__name__, fn.__doc__
    return wrapper


class _Header:
    """Messages of one object header: a list of (type, offset, size)."""

    def __init__(self, msgs):
        self.msgs = msgs

    def first(self, mtype):
        for t, off, size in self.msgs:
            if t == mtype:
                return off, size
        return None

    def all(self, mtype):
        return [(off, size) for t, off, size in self.msgs if t == mtype]


class File:
    """An HDF5 file opened read-only.  ``f[name]`` reads a dataset (a path of
    ``/``-separated names from the root) into a NumPy array."""

    def __init__(self, path):
        self.path = str(path)
        self.continuation_blocks = 0  # object-header continuations parsed so far
        self._fh = open(self.path, "rb")
        try:
            self._mm = mmap.mmap(self._fh.fileno(), 0, access=mmap.ACCESS_READ)
        except ValueError:  # an empty file cannot be mapped
            self._fh.close()
            raise HDF5Error(f"{self.path}: empty file, not HDF5") from None
        try:
            self._read_superblock()
        except Exception:
            self.close()
            raise

    # -- plumbing -------------------------------------------------------------

    def close(self):
        if getattr(self, "_mm", None) is not None:
            self._mm.close()
            self._mm = None
        if getattr(self, "_fh", None) is not None:
            self._fh.close()
            self._fh = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def _u(self, off, n):
        if off < 0 or off + n > len(self._mm):
            raise HDF5Error(f"{self.path}: read of {n} bytes at {off} beyond the end of the file")
        return int.from_bytes(self._mm[off:off + n], "little")

    def _bytes(self, off, n):
        if off < 0 or off + n > len(self._mm):
            raise HDF5Error(f"{self.path}: read of {n} bytes at {off} beyond the end of the file")
        return self._mm[off:off + n]

    def _addr(self, off):
        """An address field (size-of-offsets bytes); None when undefined."""
        v = self._u(off, self._so)
        return None if v == self._undef else v + self._base

    def _sig(self, off, sig):
        if self._bytes(off, 4) != sig:
            raise HDF5Error(f"{self.path}: expected {sig!r} at {off}")

    # -- superblock -------------------------------------------------------------

    @_guarded
    def _read_superblock(self):
        mm = self._mm
        at = 0
        while at + 8 <= len(mm) and mm[at:at + 8] != _SIGNATURE:
            at = 512 if at == 0 else 2 * at
        if at + 8 > len(mm):
            head = mm[:40]
            hint = " (a Git LFS pointer, not the data)" if head.startswith(b"version https://git-lfs") else ""
            raise HDF5Error(f"{self.path}: no HDF5 signature{hint}")
        ver = mm[at + 8]
        if ver in (0, 1):
            self._so, self._sl = mm[at + 13], mm[at + 14]
            p = at + 24 + (4 if ver == 1 else 0)
            self._undef = (1 << (8 * self._so)) - 1
            self._base = self._u(p, self._so)
            p += 4 * self._so  # base, free-space, end-of-file and driver addresses
            # root group symbol table entry: link name offset, object header address
            self._root = self._addr(p + self._so)
        elif ver in (2, 3):
            self._so, self._sl = mm[at + 9], mm[at + 10]
            self._undef = (1 << (8 * self._so)) - 1
            self._base = self._u(at + 12, self._so)
            self._root = self._addr(at + 12 + 3 * self._so)
        else:
            raise HDF5Error(f"{self.path}: superblock version {ver} is not supported")
        if self._so not in (2, 4, 8) or self._sl not in (2, 4, 8):
            raise HDF5Error(f"{self.path}: offset/length sizes {self._so}/{self._sl} not supported")
        self.superblock_version = ver

    # -- object headers -------------------------------------------------------

    def _header(self, addr):
        mm = self._mm
        msgs = []
        if self._bytes(addr, 4) == b"OHDR":
            if mm[addr + 4] != 2:
                raise HDF5Error(f"{self.path}: object header version {mm[addr + 4]} at {addr}")
            flags = mm[addr + 5]
            p = addr + 6
            if flags & 0x20:
                p += 16  # access, modification, change and birth times
            if flags & 0x10:
                p += 4  # attribute phase-change values
            csize = 1 << (flags & 3)
            size0 = self._u(p, csize)
            p += csize
            blocks = [(p, p + size0)]
            seen = set()
            order = 2 if flags & 0x04 else 0
            while blocks:
                b, e = blocks.pop(0)
                q = b
                while q + 4 + order <= e:
                    t, size = mm[q], self._u(q + 1, 2)
                    q += 4 + order
                    if q + size > e:
                        raise HDF5Error(f"{self.path}: message of {size} bytes overruns its block at {q}")
                    if t == _MSG_CONTINUATION:
                        c, n = self._addr(q), self._u(q + self._so, self._sl)
                        self._sig(c, b"OCHK")
                        if c in seen:
                            raise HDF5Error(f"{self.path}: continuation loop at {c}")
                        seen.add(c)
                        self.continuation_blocks += 1
                        blocks.append((c + 4, c + n - 4))  # signature ... checksum
                    elif t != 0:
                        msgs.append((t, q, size))
                    q += size
            return _Header(msgs)
        ver = mm[addr]
        if ver != 1:
            raise HDF5Error(f"{self.path}: object header version {ver} at {addr}")
        nmsg = self._u(addr + 2, 2)
        hsize = self._u(addr + 8, 4)
        blocks = [(addr + 16, addr + 16 + hsize)]
        seen = 0
        while blocks and seen < nmsg:
            b, e = blocks.pop(0)
            q = b
            while q + 8 <= e and seen < nmsg:
                t, size = self._u(q, 2), self._u(q + 2, 2)
                q += 8
                if q + size > e:
                    raise HDF5Error(f"{self.path}: message of {size} bytes overruns its block at {q}")
                seen += 1
                if t == _MSG_CONTINUATION:
                    c, n = self._addr(q), self._u(q + self._so, self._sl)
                    blocks.append((c, c + n))
                    self.continuation_blocks += 1
                elif t != 0:
                    msgs.append((t, q, size))
                q += size
        return _Header(msgs)

    # -- groups ---------------------------------------------------------------

    def _links(self, hdr):
        """{name: object header address} of a group's header."""
        out = {}
        st = hdr.first(_MSG_SYMBOL_TABLE)
        if st is not None:
            btree, heap = self._addr(st[0]), self._addr(st[0] + self._so)
            self._sig(heap, b"HEAP")
            data = self._addr(heap + 8 + 2 * self._sl)
            for name_off, obj in self._group_btree(btree):
                end = self._mm.find(b"\x00", data + name_off)
                out[self._bytes(data + name_off, end - data - name_off).decode()] = obj
        li = hdr.first(_MSG_LINK_INFO)
        if li is not None:
            flags = self._mm[li[0] + 1]
            heap = self._addr(li[0] + 2 + (8 if flags & 1 else 0))
            if heap is not None:
                raise HDF5Error(f"{self.path}: dense link storage (a group with more links than "
                                "fit in its header) is not supported")
        for off, size in hdr.all(_MSG_LINK):
            name, obj = self._link(off)
            if obj is not None:
                out[name] = obj
        return out

    def _group_btree(self, addr):
        self._sig(addr, b"TREE")
        mm = self._mm
        if mm[addr + 4] != 0:
            raise HDF5Error(f"{self.path}: B-tree at {addr} is not a group tree")
        level, used = mm[addr + 5], self._u(addr + 6, 2)
        p = addr + 8 + 2 * self._so + self._sl  # siblings, then key 0
        for _ in range(used):
            child = self._addr(p)
            p += self._so + self._sl
            if level > 0:
                yield from self._group_btree(child)
                continue
            self._sig(child, b"SNOD")
            nsym = self._u(child + 6, 2)
            e = child + 8
            for _ in range(nsym):
                yield self._u(e, self._so), self._addr(e + self._so)
                e += 2 * self._so + 24

    def _link(self, off):
        mm = self._mm
        if mm[off] != 1:
            raise HDF5Error(f"{self.path}: link message version {mm[off]}")
        flags = mm[off + 1]
        p = off + 2
        ltype = 0
        if flags & 0x08:
            ltype = mm[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        ls = 1 << (flags & 3)
        n = self._u(p, ls)
        p += ls
        name = self._bytes(p, n).decode()
        p += n
        if ltype != 0:  # soft and external links name no object in this file
            return name, None
        return name, self._addr(p)

    def _resolve(self, name):
        addr = self._root
        parts = [s for s in str(name).split("/") if s]
        for i, part in enumerate(parts):
            links = self._links(self._header(addr))
            if part not in links:
                where = "/" + "/".join(parts[:i])
                raise KeyError(f"{self.path}: no object {part!r} in group {where!r}")
            addr = links[part]
        return addr

    @_guarded
    def keys(self):
        """Names in the root group."""
        return list(self._links(self._header(self._root)))

    @_guarded
    def visit(self):
        """[(path, kind, attrs)] of every object below the root, depth first in
        name order, like the listing of data/explore.py:11-16 (kind is
        "group" or "dataset")."""
        out = []

        def walk(addr, prefix):
            for name, child in sorted(self._links(self._header(addr)).items()):
                h = self._header(child)
                path = prefix + name
                kind = "dataset" if h.first(_MSG_LAYOUT) else "group"
                out.append((path, kind, self._attrs(h)))
                if kind == "group":
                    walk(child, path + "/")

        walk(self._root, "")
        return out

    @property
    @_guarded
    def attrs(self):
        """Attributes of the root group."""
        return self._attrs(self._header(self._root))

    @_guarded
    def attrs_of(self, name):
        return self._attrs(self._header(self._resolve(name)))

    # -- messages -------------------------------------------------------------

    def _dataspace(self, off):
        mm = self._mm
        ver, rank, flags = mm[off], mm[off + 1], mm[off + 2]
        if ver == 1:
            p = off + 8
        elif ver == 2:
            if mm[off + 3] == 2:
                return None  # null dataspace: no elements
            p = off + 4
        else:
            raise HDF5Error(f"{self.path}: dataspace version {ver}")
        return tuple(self._u(p + i * self._sl, self._sl) for i in range(rank))

    def _datatype(self, off):
        mm = self._mm
        cls, ver = mm[off] & 0x0F, mm[off] >> 4
        b0 = mm[off + 1]
        size = self._u(off + 4, 4)
        if ver not in (1, 2, 3):
            raise HDF5Error(f"{self.path}: datatype version {ver}")
        if cls in (0, 1):
            if b0 & 0x40:
                raise HDF5Error(f"{self.path}: VAX byte order is not supported")
            order = ">" if b0 & 1 else "<"
            if cls == 0:
                if size not in (1, 2, 4, 8):
                    raise HDF5Error(f"{self.path}: {size}-byte integers are not supported")
                return np.dtype(f"{order}{'i' if b0 & 0x08 else 'u'}{size}")
            if size not in (2, 4, 8):
                raise HDF5Error(f"{self.path}: {size}-byte floats are not supported")
            return np.dtype(f"{order}f{size}")
        if cls == 3:
            return np.dtype(f"S{size}")
        names = {2: "time", 4: "bitfield", 5: "opaque", 6: "compound", 7: "reference",
                 8: "enumerated", 9: "variable-length", 10: "array"}
        raise HDF5Error(f"{self.path}: datatype class {names.get(cls, cls)} is not supported")

    def _filters(self, hdr):
        fm = hdr.first(_MSG_FILTERS)
        if fm is None:
            return []
        mm = self._mm
        off = fm[0]
        ver, n = mm[off], mm[off + 1]
        p = off + (8 if ver == 1 else 2)
        out = []
        for _ in range(n):
            fid = self._u(p, 2)
            if ver == 1 or fid >= 256:
                nlen = self._u(p + 2, 2)
                p += 4
            else:
                nlen = 0
                p += 2
            flags, nval = self._u(p, 2), self._u(p + 2, 2)
            p += 4
            p += ((nlen + 7) // 8) * 8 if ver == 1 else nlen
            vals = [self._u(p + 4 * i, 4) for i in range(nval)]
            p += 4 * nval + (4 if ver == 1 and nval % 2 else 0)
            if fid not in (_FILTER_DEFLATE, _FILTER_SHUFFLE, _FILTER_FLETCHER32):
                raise HDF5Error(f"{self.path}: filter {fid} is not supported")
            out.append((fid, vals))
        return out

    def _unfilter(self, raw, filters, mask, itemsize):
        data = bytes(raw)
        for i in range(len(filters) - 1, -1, -1):
            if mask & (1 << i):
                continue
            fid, vals = filters[i]
            if fid == _FILTER_DEFLATE:
                data = zlib.decompress(data)
            elif fid == _FILTER_FLETCHER32:
                data = data[:-4]
            elif fid == _FILTER_SHUFFLE:
                es = vals[0] if vals else itemsize
                if es > 1:
                    n = len(data) // es
                    body = np.frombuffer(data, np.uint8, n * es).reshape(es, n).T.tobytes()
                    data = body + data[n * es:]
        return data

    def _attrs(self, hdr):
        out = {}
        mm = self._mm
        for off, size in hdr.all(_MSG_ATTRIBUTE):
            ver = mm[off]
            nsz, tsz, ssz = self._u(off + 2, 2), self._u(off + 4, 2), self._u(off + 6, 2)
            p = off + (9 if ver == 3 else 8)
            pad = (lambda n: (n + 7) // 8 * 8) if ver == 1 else (lambda n: n)
            name = self._bytes(p, nsz).rstrip(b"\x00").decode()
            p += pad(nsz)
            try:
                dt = self._datatype(p)
                shape = self._dataspace(p + pad(tsz))
            except HDF5Error as e:
                out[name] = f"<{e}>"
                continue
            p += pad(tsz) + pad(ssz)
            if shape is None:
                out[name] = None
                continue
            count = int(np.prod(shape)) if shape else 1
            a = np.frombuffer(self._bytes(p, count * dt.itemsize), dt, count)
            a = a.astype(dt.newbyteorder("=")) if dt.kind in "iuf" else a.copy()
            out[name] = a.reshape(shape) if shape else a[0]
        return out

    # -- datasets -------------------------------------------------------------

    def __getitem__(self, name):
        return self.read(name)

    @_guarded
    def read(self, name):
        hdr = self._header(self._resolve(name))
        lay = hdr.first(_MSG_LAYOUT)
        if lay is None:
            raise HDF5Error(f"{self.path}: {name!r} is a group, not a dataset")
        ds, dt = hdr.first(_MSG_DATASPACE), hdr.first(_MSG_DATATYPE)
        if ds is None or dt is None:
            raise HDF5Error(f"{self.path}: dataset {name!r} lacks its dataspace or datatype")
        shape = self._dataspace(ds[0])
        dtype = self._datatype(dt[0])
        native = dtype.newbyteorder("=") if dtype.kind in "iuf" else dtype
        if shape is None:
            return np.empty((0,), native)
        count = int(np.prod(shape, dtype=object)) if shape else 1
        nbytes = count * dtype.itemsize
        if nbytes > _MAX_DATASET_BYTES:
            raise HDF5Error(f"{self.path}: {name!r} claims {nbytes} bytes")
        mm = self._mm
        off = lay[0]
        ver = mm[off]
        if ver in (3, 4):
            cls = mm[off + 1]
            if cls == 0:
                n = self._u(off + 2, 2)
                a = np.array(np.frombuffer(self._bytes(off + 4, n), dtype, count), native)
            elif cls == 1:
                addr = self._addr(off + 2)
                if addr is None:  # never written: the fill value, zero by default
                    return np.zeros(shape, native)
                a = self._copy_out(addr, dtype, count, native)
            elif cls == 2 and ver == 3:
                rank = mm[off + 2] - 1
                btree = self._addr(off + 3)
                p = off + 3 + self._so
                chunk = tuple(self._u(p + 4 * i, 4) for i in range(rank))
                return self._read_chunked(btree, shape, chunk, dtype, hdr).astype(native)
            else:
                kinds = {2: "chunked (version-4 index)", 3: "virtual"}
                raise HDF5Error(f"{self.path}: {kinds.get(cls, cls)} layout of {name!r} "
                                "is not supported")
        elif ver in (1, 2):
            rank, cls = mm[off + 1], mm[off + 2]
            p = off + 8
            if cls == 1:
                addr = self._addr(p)
                if addr is None:
                    return np.zeros(shape, native)
                a = self._copy_out(addr, dtype, count, native)
            elif cls == 2:
                btree = self._addr(p)
                p += self._so
                chunk = tuple(self._u(p + 4 * i, 4) for i in range(rank - 1))
                return self._read_chunked(btree, shape, chunk, dtype, hdr).astype(native)
            elif cls == 0:
                p += 4 * rank
                n = self._u(p, 4)
                a = np.array(np.frombuffer(self._bytes(p + 4, n), dtype, count), native)
            else:
                raise HDF5Error(f"{self.path}: layout class {cls} of {name!r}")
        else:
            raise HDF5Error(f"{self.path}: layout message version {ver}")
        if a.size != count:
            raise HDF5Error(f"{self.path}: {name!r} holds {a.size} elements, expected {count}")
        return a.reshape(shape)

    def _copy_out(self, addr, dtype, count, native):
        """count elements at addr, copied out of the map in native byte order
        (the temporary view of the map is gone when this returns, so the
        file can be closed)."""
        if addr + count * dtype.itemsize > len(self._mm):
            raise HDF5Error(f"{self.path}: data at {addr} run past the end of the file")
        return np.array(np.frombuffer(self._mm, dtype, count, addr), native)

    def _read_chunked(self, btree, shape, chunk, dtype, hdr):
        rank = len(shape)
        if len(chunk) != rank:
            raise HDF5Error(f"{self.path}: chunk rank {len(chunk)} vs dataset rank {rank}")
        filters = self._filters(hdr)
        out = np.zeros(shape, dtype)
        ccount = int(np.prod(chunk))
        if btree is None:
            return out
        for size, mask, start, addr in self._chunk_btree(btree, rank):
            raw = self._unfilter(self._bytes(addr, size), filters, mask, dtype.itemsize)
            if len(raw) < ccount * dtype.itemsize:
                raise HDF5Error(f"{self.path}: chunk at {addr} holds {len(raw)} bytes, "
                                f"expected {ccount * dtype.itemsize}")
            c = np.frombuffer(raw, dtype, ccount).reshape(chunk)
            dst = tuple(slice(s, min(s + n, d)) for s, n, d in zip(start, chunk, shape))
            src = tuple(slice(0, d.stop - d.start) for d in dst)
            out[dst] = c[src]
        return out

    def _chunk_btree(self, addr, rank):
        self._sig(addr, b"TREE")
        mm = self._mm
        if mm[addr + 4] != 1:
            raise HDF5Error(f"{self.path}: B-tree at {addr} is not a chunk tree")
        level, used = mm[addr + 5], self._u(addr + 6, 2)
        ksize = 8 + 8 * (rank + 1)
        p = addr + 8 + 2 * self._so
        for _ in range(used):
            size, mask = self._u(p, 4), self._u(p + 4, 4)
            start = tuple(self._u(p + 8 + 8 * i, 8) for i in range(rank))
            child = self._addr(p + ksize)
            p += ksize + self._so
            if level > 0:
                yield from self._chunk_btree(child, rank)
            else:
                yield size, mask, start, child


def read(path, name):
    """One dataset of an HDF5 file as a NumPy array (native byte order)."""
    with File(path) as f:
        return f.read(name)


def load_psi(path, scale: float = 1e5, re: str = "psiRe", im: str = "psiIm") -> np.ndarray:
    """The reference's cosmological initial condition, complex128 on the host:
    psi = (psiRe + i psiIm) * scale (quantum-cosmo.py:242-244, where the file
    is in units of 1e10 Msun/kpc^2 and scale = 1e5).  The reference forms psi
    in complex64 (JAX's default precision); here the product is taken in
    FP64, so the two agree to complex64 rounding."""
    with File(path) as f:
        r, i = f.read(re), f.read(im)
    if r.shape != i.shape:
        raise HDF5Error(f"{path}: {re} {r.shape} and {im} {i.shape} differ in shape")
    psi = r.astype(np.float64) + 1j * i.astype(np.float64)
    psi *= scale
    return psi
