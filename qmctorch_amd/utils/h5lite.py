"""Pure-Python reader / writer for the subset of HDF5 that QMCTorch's files use.

The reference reads and writes its molecule caches, observables and trained
models through h5py (utils/hdf5_utils.py:27-483, wavefunction/wf_base.py:257-277,
scf/molecule.py:394-424); h5py is not importable for this interpreter, so this
module implements the part of the HDF5 File Format Specification (version 3.0)
those files need, with an h5py-like surface (File, Group, Dataset, attrs).

Read:  superblock versions 0 / 1 with 8-byte offsets and lengths; version-1
       object headers (with continuation blocks); groups stored as symbol
       tables (version-1 B-tree of SNOD leaves + a local heap of names) or as
       compact link messages; datasets with contiguous or compact storage;
       attribute messages (versions 1-3); datatypes fixed-point, IEEE floating
       point, fixed-length strings, enumerations (h5py stores bool as the enum
       FALSE/TRUE over int8) and variable-length strings (global heap).
Write: the same layout h5py produces with its default ("earliest") format:
       superblock 0, version-1 object headers, symbol-table groups, contiguous
       datasets; strings are written as fixed-length byte strings.
Not supported (reading raises H5Error): chunked or filtered storage,
version-2 object headers, dense link / attribute storage (fractal heaps),
compound, array, reference and variable-length sequence types.

A File opened with mode "a", "r+" or "w" holds the whole tree in memory and
rewrites the file on close(); QMCTorch's files are small (kilobytes to a few
megabytes) and are written once per run.
"""
import os
from typing import Dict, Iterator, Tuple

import numpy as np

UNDEF = 0xFFFFFFFFFFFFFFFF
SIGNATURE = b"\x89HDF\r\n\x1a\n"
_INTERNAL_K = 16  # group B-tree internal node K (HDF5 default)


class H5Error(Exception):
    """The file uses an HDF5 feature outside the supported subset, or is corrupt."""


# ---------------------------------------------------------------- tree model
class Dataset:
    """An n-dimensional array with attributes (h5py.Dataset subset)."""

    def __init__(self, data, attrs=None):
        self._data = data
        self.attrs: Dict[str, object] = dict(attrs or {})

    def __getitem__(self, key):
        d = self._data
        if key is Ellipsis or (isinstance(key, tuple) and len(key) == 0):
            if isinstance(d, np.ndarray) and d.ndim == 0:
                return d[()]
            return d
        return np.asarray(d)[key]

    def __array__(self, dtype=None, copy=None):
        a = np.asarray(self._data)
        return a.astype(dtype) if dtype is not None else a

    def __len__(self):
        return len(self._data)

    @property
    def shape(self) -> Tuple[int, ...]:
        return np.shape(self._data)

    @property
    def dtype(self):
        return np.asarray(self._data).dtype


class Group:
    """Named children (groups and datasets) and attributes (h5py.Group subset).
    Iteration is in byte order of the names, as h5py lists symbol-table groups."""

    def __init__(self, attrs=None):
        self._children: Dict[str, object] = {}
        self.attrs: Dict[str, object] = dict(attrs or {})

    # -- path access
    def _walk(self, path: str, create: bool = False):
        node = self
        parts = [p for p in path.split("/") if p]
        for i, p in enumerate(parts[:-1]):
            nxt = node._children.get(p) if isinstance(node, Group) else None
            if nxt is None:
                if not create:
                    raise KeyError(path)
                nxt = node._children[p] = Group()
            if not isinstance(nxt, Group):
                raise KeyError("%s is not a group" % "/".join(parts[: i + 1]))
            node = nxt
        return node, (parts[-1] if parts else "")

    def __getitem__(self, path: str):
        parent, name = self._walk(path)
        if not name:
            return parent
        if name not in parent._children:
            raise KeyError(path)
        return parent._children[name]

    def __contains__(self, path: str) -> bool:
        try:
            self[path]
            return True
        except KeyError:
            return False

    def __setitem__(self, path: str, value):
        self.create_dataset(path, data=value)

    def __delitem__(self, path: str):
        parent, name = self._walk(path)
        del parent._children[name]

    def keys(self):
        return sorted(self._children, key=lambda s: s.encode("utf-8"))

    def __iter__(self) -> Iterator[str]:
        return iter(self.keys())

    def __len__(self) -> int:
        return len(self._children)

    def items(self):
        return [(k, self._children[k]) for k in self.keys()]

    def values(self):
        return [self._children[k] for k in self.keys()]

    # -- creation
    def create_group(self, path: str) -> "Group":
        parent, name = self._walk(path, create=True)
        if name in parent._children:
            raise ValueError("Unable to create group (name already exists): %s" % path)
        g = parent._children[name] = Group()
        return g

    def require_group(self, path: str) -> "Group":
        if path in self:
            g = self[path]
            if not isinstance(g, Group):
                raise TypeError("%s exists and is not a group" % path)
            return g
        return self.create_group(path)

    def create_dataset(self, path: str, shape=None, dtype=None, data=None) -> Dataset:
        parent, name = self._walk(path, create=True)
        if name in parent._children:
            raise ValueError("Unable to create dataset (name already exists): %s" % path)
        if data is None:
            data = np.zeros(shape if shape is not None else (), dtype=dtype or np.float64)
        arr = _as_storable(data)
        if dtype is not None and arr.dtype.kind not in "SO":
            arr = arr.astype(dtype)
        ds = parent._children[name] = Dataset(arr)
        return ds


def _as_storable(value) -> np.ndarray:
    """numpy array of a supported kind (int, uint, float, bool, bytes)."""
    if isinstance(value, Dataset):
        value = value._data
    if hasattr(value, "detach"):  # torch tensor
        value = value.detach().cpu().numpy()
    a = np.asarray(value)
    if a.dtype.kind == "U":
        a = np.char.encode(a, "utf-8")
    elif a.dtype.kind == "O":
        flat = [x.encode("utf-8") if isinstance(x, str) else bytes(x) for x in a.ravel()]
        a = np.array(flat, dtype="S").reshape(a.shape) if flat else np.zeros(a.shape, dtype="S1")
    if a.dtype.kind == "S" and a.dtype.itemsize == 0:
        a = a.astype("S1")
    if a.dtype.kind not in "iufbS":
        raise TypeError("cannot store numpy dtype %s in HDF5" % a.dtype)
    return a


class File(Group):
    """HDF5 file (h5py.File subset).  Modes: "r" (read), "r+"/"a" (read if it
    exists, rewrite on close), "w" (new, truncating), "w-"/"x" (new, must not
    exist)."""

    def __init__(self, name, mode: str = "r"):
        super().__init__()
        self.filename = os.fspath(name)
        self.mode = mode
        exists = os.path.isfile(self.filename)
        if mode in ("w-", "x") and exists:
            raise FileExistsError(self.filename)
        if mode in ("r", "r+") and not exists:
            raise FileNotFoundError(self.filename)
        if mode not in ("r", "r+", "a", "w", "w-", "x"):
            raise ValueError("invalid mode %r" % mode)
        if exists and mode in ("r", "r+", "a"):
            with open(self.filename, "rb") as fh:
                root = _Reader(fh.read()).read_root()
            self._children, self.attrs = root._children, root.attrs
        self._open = True
        if mode in ("w", "w-", "x"):
            self.flush()

    def flush(self):
        if self.mode == "r":
            return
        data = _Writer().write(self)
        tmp = self.filename + ".tmp%d" % os.getpid()
        with open(tmp, "wb") as fh:
            fh.write(data)
        os.replace(tmp, self.filename)

    def close(self):
        if self._open:
            self.flush()
            self._open = False

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
        return False


# -------------------------------------------------------------------- reader
def _pad8(n: int) -> int:
    return (n + 7) & ~7


class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if buf[:8] != SIGNATURE:
            raise H5Error("not an HDF5 file (no signature at offset 0)")
        ver = buf[8]
        if ver not in (0, 1):
            raise H5Error("superblock version %d is not supported" % ver)
        if buf[13] != 8 or buf[14] != 8:
            raise H5Error("only 8-byte offsets and lengths are supported")
        p = 24 + (4 if ver == 1 else 0)
        self.base = self.u(p, 8)
        self.root_entry = p + 32
        self._gheap: Dict[int, Dict[int, bytes]] = {}

    def u(self, off: int, n: int) -> int:
        return int.from_bytes(self.b[off:off + n], "little")

    def at(self, addr: int) -> int:
        return self.base + addr

    # -- object headers
    def messages(self, addr: int):
        p = self.at(addr)
        if self.b[p:p + 4] == b"OHDR":
            raise H5Error("version-2 object headers are not supported")
        if self.b[p] != 1:
            raise H5Error("object header version %d is not supported" % self.b[p])
        blocks = [(p + 16, self.u(p + 8, 4))]
        seen = {p + 16}
        out = []
        while blocks:
            start, size = blocks.pop(0)
            if start + size > len(self.b):
                raise H5Error("object header block past the end of the file")
            q, end = start, start + size
            while q + 8 <= end:
                mtype, msize, flags = self.u(q, 2), self.u(q + 2, 2), self.b[q + 4]
                d = q + 8
                if mtype == 0x10:  # continuation
                    nxt = self.at(self.u(d, 8))
                    if nxt in seen:
                        raise H5Error("object header continuation cycle")
                    seen.add(nxt)
                    blocks.append((nxt, self.u(d + 8, 8)))
                elif mtype != 0:
                    if flags & 0x02:
                        raise H5Error("shared header messages are not supported")
                    out.append((mtype, d, msize))
                q = d + msize
        return out

    def read_root(self) -> Group:
        return self.read_object(self.u(self.root_entry + 8, 8))

    def read_object(self, addr: int):
        msgs = self.messages(addr)
        types = {m[0] for m in msgs}
        attrs = {}
        for t, d, n in msgs:
            if t == 0x0C:
                k, v = self.attribute(d)
                attrs[k] = v
            elif t == 0x15 and self.u(d + 2, 8) != UNDEF:  # attribute info with a fractal heap
                raise H5Error("dense attribute storage is not supported")
        if 0x11 in types or 0x06 in types or 0x02 in types:
            g = Group(attrs)
            for t, d, n in msgs:
                if t == 0x11:
                    for name, child in self.symbol_table(self.u(d, 8), self.u(d + 8, 8)):
                        g._children[name] = self.read_object(child)
                elif t == 0x06:
                    name, child = self.link(d)
                    if child is not None:
                        g._children[name] = self.read_object(child)
                elif t == 0x02:
                    flags = self.b[d + 1]
                    q = d + 2 + (8 if flags & 1 else 0)
                    if self.u(q, 8) != UNDEF:
                        raise H5Error("dense link storage is not supported")
            return g
        if 0x08 not in types:
            raise H5Error("object at %d is neither a group nor a dataset" % addr)
        shape = dt = lay = None
        for t, d, n in msgs:
            if t == 0x01:
                shape = self.dataspace(d)
            elif t == 0x03:
                dt = self.datatype(d)[0]
            elif t == 0x08:
                lay = (d, n)
            elif t == 0x0B:
                raise H5Error("filtered (compressed) datasets are not supported")
        return Dataset(self.dataset_data(shape, dt, *lay), attrs)

    # -- groups
    def heap_string(self, data: int, off: int) -> str:
        end = self.b.index(b"\0", data + off)
        return self.b[data + off:end].decode("utf-8")

    def symbol_table(self, btree: int, heap: int):
        hp = self.at(heap)
        if self.b[hp:hp + 4] != b"HEAP":
            raise H5Error("bad local heap signature")
        data = self.at(self.u(hp + 24, 8))
        out = []
        visited = set()

        def node(a):
            p = self.at(a)
            if p in visited:
                raise H5Error("group B-tree cycle")
            visited.add(p)
            if self.b[p:p + 4] != b"TREE" or self.b[p + 4] != 0:
                raise H5Error("bad group B-tree node")
            level, n = self.b[p + 5], self.u(p + 6, 2)
            q = p + 24
            for _ in range(n):
                child = self.u(q + 8, 8)
                q += 16
                node(child) if level > 0 else snod(child)

        def snod(a):
            p = self.at(a)
            if self.b[p:p + 4] != b"SNOD":
                raise H5Error("bad symbol table node")
            for i in range(self.u(p + 6, 2)):
                e = p + 8 + 40 * i
                out.append((self.heap_string(data, self.u(e, 8)), self.u(e + 8, 8)))

        if btree != UNDEF:
            node(btree)
        return out

    def link(self, d: int):
        flags = self.b[d + 1]
        q = d + 2
        ltype = 0
        if flags & 0x08:
            ltype = self.b[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        nlen_size = 1 << (flags & 3)
        nlen = self.u(q, nlen_size)
        q += nlen_size
        name = self.b[q:q + nlen].decode("utf-8")
        q += nlen
        if ltype != 0:  # soft / external links are not followed
            return name, None
        return name, self.u(q, 8)

    # -- dataspace / datatype / layout
    def dataspace(self, d: int):
        ver, rank = self.b[d], self.b[d + 1]
        if ver == 1:
            q = d + 8
        elif ver == 2:
            if self.b[d + 3] == 2:  # null dataspace
                return None
            q = d + 4
        else:
            raise H5Error("dataspace version %d is not supported" % ver)
        return tuple(self.u(q + 8 * i, 8) for i in range(rank))

    def datatype(self, d: int):
        """(descriptor, encoded length).  Descriptors: a numpy dtype,
        ("bool", base dtype), ("enum", base dtype) or ("vstr",)."""
        cls = self.b[d] & 0x0F
        bf = self.u(d + 1, 3)
        size = self.u(d + 4, 4)
        order = ">" if bf & 1 else "<"
        if cls == 0:
            return np.dtype("%s%s%d" % (order, "i" if bf & 0x08 else "u", size)), 12
        if cls == 1:
            if size not in (2, 4, 8):
                raise H5Error("floating point of %d bytes is not supported" % size)
            return np.dtype("%sf%d" % (order, size)), 20
        if cls == 3:
            return np.dtype("S%d" % size), 8
        if cls == 8:
            ver = self.b[d] >> 4
            base, blen = self.datatype(d + 8)
            nmem = bf & 0xFFFF
            q = d + 8 + blen
            names = []
            for _ in range(nmem):
                end = self.b.index(b"\0", q)
                names.append(self.b[q:end])
                q = d + 8 + blen + _pad8(end + 1 - (d + 8 + blen)) if ver < 3 else end + 1
            vals = np.frombuffer(self.b[q:q + nmem * base.itemsize], dtype=base)
            q += nmem * base.itemsize
            mapping = dict(zip(names, vals.tolist()))
            kind = "bool" if mapping == {b"FALSE": 0, b"TRUE": 1} else "enum"
            return (kind, base), q - d
        if cls == 9:
            if bf & 0x0F != 1:
                raise H5Error("variable-length sequences are not supported")
            _, blen = self.datatype(d + 8)
            return ("vstr",), 8 + blen
        raise H5Error("datatype class %d is not supported" % cls)

    @staticmethod
    def _itemsize(dt) -> int:
        if isinstance(dt, np.dtype):
            return dt.itemsize
        if dt[0] == "vstr":
            return 16
        return dt[1].itemsize

    def gheap_object(self, coll: int, index: int) -> bytes:
        objs = self._gheap.get(coll)
        if objs is None:
            p = self.at(coll)
            if self.b[p:p + 4] != b"GCOL":
                raise H5Error("bad global heap collection")
            end, q, objs = p + self.u(p + 8, 8), p + 16, {}
            while q + 16 <= end:
                idx, n = self.u(q, 2), self.u(q + 8, 8)
                if idx == 0:
                    break
                objs[idx] = bytes(self.b[q + 16:q + 16 + n])
                q += 16 + _pad8(n)
            self._gheap[coll] = objs
        return objs[index]

    def decode(self, raw: bytes, shape, dt):
        if shape is None:
            return np.zeros((0,))
        n = int(np.prod(shape)) if shape else 1
        if isinstance(dt, np.dtype):
            a = np.frombuffer(raw, dtype=dt, count=n).copy()
            a = a.astype(dt.newbyteorder("=")) if dt.byteorder == ">" else a
            return a.reshape(shape)
        if dt[0] in ("bool", "enum"):
            a = np.frombuffer(raw, dtype=dt[1], count=n).copy().reshape(shape)
            return a.astype(bool) if dt[0] == "bool" else a
        vals = []
        for i in range(n):
            e = raw[16 * i:16 * i + 16]
            length = int.from_bytes(e[0:4], "little")
            coll = int.from_bytes(e[4:12], "little")
            idx = int.from_bytes(e[12:16], "little")
            vals.append(b"" if length == 0 or coll in (0, UNDEF) else self.gheap_object(coll, idx)[:length])
        if shape == ():
            return np.array(vals[0], dtype=object)
        return np.array(vals, dtype=object).reshape(shape)

    def dataset_data(self, shape, dt, d: int, n: int):
        ver = self.b[d]
        nbytes = (int(np.prod(shape)) if shape else 1) * self._itemsize(dt) if shape is not None else 0
        if ver == 3:
            cls = self.b[d + 1]
            if cls == 0:
                size = self.u(d + 2, 2)
                raw = self.b[d + 4:d + 4 + size]
            elif cls == 1:
                addr = self.u(d + 2, 8)
                raw = bytes(nbytes) if addr == UNDEF else self.b[self.at(addr):self.at(addr) + nbytes]
            else:
                raise H5Error("chunked storage is not supported")
        elif ver in (1, 2):
            rank, cls = self.b[d + 1], self.b[d + 2]
            q = d + 8
            if cls == 1:
                addr = self.u(q, 8)
                raw = bytes(nbytes) if addr == UNDEF else self.b[self.at(addr):self.at(addr) + nbytes]
            elif cls == 0:
                q += 4 * rank
                size = self.u(q, 4)
                raw = self.b[q + 4:q + 4 + size]
            else:
                raise H5Error("chunked storage is not supported")
        else:
            raise H5Error("data layout version %d is not supported" % ver)
        return self.decode(bytes(raw), shape, dt)

    def attribute(self, d: int):
        ver = self.b[d]
        nlen, tlen, slen = self.u(d + 2, 2), self.u(d + 4, 2), self.u(d + 6, 2)
        if ver == 1:
            q = d + 8
            name = self.b[q:q + nlen].rstrip(b"\0").decode("utf-8")
            q += _pad8(nlen)
            dt = self.datatype(q)[0]
            q += _pad8(tlen)
            shape = self.dataspace(q)
            q += _pad8(slen)
        elif ver in (2, 3):
            q = d + 8 + (1 if ver == 3 else 0)
            name = self.b[q:q + nlen].rstrip(b"\0").decode("utf-8")
            q += nlen
            dt = self.datatype(q)[0]
            q += tlen
            shape = self.dataspace(q)
            q += slen
        else:
            raise H5Error("attribute message version %d is not supported" % ver)
        nbytes = (int(np.prod(shape)) if shape else 1) * self._itemsize(dt) if shape is not None else 0
        val = self.decode(bytes(self.b[q:q + nbytes]), shape, dt)
        if not isinstance(dt, np.dtype) and dt[0] == "vstr":  # h5py returns str for these
            val = np.vectorize(lambda x: x.decode("utf-8"), otypes=[object])(val)
        if isinstance(val, np.ndarray) and val.ndim == 0:
            val = val[()]
        return name, val


# -------------------------------------------------------------------- writer
def _u(v: int, n: int) -> bytes:
    return int(v).to_bytes(n, "little")


def _dtype_message(a: np.ndarray) -> bytes:
    k, size = a.dtype.kind, a.dtype.itemsize
    if k in "iu":
        return bytes([0x10, 0x08 if k == "i" else 0, 0, 0]) + _u(size, 4) + _u(0, 2) + _u(8 * size, 2)
    if k == "f":
        sign, eloc, esize, msize, bias = {2: (15, 10, 5, 10, 15), 4: (31, 23, 8, 23, 127),
                                          8: (63, 52, 11, 52, 1023)}[size]
        return (bytes([0x11, 0x20, sign, 0]) + _u(size, 4) + _u(0, 2) + _u(8 * size, 2)
                + bytes([eloc, esize, 0, msize]) + _u(bias, 4))
    if k == "S":
        return bytes([0x13, 0x01, 0, 0]) + _u(size, 4)  # null-padded ASCII
    if k == "b":  # h5py's bool: enum {FALSE: 0, TRUE: 1} over int8
        base = bytes([0x10, 0x08, 0, 0]) + _u(1, 4) + _u(0, 2) + _u(8, 2)
        names = b"FALSE\0\0\0" + b"TRUE\0\0\0\0"
        return bytes([0x18, 2, 0, 0]) + _u(1, 4) + base + names + bytes([0, 1])
    raise TypeError("cannot store numpy dtype %s in HDF5" % a.dtype)


def _raw(a: np.ndarray) -> bytes:
    if a.dtype.kind == "b":
        return a.astype(np.int8).tobytes()
    if a.dtype.kind in "iuf" and a.dtype.byteorder == ">":
        a = a.astype(a.dtype.newbyteorder("<"))
    return np.ascontiguousarray(a).tobytes()


def _dataspace_message(shape) -> bytes:
    return bytes([1, len(shape), 0, 0]) + bytes(4) + b"".join(_u(n, 8) for n in shape)


def _message(mtype: int, data: bytes, flags: int = 0) -> bytes:
    data = data + bytes(_pad8(len(data)) - len(data))
    if len(data) > 0xFFFF:
        raise H5Error("header message of %d bytes exceeds 64 KiB" % len(data))
    return _u(mtype, 2) + _u(len(data), 2) + bytes([flags, 0, 0, 0]) + data


def _attribute_message(name: str, value) -> bytes:
    a = _as_storable(value)
    nm = name.encode("utf-8") + b"\0"
    dt, ds = _dtype_message(a), _dataspace_message(a.shape)
    body = (bytes([1, 0]) + _u(len(nm), 2) + _u(len(dt), 2) + _u(len(ds), 2)
            + nm + bytes(_pad8(len(nm)) - len(nm)) + dt + bytes(_pad8(len(dt)) - len(dt))
            + ds + bytes(_pad8(len(ds)) - len(ds)) + _raw(a))
    return _message(0x0C, body)


class _Writer:
    def __init__(self):
        self.buf = bytearray()
        self.leaf_k = 4

    def alloc(self, data: bytes) -> int:
        self.buf += bytes(_pad8(len(self.buf)) - len(self.buf))
        addr = len(self.buf)
        self.buf += data
        return addr

    def header(self, msgs) -> int:
        body = b"".join(msgs)
        prefix = bytes([1, 0]) + _u(len(msgs), 2) + _u(1, 4) + _u(len(body), 4) + bytes(4)
        return self.alloc(prefix + body)

    def write(self, root: Group) -> bytes:
        def most(g):
            m = len(g._children)
            for c in g._children.values():
                if isinstance(c, Group):
                    m = max(m, most(c))
            return m

        self.leaf_k = max(4, (most(root) + 1) // 2)
        self.alloc(bytes(96))  # superblock, filled in below
        hdr, btree, heap = self.group(root)
        sb = (SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + _u(self.leaf_k, 2) + _u(_INTERNAL_K, 2) + _u(0, 4)
              + _u(0, 8) + _u(UNDEF, 8) + _u(len(self.buf), 8) + _u(UNDEF, 8)
              + _u(0, 8) + _u(hdr, 8) + _u(1, 4) + _u(0, 4) + _u(btree, 8) + _u(heap, 8))
        self.buf[0:96] = sb
        return bytes(self.buf)

    def dataset(self, ds: Dataset) -> int:
        a = _as_storable(ds._data)
        raw = _raw(a)
        addr = self.alloc(raw) if raw else UNDEF
        msgs = [_message(0x01, _dataspace_message(a.shape)),
                _message(0x03, _dtype_message(a), flags=1),
                _message(0x05, bytes([2, 2, 2, 0]), flags=1),  # fill value: late allocation, none defined
                _message(0x08, bytes([3, 1]) + _u(addr, 8) + _u(len(raw), 8))]
        msgs += [_attribute_message(k, v) for k, v in ds.attrs.items()]
        return self.header(msgs)

    def group(self, g: Group):
        entries = []
        for name in g.keys():
            c = g._children[name]
            addr = self.group(c)[0] if isinstance(c, Group) else self.dataset(c)
            entries.append((name, addr))
        heap_data, offs = bytearray(8), []
        for name, _ in entries:
            offs.append(len(heap_data))
            nb = name.encode("utf-8") + b"\0"
            heap_data += nb + bytes(_pad8(len(nb)) - len(nb))
        data_addr = self.alloc(bytes(heap_data))
        # free-list head 1 = no free block (offsets of real blocks are multiples of 8)
        heap = self.alloc(b"HEAP" + bytes(4) + _u(len(heap_data), 8) + _u(1, 8) + _u(data_addr, 8))
        cap = 2 * self.leaf_k
        snods = []
        for i in range(0, len(entries), cap):
            chunk = list(zip(entries[i:i + cap], offs[i:i + cap]))
            body = b"".join(_u(off, 8) + _u(addr, 8) + bytes(24) for (_, addr), off in chunk)
            node = b"SNOD" + bytes([1, 0]) + _u(len(chunk), 2) + body
            snods.append((self.alloc(node + bytes(8 + 40 * cap - len(node))), chunk[-1][1]))
        if len(snods) > 2 * _INTERNAL_K:
            raise H5Error("group with more than %d members" % (2 * _INTERNAL_K * cap))
        tree = b"TREE" + bytes([0, 0]) + _u(len(snods), 2) + _u(UNDEF, 8) + _u(UNDEF, 8) + _u(0, 8)
        for addr, last in snods:
            tree += _u(addr, 8) + _u(last, 8)
        size = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8
        btree = self.alloc(tree + bytes(size - len(tree)))
        msgs = [_message(0x11, _u(btree, 8) + _u(heap, 8))]
        msgs += [_attribute_message(k, v) for k, v in g.attrs.items()]
        return self.header(msgs), btree, heap
