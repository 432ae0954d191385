"""A small HDF5 reader and writer (numpy only), for the phantom, beam-profile and
scan files of the reference's HDF5 formats (src/utils/HDFConvenience.cpp:45-196,
written by its phantom scripts with h5py, e.g. phantom_scripts/cyst_phantom_2d.py:35-40).
Neither h5py nor the HDF5 library is a dependency of this package.

Written from the HDF5 File Format Specification (version 3.0), covering what
such files use:
  read:  superblock versions 0-3; object headers version 1 and 2 with
         continuation blocks; groups stored as symbol tables (version-1 B-tree,
         local heap, symbol-table nodes) or as compact link messages; datasets
         with a scalar or simple dataspace, fixed-point or IEEE floating-point
         elements of either byte order, compact, contiguous or (unfiltered,
         version-3 layout) chunked storage;
  write: superblock version 0, version-1 object headers, one root group as a
         symbol table, contiguous little-endian numeric datasets -- the layout
         h5py writes by default.
Anything else (filters such as gzip, dense link storage, variable-length or
compound types, attributes' values) raises H5Error naming what is missing.
"""
from __future__ import annotations

import struct
from typing import Dict

import numpy as np

__all__ = ["H5Error", "read", "write"]

_SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(ValueError):
    """The file is not HDF5, or uses a feature this reader does not cover."""


# --------------------------------------------------------------------------- reader
class _Reader:
    def __init__(self, buf: bytes):
        self.b = buf
        if len(buf) < 48:
            raise H5Error("file too short for an HDF5 superblock")
        # the superblock may sit at 0, 512, 1024, ... (a user block before it)
        base = 0
        while buf[base:base + 8] != _SIGNATURE:
            base = 512 if base == 0 else base * 2
            if base + 8 > len(buf):
                raise H5Error("no HDF5 signature")
        ver = buf[base + 8]
        if ver in (0, 1):
            self.so, self.sl = buf[base + 13], buf[base + 14]
            p = base + 24 + (4 if ver == 1 else 0)
            self.base_addr = self.uint(p, self.so)
            p += 4 * self.so  # base, free-space, end-of-file, driver addresses
            # root group symbol-table entry: name offset, object header address, ...
            self.root = self.uint(p + self.so, self.so)
        elif ver in (2, 3):
            self.so, self.sl = buf[base + 9], buf[base + 10]
            self.base_addr = self.uint(base + 12, self.so)
            self.root = self.uint(base + 12 + 3 * self.so, self.so)
        else:
            raise H5Error(f"superblock version {ver} not supported")
        if self.so not in (2, 4, 8) or self.sl not in (2, 4, 8):
            raise H5Error("unsupported size of offsets / lengths")

    # ---- primitives
    def uint(self, p: int, n: int) -> int:
        return int.from_bytes(self.b[p:p + n], "little")

    def at(self, a: int) -> int:
        return self.base_addr + a

    # ---- object headers: list of (type, data bytes)
    def messages(self, a: int):
        p = self.at(a)
        out = []
        if self.b[p:p + 4] == b"OHDR":
            ver, flags = self.b[p + 4], self.b[p + 5]
            if ver != 2:
                raise H5Error(f"object header version {ver} not supported")
            q = p + 6
            if flags & 0x20:
                q += 16
            if flags & 0x10:
                q += 4
            w = 1 << (flags & 3)
            size = self.uint(q, w)
            q += w
            self._v2_chunk(q, q + size, flags, out)
        else:
            if self.b[p] != 1:
                raise H5Error(f"object header version {self.b[p]} not supported")
            size = self.uint(p + 8, 4)
            self._v1_chunk(p + 16, p + 16 + size, out)
        # continuation blocks (message 0x10) append their messages; a block met
        # twice (a damaged file) would repeat forever
        k = 0
        blocks = set()
        while k < len(out):
            t, d = out[k]
            if t == 0x10:
                ca = int.from_bytes(d[:self.so], "little")
                cl = int.from_bytes(d[self.so:self.so + self.sl], "little")
                if ca in blocks or not 0 <= self.at(ca) < len(self.b):
                    raise H5Error("damaged object header continuation")
                blocks.add(ca)
                c = self.at(ca)
                if self.b[c:c + 4] == b"OCHK":
                    self._v2_chunk(c + 4, c + cl - 4, self._v2_flags(a), out)
                else:
                    self._v1_chunk(c, c + cl, out)
            k += 1
        return out

    def _v2_flags(self, a: int) -> int:
        return self.b[self.at(a) + 5]

    def _v1_chunk(self, p: int, end: int, out):
        while p + 8 <= end:
            t, size = self.uint(p, 2), self.uint(p + 2, 2)
            out.append((t, self.b[p + 8:p + 8 + size]))
            p += 8 + size

    def _v2_chunk(self, p: int, end: int, flags: int, out):
        hdr = 4 + (2 if flags & 0x04 else 0)
        while p + hdr <= end:
            t, size = self.b[p], self.uint(p + 1, 2)
            out.append((t, self.b[p + hdr:p + hdr + size]))
            p += hdr + size

    # ---- groups: {name: object header address}
    def links(self, a: int) -> Dict[str, int]:
        res: Dict[str, int] = {}
        for t, d in self.messages(a):
            if t == 0x11:  # symbol table: v1 B-tree + local heap
                btree = int.from_bytes(d[:self.so], "little")
                heap = int.from_bytes(d[self.so:2 * self.so], "little")
                self._walk_group_btree(btree, self._heap_data(heap), res)
            elif t == 0x06:  # link message (compact storage)
                name, target = self._link(d)
                if target is not None:
                    res[name] = target
            elif t == 0x02:  # link info
                fheap = int.from_bytes(d[2 + (8 if d[1] & 1 else 0):][:self.so], "little")
                if fheap != (1 << (8 * self.so)) - 1:
                    raise H5Error("groups in dense link storage (fractal heap) are not supported")
        return res

    def _heap_data(self, a: int) -> bytes:
        p = self.at(a)
        if self.b[p:p + 4] != b"HEAP":
            raise H5Error("local heap signature missing")
        size = self.uint(p + 8, self.sl)
        data = self.uint(p + 8 + 2 * self.sl, self.so)
        return self.b[self.at(data):self.at(data) + size]

    def _walk_group_btree(self, a: int, heap: bytes, res: Dict[str, int]):
        p = self.at(a)
        if self.b[p:p + 4] != b"TREE" or self.b[p + 4] != 0:
            raise H5Error("group B-tree node expected")
        level, used = self.b[p + 5], self.uint(p + 6, 2)
        q = p + 8 + 2 * self.so + self.sl  # past the siblings and key 0
        for _ in range(used):
            child = self.uint(q, self.so)
            q += self.so + self.sl
            if level > 0:
                self._walk_group_btree(child, heap, res)
            else:
                self._symbol_node(child, heap, res)

    def _symbol_node(self, a: int, heap: bytes, res: Dict[str, int]):
        p = self.at(a)
        if self.b[p:p + 4] != b"SNOD":
            raise H5Error("symbol table node signature missing")
        n = self.uint(p + 6, 2)
        e = p + 8
        for _ in range(n):
            name_off = self.uint(e, self.so)
            obj = self.uint(e + self.so, self.so)
            end = heap.index(b"\0", name_off)
            res[heap[name_off:end].decode("utf-8")] = obj
            e += 2 * self.so + 24

    def _link(self, d: bytes):
        flags = d[1]
        q = 2
        ltype = 0
        if flags & 0x08:
            ltype = d[q]
            q += 1
        if flags & 0x04:
            q += 8
        if flags & 0x10:
            q += 1
        w = 1 << (flags & 3)
        n = int.from_bytes(d[q:q + w], "little")
        q += w
        name = d[q:q + n].decode("utf-8")
        q += n
        if ltype != 0:
            return name, None  # soft or external link: not followed
        return name, int.from_bytes(d[q:q + self.so], "little")

    # ---- datasets
    def is_group(self, msgs) -> bool:
        return any(t in (0x11, 0x06, 0x02) for t, _ in msgs)

    def dataset(self, msgs) -> np.ndarray:
        shape = dtype = layout = None
        for t, d in msgs:
            if t == 0x01:
                shape = self._dataspace(d)
            elif t == 0x03:
                dtype = self._datatype(d)
            elif t == 0x08:
                layout = d
            elif t == 0x0B:
                raise H5Error("filtered (compressed) datasets are not supported")
        if shape is None or dtype is None or layout is None:
            raise H5Error("dataset without dataspace, datatype or layout")
        count = int(np.prod(shape, dtype=np.int64)) if shape else 1
        raw = self._raw(layout, count * dtype.itemsize, shape, dtype.itemsize)
        arr = np.frombuffer(raw, dtype=dtype, count=count).reshape(shape)
        return arr.astype(dtype.newbyteorder("="))

    def _dataspace(self, d: bytes):
        ver, rank, flags = d[0], d[1], d[2]
        if ver == 1:
            q = 8
        elif ver == 2:
            if d[3] == 2:
                raise H5Error("null dataspace")
            q = 4
        else:
            raise H5Error(f"dataspace version {ver} not supported")
        return tuple(int.from_bytes(d[q + i * self.sl:q + (i + 1) * self.sl], "little") for i in range(rank))

    @staticmethod
    def _datatype(d: bytes) -> np.dtype:
        cls, bits, size = d[0] & 0x0F, d[1], int.from_bytes(d[4:8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 0 and size in (1, 2, 4, 8):
            return np.dtype(f"{order}{'i' if bits & 0x08 else 'u'}{size}")
        if cls == 1 and size in (2, 4, 8):
            return np.dtype(f"{order}f{size}")
        raise H5Error(f"datatype class {cls} of {size} bytes not supported")

    def _raw(self, d: bytes, nbytes: int, shape, elem: int) -> bytes:
        ver = d[0]
        if ver == 3:
            cls = d[1]
            if cls == 0:  # compact
                size = int.from_bytes(d[2:4], "little")
                return d[4:4 + size][:nbytes]
            if cls == 1:  # contiguous
                a = int.from_bytes(d[2:2 + self.so], "little")
                if a == (1 << (8 * self.so)) - 1:
                    return bytes(nbytes)  # never written: all zero (default fill)
                return self.b[self.at(a):self.at(a) + nbytes]
            if cls == 2:  # chunked, version-1 B-tree index
                rank = d[2] - 1
                a = int.from_bytes(d[3:3 + self.so], "little")
                q = 3 + self.so
                cdims = [int.from_bytes(d[q + 4 * i:q + 4 * i + 4], "little") for i in range(rank)]
                return self._chunked(a, shape, cdims, elem)
            raise H5Error(f"layout class {cls} not supported")
        if ver in (1, 2):
            rank, cls = d[1], d[2]
            q = 8
            if cls == 1:
                a = int.from_bytes(d[q:q + self.so], "little")
                return self.b[self.at(a):self.at(a) + nbytes]
            if cls == 0:
                q += 4 * rank
                size = int.from_bytes(d[q:q + 4], "little")
                return d[q + 4:q + 4 + size][:nbytes]
            raise H5Error("chunked storage in an old layout message is not supported")
        raise H5Error(f"layout message version {ver} not supported")

    def _chunked(self, a: int, shape, cdims, elem: int) -> bytes:
        out = np.zeros(int(np.prod(shape, dtype=np.int64)) * elem, np.uint8).reshape(*shape, elem) if shape else \
            np.zeros((elem,), np.uint8)
        rank = len(shape)

        def walk(node):
            p = self.at(node)
            if self.b[p:p + 4] != b"TREE" or self.b[p + 4] != 1:
                raise H5Error("chunk B-tree node expected")
            level, used = self.b[p + 5], self.uint(p + 6, 2)
            q = p + 8 + 2 * self.so
            ksize = 8 + 8 * (rank + 1)
            for _ in range(used):
                size = self.uint(q, 4)
                if self.uint(q + 4, 4):
                    raise H5Error("filtered chunks are not supported")
                offs = [self.uint(q + 8 + 8 * i, 8) for i in range(rank)]
                child = self.uint(q + ksize, self.so)
                if level > 0:
                    walk(child)
                else:
                    c = self.at(child)
                    chunk = np.frombuffer(self.b[c:c + size], np.uint8).reshape(*cdims, elem)
                    sl = tuple(slice(o, min(o + n, s)) for o, n, s in zip(offs, cdims, shape))
                    out[sl] = chunk[tuple(slice(0, x.stop - x.start) for x in sl)]
                q += ksize + self.so

        walk(a)
        return out.tobytes()

    # ---- whole file
    def read_all(self) -> Dict[str, np.ndarray]:
        res: Dict[str, np.ndarray] = {}
        seen = set()

        def visit(a: int, prefix: str):
            if a in seen:
                return
            seen.add(a)
            for name, obj in sorted(self.links(a).items()):
                msgs = self.messages(obj)
                if self.is_group(msgs):
                    visit(obj, prefix + name + "/")
                else:
                    res[prefix + name] = self.dataset(msgs)

        visit(self.root, "")
        return res


def read(path: str) -> Dict[str, np.ndarray]:
    """Every dataset of an HDF5 file, keyed by its path ("data", "group/x", ...).
    Scalar datasets come back as 0-d arrays."""
    with open(path, "rb") as f:
        return _Reader(f.read()).read_all()


# --------------------------------------------------------------------------- writer
def _pad8(n: int) -> int:
    return (n + 7) & ~7


def _msg(mtype: int, data: bytes, flags: int = 0) -> bytes:
    data = data + bytes(_pad8(len(data)) - len(data))
    return struct.pack("<HHB3x", mtype, len(data), flags) + data


def _datatype_msg(dt: np.dtype) -> bytes:
    size = dt.itemsize
    if dt.kind == "f":
        # class 1, version 1; little-endian IEEE; sign bit position, exponent and
        # mantissa fields and bias as the IEEE formats define them
        spec = {2: (15, 10, 5, 0, 10, 15), 4: (31, 23, 8, 0, 23, 127), 8: (63, 52, 11, 0, 52, 1023)}
        sign, eloc, esize, mloc, msize, bias = spec[size]
        return struct.pack("<BBBBIHHBBBBI", 0x11, 0x20, sign, 0, size, 0, 8 * size, eloc, esize, mloc, msize, bias)
    if dt.kind in "iu":
        return struct.pack("<BBBBIHH", 0x10, 0x08 if dt.kind == "i" else 0, 0, 0, size, 0, 8 * size)
    raise H5Error(f"cannot write elements of type {dt}")


def write(path: str, datasets: Dict[str, np.ndarray]) -> None:
    """Write numeric arrays (and scalars) as root-level datasets of a new HDF5 file:
    superblock version 0, the root group as a symbol table, contiguous
    little-endian storage -- the layout h5py writes by default."""
    items = []
    for name, value in datasets.items():
        if not name or "/" in name or "\0" in name:
            raise H5Error(f"invalid dataset name {name!r}")
        arr = np.asarray(value)
        if arr.dtype == np.bool_:
            arr = arr.astype(np.uint8)
        if arr.dtype.kind not in "fiu":
            raise H5Error(f"dataset {name!r}: cannot write elements of type {arr.dtype}")
        items.append((name, np.asarray(arr, dtype=arr.dtype.newbyteorder("<"), order="C")))  # keeps 0-d
    items.sort(key=lambda kv: kv[0].encode("utf-8"))
    K_LEAF, K_NODE = 4, 16
    per_node = 2 * K_LEAF
    if len(items) > per_node * 2 * K_NODE:
        raise H5Error("too many datasets for a single-level group")
    nodes = [items[i:i + per_node] for i in range(0, len(items), per_node)] or [[]]

    # local heap: "" at 0, then every name, null-terminated and padded to 8
    heap = bytearray(8)
    name_off = {}
    for name, _ in items:
        name_off[name] = len(heap)
        raw = name.encode("utf-8") + b"\0"
        heap += raw + bytes(_pad8(len(raw)) - len(raw))

    SB, ROOT = 0, 96
    BTREE = ROOT + 40
    BTREE_SIZE = 24 + 2 * K_NODE * 8 + (2 * K_NODE + 1) * 8
    HEAP = BTREE + BTREE_SIZE
    HEAP_DATA = HEAP + 32
    SNOD0 = HEAP_DATA + len(heap)
    SNOD_SIZE = 8 + per_node * 40
    p = SNOD0 + len(nodes) * SNOD_SIZE
    headers = {}
    for name, arr in items:
        dspace = struct.pack("<BBB5x", 1, arr.ndim, 0) + b"".join(struct.pack("<Q", n) for n in arr.shape)
        body = (_msg(0x01, dspace) + _msg(0x03, _datatype_msg(arr.dtype), 1) +
                _msg(0x05, bytes([2, 2, 2, 1, 0, 0, 0, 0]), 1))
        layout_len = len(_msg(0x08, bytes(18)))
        headers[name] = (p, body)
        p += 16 + len(body) + layout_len
    data_at = {}
    for name, arr in items:
        p = _pad8(p)
        data_at[name] = p
        p += arr.nbytes
    eof = p

    out = bytearray(eof)
    # superblock version 0 and the root group's symbol-table entry
    out[0:56] = (_SIGNATURE + bytes([0, 0, 0, 0, 0, 8, 8, 0]) + struct.pack("<HHI", K_LEAF, K_NODE, 0) +
                 struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF))
    out[56:96] = struct.pack("<QQII", 0, ROOT, 1, 0) + struct.pack("<QQ", BTREE, HEAP)
    # root object header: one symbol-table message
    msg = _msg(0x11, struct.pack("<QQ", BTREE, HEAP))
    out[ROOT:ROOT + 16 + len(msg)] = struct.pack("<BBHII4x", 1, 0, 1, 1, len(msg)) + msg
    # group B-tree: one leaf-level node over the symbol-table nodes; key i is the
    # heap offset of the last name in child i - 1 (key 0: the empty name)
    bt = bytearray(b"TREE" + struct.pack("<BBH", 0, 0, len(nodes)) + struct.pack("<QQ", _UNDEF, _UNDEF))
    bt += struct.pack("<Q", 0)
    for k, node in enumerate(nodes):
        bt += struct.pack("<QQ", SNOD0 + k * SNOD_SIZE, name_off[node[-1][0]] if node else 0)
    out[BTREE:BTREE + len(bt)] = bt
    # local heap (no free block)
    out[HEAP:HEAP_DATA] = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), _UNDEF, HEAP_DATA)
    out[HEAP_DATA:HEAP_DATA + len(heap)] = heap
    # symbol-table nodes
    for k, node in enumerate(nodes):
        s = SNOD0 + k * SNOD_SIZE
        sn = bytearray(b"SNOD" + struct.pack("<BBH", 1, 0, len(node)))
        for name, _ in node:
            sn += struct.pack("<QQII16x", name_off[name], headers[name][0], 0, 0)
        out[s:s + len(sn)] = sn
    # dataset object headers (dataspace, datatype, fill value, contiguous layout)
    for name, arr in items:
        at, body = headers[name]
        layout = _msg(0x08, struct.pack("<BBQQ", 3, 1, data_at[name], arr.nbytes), 1)
        hdr = struct.pack("<BBHII4x", 1, 0, 4, 1, len(body) + len(layout)) + body + layout
        out[at:at + len(hdr)] = hdr
        out[data_at[name]:data_at[name] + arr.nbytes] = arr.tobytes()
    with open(path, "wb") as f:
        f.write(bytes(out))
