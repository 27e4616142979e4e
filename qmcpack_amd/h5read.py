"""Minimal read-only HDF5 reader for orbital files (no HDF5 library in this image).

Reads what the ESHDF orbital files written by the plane-wave converters use: superblock versions
0 / 1 with symbol-table groups (version-1 B-trees of group nodes, local heaps) and superblock 2 / 3
groups with link messages; version-1 and version-2 object headers with continuation blocks;
datasets of fixed-point and IEEE floating-point elements (either byte order) and fixed-length
strings, stored compact, contiguous or in chunks without filters.  Anything else (filters,
variable-length types, dense link storage) raises an error naming it.

Layout facts follow the published HDF5 File Format Specification (version 3.0): superblock
(section II.A), symbol table entry (III.C), B-tree version 1 and group nodes (III.A.1, III.B),
local heap (III.D), object headers (IV.A.1), dataspace, datatype, data layout, link, symbol table
and continuation messages (IV.A.2.b, d, i, g, p, q).
"""
from __future__ import annotations

import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Error(RuntimeError):
    pass


class Dataset:
    def __init__(self, f: "File", addr: int):
        self.f, self.addr = f, addr
        self.shape, self.dtype, self.layout = None, None, None
        for mtype, body in f._messages(addr):
            if mtype == 0x0001:
                self.shape = f._dataspace(body)
            elif mtype == 0x0003:
                self.dtype = f._datatype(body)
            elif mtype == 0x0008:
                self.layout = f._layout(body)
            elif mtype == 0x000B:
                raise H5Error("filtered (compressed) datasets are not supported")
        if self.shape is None or self.dtype is None or self.layout is None:
            raise H5Error("object at %#x is not a dataset" % addr)

    def read(self) -> np.ndarray:
        kind = self.layout[0]
        n = int(np.prod(self.shape)) if self.shape else 1
        nbytes = n * self.dtype.itemsize
        if kind == "compact":
            raw = self.layout[1][:nbytes]
        elif kind == "contiguous":
            addr, size = self.layout[1], self.layout[2]
            raw = self.f.buf[addr:addr + nbytes] if addr != UNDEF else bytes(nbytes)
        else:
            raw = self.f._read_chunked(self.layout, self.shape, self.dtype)
        arr = np.frombuffer(raw, dtype=self.dtype, count=n).reshape(self.shape if self.shape else ())
        if arr.dtype.kind == "S":
            return arr
        return arr.astype(arr.dtype.newbyteorder("="))


class File:
    """An HDF5 file held in memory; f["/a/b"] returns the dataset's array, f.keys("/a") the names."""

    def __init__(self, path: str):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        base = b.find(SIGNATURE)
        if base != 0:
            raise H5Error("not an HDF5 file (signature at offset 0 expected)")
        ver = b[8]
        if ver in (0, 1):
            self.so, self.sl = b[13], b[14]
            p = 24 + (4 if ver == 1 else 0)
            p += 4 * self.so  # base, free space, end of file, driver info addresses
            self.root = self._symbol_entry(p)[1]
        elif ver in (2, 3):
            self.so, self.sl = b[9], b[10]
            p = 12 + 3 * self.so
            self.root = self._off(p)
        else:
            raise H5Error("superblock version %d" % ver)
        if self.so != 8 or self.sl != 8:
            raise H5Error("only 8-byte offsets and lengths are supported")

    # -------------------------------------------------------------- primitives
    def _off(self, p):
        return struct.unpack_from("<Q", self.buf, p)[0]

    def _symbol_entry(self, p):
        name_off, header = struct.unpack_from("<QQ", self.buf, p)
        cache = struct.unpack_from("<I", self.buf, p + 16)[0]
        scratch = struct.unpack_from("<QQ", self.buf, p + 24) if cache == 1 else None
        return name_off, header, cache, scratch

    def _messages(self, addr):
        """(type, body) of every message of the object header at addr, continuations followed."""
        b = self.buf
        out = []
        if b[addr:addr + 4] == b"OHDR":
            ver, flags = b[addr + 4], b[addr + 5]
            if ver != 2:
                raise H5Error("object header version %d" % ver)
            p = addr + 6
            if flags & 0x20:
                p += 16  # times
            if flags & 0x10:
                p += 4  # attribute phase change values
            size_len = 1 << (flags & 3)
            size = int.from_bytes(b[p:p + size_len], "little")
            p += size_len
            blocks = [(p, p + size)]
            track = bool(flags & 0x04)
            while blocks:
                s, e = blocks.pop(0)
                q = s
                while q + 4 <= e:
                    mtype, msize, mflags = b[q], struct.unpack_from("<H", b, q + 1)[0], b[q + 3]
                    q += 4 + (2 if track else 0)
                    body = b[q:q + msize]
                    q += msize
                    if mtype == 0x10:
                        caddr, clen = struct.unpack_from("<QQ", body, 0)
                        blocks.append((caddr + 4, caddr + clen - 4))  # "OCHK" ... checksum
                    elif mtype != 0:
                        out.append((mtype, body))
            return out
        ver = b[addr]
        if ver != 1:
            raise H5Error("object header version %d at %#x" % (ver, addr))
        nmsg = struct.unpack_from("<H", b, addr + 2)[0]
        size = struct.unpack_from("<I", b, addr + 8)[0]
        blocks = [(addr + 16, addr + 16 + size)]
        while blocks and len(out) < nmsg:
            s, e = blocks.pop(0)
            q = s
            while q + 8 <= e and len(out) < nmsg:
                mtype, msize = struct.unpack_from("<HH", b, q)
                body = b[q + 8:q + 8 + msize]
                q += 8 + msize
                if mtype == 0x10:
                    caddr, clen = struct.unpack_from("<QQ", body, 0)
                    blocks.append((caddr, caddr + clen))
                out.append((mtype, body))
        return out

    def _dataspace(self, body):
        ver, ndim, flags = body[0], body[1], body[2]
        p = 8 if ver == 1 else 4
        return tuple(struct.unpack_from("<%dQ" % ndim, body, p)) if ndim else ()

    def _datatype(self, body):
        cls, ver = body[0] & 0x0F, body[0] >> 4
        bits = body[1] | (body[2] << 8) | (body[3] << 16)
        size = struct.unpack_from("<I", body, 4)[0]
        order = ">" if bits & 1 else "<"
        if cls == 0:  # fixed point
            return np.dtype("%s%s%d" % (order, "i" if bits & 8 else "u", size))
        if cls == 1:  # IEEE floating point
            return np.dtype("%sf%d" % (order, size))
        if cls == 3:  # fixed-length string
            return np.dtype("S%d" % size)
        raise H5Error("datatype class %d (version %d)" % (cls, ver))

    def _layout(self, body):
        ver = body[0]
        if ver in (1, 2):
            ndim, cls = body[1], body[2]
            p = 8
            addr = None
            if cls in (1, 2):
                addr = self._off(p)
                p += 8
            dims = struct.unpack_from("<%dI" % ndim, body, p)
            p += 4 * ndim
            if cls == 0:
                n = struct.unpack_from("<I", body, p)[0]
                return ("compact", body[p + 4:p + 4 + n])
            if cls == 1:
                return ("contiguous", addr, int(np.prod(dims)))
            return ("chunked", addr, dims)
        if ver == 3:
            cls = body[1]
            if cls == 0:
                n = struct.unpack_from("<H", body, 2)[0]
                return ("compact", body[4:4 + n])
            if cls == 1:
                addr, size = struct.unpack_from("<QQ", body, 2)
                return ("contiguous", addr, size)
            if cls == 2:
                ndim = body[2]
                addr = self._off(3)
                dims = struct.unpack_from("<%dI" % ndim, body, 11)
                return ("chunked", addr, dims)
        raise H5Error("data layout message version %d" % ver)

    def _read_chunked(self, layout, shape, dtype):
        """Chunks indexed by a version-1 B-tree (layout versions 1-3), no filters."""
        _, btree, cdims = layout
        nd = len(shape)
        out = np.zeros(shape, dtype=dtype)
        chunk = tuple(cdims[:nd])
        csize = int(np.prod(chunk)) * dtype.itemsize

        def walk(addr):
            b = self.buf
            if b[addr:addr + 4] != b"TREE" or b[addr + 4] != 1:
                raise H5Error("chunk B-tree node expected at %#x" % addr)
            level, used = b[addr + 5], struct.unpack_from("<H", b, addr + 6)[0]
            p = addr + 24
            ksize = 8 + 8 * (nd + 1)
            for _ in range(used):
                nbytes, mask = struct.unpack_from("<II", b, p)
                offs = struct.unpack_from("<%dQ" % (nd + 1), b, p + 8)[:nd]
                child = self._off(p + ksize)
                p += ksize + 8
                if level > 0:
                    walk(child)
                    continue
                if mask:
                    raise H5Error("filtered chunks are not supported")
                data = np.frombuffer(b[child:child + csize], dtype=dtype).reshape(chunk)
                sl = tuple(slice(o, min(o + c, s)) for o, c, s in zip(offs, chunk, shape))
                out[sl] = data[tuple(slice(0, x.stop - x.start) for x in sl)]

        walk(btree)
        return out.tobytes()

    # -------------------------------------------------------------- groups
    def _children(self, addr):
        """{name: object header address} of the group at addr."""
        out = {}
        for mtype, body in self._messages(addr):
            if mtype == 0x0011:  # symbol table: v1 B-tree + local heap
                btree, heap = struct.unpack_from("<QQ", body, 0)
                out.update(self._symbol_table(btree, heap))
            elif mtype == 0x0006:  # link message (hard links)
                name, target = self._link(body)
                if target is not None:
                    out[name] = target
            elif mtype == 0x0002:  # link info: dense storage
                fheap = self._off(2 + (8 if body[1] & 1 else 0))
                if fheap != UNDEF:
                    raise H5Error("dense link storage is not supported")
        return out

    def _heap_data(self, heap):
        b = self.buf
        if b[heap:heap + 4] != b"HEAP":
            raise H5Error("local heap expected at %#x" % heap)
        return self._off(heap + 24)

    def _symbol_table(self, btree, heap):
        b = self.buf
        data = self._heap_data(heap)
        out = {}

        def name_at(off):
            e = b.index(b"\x00", data + off)
            return b[data + off:e].decode()

        def walk(addr):
            if b[addr:addr + 4] != b"TREE" or b[addr + 4] != 0:
                raise H5Error("group B-tree node expected at %#x" % addr)
            level, used = b[addr + 5], struct.unpack_from("<H", b, addr + 6)[0]
            p = addr + 24 + 8  # first key
            for _ in range(used):
                child = self._off(p)
                p += 16
                if level > 0:
                    walk(child)
                    continue
                if b[child:child + 4] != b"SNOD":
                    raise H5Error("symbol table node expected at %#x" % child)
                nsym = struct.unpack_from("<H", b, child + 6)[0]
                for i in range(nsym):
                    name_off, header, _, _ = self._symbol_entry(child + 8 + 40 * i)
                    out[name_at(name_off)] = header

        walk(btree)
        return out

    def _link(self, body):
        flags = body[1]
        p = 2
        ltype = 0
        if flags & 0x08:
            ltype = body[p]
            p += 1
        if flags & 0x04:
            p += 8
        if flags & 0x10:
            p += 1
        nlen_size = 1 << (flags & 3)
        nlen = int.from_bytes(body[p:p + nlen_size], "little")
        p += nlen_size
        name = body[p:p + nlen].decode()
        p += nlen
        return name, (self._off(p) if ltype == 0 else None)

    def _resolve(self, path):
        addr = self.root
        for part in [x for x in path.split("/") if x]:
            ch = self._children(addr)
            if part not in ch:
                raise KeyError(path)
            addr = ch[part]
        return addr

    def keys(self, path="/"):
        return sorted(self._children(self._resolve(path)))

    def __contains__(self, path):
        try:
            self._resolve(path)
            return True
        except KeyError:
            return False

    def __getitem__(self, path) -> np.ndarray:
        return Dataset(self, self._resolve(path)).read()
