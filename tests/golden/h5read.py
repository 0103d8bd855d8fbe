"""Minimal HDF5 reader, written from the published HDF5 file format
specification (format version 0 superblock, version 1 object headers,
symbol-table groups: version 1 B-tree + local heap + symbol nodes, contiguous
or compact numeric datasets).

It reads what pylbm's H5File writes with h5py's default format
(pylbm/hdf5.py:44, 118, 204, 240: one file, ``create_dataset`` of float64
arrays, no chunking, no compression).  h5py is not installed in this
environment, so the fixtures of the reference's tests (tests/reference/*.h5)
are converted with this reader.  It is test infrastructure and raises
``NotImplementedError`` for anything outside that subset.
"""
import struct

import numpy as np

SIGNATURE = b"\x89HDF\r\n\x1a\n"
UNDEF = 0xFFFFFFFFFFFFFFFF


class H5Reader:
    def __init__(self, path):
        with open(path, "rb") as fh:
            self.buf = fh.read()
        b = self.buf
        if b[:8] != SIGNATURE:
            raise ValueError("%s: not an HDF5 file" % path)
        version = b[8]
        if version not in (0, 1):
            raise NotImplementedError("superblock version %d" % version)
        self.so, self.sl = b[13], b[14]  # sizes of offsets and lengths
        if (self.so, self.sl) != (8, 8):
            raise NotImplementedError("offset/length sizes %d/%d" % (self.so, self.sl))
        pos = 24 if version == 0 else 28
        self.base = self._u(pos, 8)
        pos += 4 * 8  # base, free-space, end-of-file, driver-info addresses
        # root group symbol table entry
        self.root_header = self._u(pos + 8, 8)

    # ---------------------------------------------------------------- raw
    def _u(self, pos, n):
        return int.from_bytes(self.buf[pos:pos + n], "little")

    def _cstr(self, pos):
        end = self.buf.index(b"\x00", pos)
        return self.buf[pos:end].decode()

    # ------------------------------------------------------------ headers
    def messages(self, addr):
        """(type, flags, data bytes) of every message of a version 1 object header."""
        b = self.buf
        if b[addr] != 1:
            raise NotImplementedError("object header version %d at %#x" % (b[addr], addr))
        nmsg = self._u(addr + 2, 2)
        size = self._u(addr + 8, 4)
        blocks = [(addr + 16, size)]
        out = []
        while blocks and len(out) < nmsg:
            start, length = blocks.pop(0)
            pos, end = start, start + length
            while pos + 8 <= end and len(out) < nmsg:
                mtype = self._u(pos, 2)
                msize = self._u(pos + 2, 2)
                flags = b[pos + 4]
                data = b[pos + 8:pos + 8 + msize]
                out.append((mtype, flags, data))
                if mtype == 0x10:  # continuation: another block of messages
                    blocks.append((int.from_bytes(data[0:8], "little"), int.from_bytes(data[8:16], "little")))
                pos += 8 + msize
        return out

    # -------------------------------------------------------------- groups
    def _heap_data(self, heap_addr):
        if self.buf[heap_addr:heap_addr + 4] != b"HEAP":
            raise ValueError("bad local heap at %#x" % heap_addr)
        return self._u(heap_addr + 24, 8)

    def _btree_snods(self, node):
        b = self.buf
        if b[node:node + 4] != b"TREE":
            raise ValueError("bad B-tree node at %#x" % node)
        ntype, level, used = b[node + 4], b[node + 5], self._u(node + 6, 2)
        if ntype != 0:
            raise NotImplementedError("B-tree node type %d" % ntype)
        pos = node + 8 + 16
        children = []
        for _ in range(used):
            pos += 8  # key (heap offset)
            children.append(self._u(pos, 8))
            pos += 8
        if level == 0:
            return children
        out = []
        for c in children:
            out.extend(self._btree_snods(c))
        return out

    def group_entries(self, header_addr):
        """[(name, object header address)] of a symbol-table group."""
        stab = [d for t, _, d in self.messages(header_addr) if t == 0x11]
        if not stab:
            raise NotImplementedError("group at %#x has no symbol table message" % header_addr)
        btree = int.from_bytes(stab[0][0:8], "little")
        heap = self._heap_data(int.from_bytes(stab[0][8:16], "little"))
        out = []
        for snod in self._btree_snods(btree):
            if self.buf[snod:snod + 4] != b"SNOD":
                raise ValueError("bad symbol node at %#x" % snod)
            nsym = self._u(snod + 6, 2)
            pos = snod + 8
            for _ in range(nsym):
                name = self._cstr(heap + self._u(pos, 8))
                out.append((name, self._u(pos + 8, 8)))
                pos += 40
        return out

    # ------------------------------------------------------------ datasets
    @staticmethod
    def _dataspace(d):
        version, rank, flags = d[0], d[1], d[2]
        pos = 8 if version == 1 else 4
        return tuple(int.from_bytes(d[pos + 8 * i:pos + 8 * i + 8], "little") for i in range(rank))

    @staticmethod
    def _datatype(d):
        cls, bits, size = d[0] & 0x0F, d[1], int.from_bytes(d[4:8], "little")
        order = ">" if bits & 1 else "<"
        if cls == 1:
            return np.dtype("%sf%d" % (order, size))
        if cls == 0:
            return np.dtype("%s%s%d" % (order, "i" if bits & 0x08 else "u", size))
        raise NotImplementedError("datatype class %d" % cls)

    def dataset(self, header_addr):
        shape = dtype = None
        raw = None
        for t, _, d in self.messages(header_addr):
            if t == 0x01:
                shape = self._dataspace(d)
            elif t == 0x03:
                dtype = self._datatype(d)
            elif t == 0x08:
                version = d[0]
                if version == 3:
                    cls = d[1]
                    if cls == 1:
                        addr, size = int.from_bytes(d[2:10], "little"), int.from_bytes(d[10:18], "little")
                        raw = self.buf[addr:addr + size] if addr != UNDEF else None
                    elif cls == 0:
                        size = int.from_bytes(d[2:4], "little")
                        raw = d[4:4 + size]
                    else:
                        raise NotImplementedError("chunked dataset layout")
                else:
                    raise NotImplementedError("data layout message version %d" % version)
            elif t in (0x0B, 0x0C) and t == 0x0B:
                raise NotImplementedError("filtered dataset")
        if shape is None or dtype is None:
            return None
        n = int(np.prod(shape)) if shape else 1
        if raw is None:
            return np.zeros(shape, dtype=dtype.newbyteorder("="))
        return np.frombuffer(raw, dtype=dtype, count=n).reshape(shape).astype(dtype.newbyteorder("="))

    def read_all(self, header_addr=None, prefix=""):
        """{path: ndarray} of every dataset below a group (default: the root)."""
        out = {}
        for name, addr in self.group_entries(self.root_header if header_addr is None else header_addr):
            types = {t for t, _, _ in self.messages(addr)}
            if 0x11 in types:
                out.update(self.read_all(addr, prefix + name + "/"))
            else:
                arr = self.dataset(addr)
                if arr is not None:
                    out[prefix + name] = arr
        return out


def read_h5(path):
    """{dataset path: ndarray} of an HDF5 file."""
    return H5Reader(path).read_all()
