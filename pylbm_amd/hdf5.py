"""HDF5 + XDMF output of simulation fields (pylbm/hdf5.py:19-325).

``H5File`` keeps pylbm's interface -- ``H5File(mpi_topo, filename, path,
timestep)``, ``set_grid(x, y, z)``, ``add_scalar(name, f, *fargs)``,
``add_vector(name, f, *fargs)``, ``save()`` -- and its file contents: datasets
``x_0 .. x_{dim-1}`` with the coordinates, scalars stored transposed
(dataset shape = global size reversed, pylbm/hdf5.py:190-205), vectors as
(reversed shape + [3]) with component i in the last index (2D vectors leave
component 2 at zero, pylbm/hdf5.py:229-245), and an ``.xdmf`` file describing
the rectilinear grid and the attributes (pylbm/hdf5.py:247-325).

pylbm writes through h5py; here the file is written directly from the HDF5
file format specification (no HDF5 library is needed): superblock version 0,
version 1 object headers, one symbol-table group (version 1 B-tree, local heap
and one symbol node), and contiguous little-endian float64 datasets.  One
process writes the whole domain (``mpi_topo`` is accepted and not used: this
backend runs on one GPU).
"""
import math
import os
import struct

import numpy as np

__all__ = ["H5File", "write_h5"]

_SIGNATURE = b"\x89HDF\r\n\x1a\n"
_UNDEF = 0xFFFFFFFFFFFFFFFF
_INTERNAL_K = 16          # group B-tree internal node K (the library default)
_FREE_END = 1             # end of a local heap free list on disk


def _pad8(n):
    return (n + 7) & ~7


def _message(mtype, data, flags=0):
    data = data + b"\x00" * (_pad8(len(data)) - len(data))
    return struct.pack("<HHB3x", mtype, len(data), flags) + data


def _object_header(messages):
    body = b"".join(messages)
    return struct.pack("<BBHII4x", 1, 0, len(messages), 1, len(body)) + body


def _dataset_messages(shape, addr, nbytes):
    rank = len(shape)
    dims = b"".join(struct.pack("<Q", int(n)) for n in shape)
    # dataspace version 1 with maximum dimensions equal to the dimensions
    space = struct.pack("<BBB5x", 1, rank, 1) + dims + dims
    # IEEE 754 little-endian double: class 1 version 1; mantissa normalisation
    # "implied msb" (2), sign bit 63; offset 0, precision 64, exponent at 52 of
    # 11 bits, mantissa at 0 of 52 bits, bias 1023
    dtype = struct.pack("<BBBBI", 0x11, 0x20, 63, 0, 8) + struct.pack("<HHBBBBI", 0, 64, 52, 11, 0, 52, 1023)
    # fill value version 2: allocation late, write if set, default value
    fill = struct.pack("<BBBBI", 2, 2, 2, 1, 0)
    # data layout version 3, contiguous
    layout = struct.pack("<BBQQ", 3, 1, addr, nbytes)
    return [_message(0x0001, space), _message(0x0003, dtype, flags=1), _message(0x0005, fill, flags=1),
            _message(0x0008, layout)]


def write_h5(path, datasets):
    """Write ``datasets`` ({name: float64 array}) as the root group of a new HDF5 file."""
    names = sorted(datasets)                      # symbol nodes hold names in strcmp order
    if any(not n or "/" in n or "\x00" in n for n in names):
        raise ValueError("dataset names must be non-empty and contain no '/'")
    arrays = {n: np.ascontiguousarray(np.asarray(datasets[n], dtype="<f8")) for n in names}
    leaf_k = max(4, math.ceil(len(names) / 2))

    # local heap data segment: "" at 0, the names, then one free block
    heap = bytearray(b"\x00" * 8)
    name_off = {}
    for n in names:
        name_off[n] = len(heap)
        raw = n.encode() + b"\x00"
        heap += raw + b"\x00" * (_pad8(len(raw)) - len(raw))
    free_off = len(heap)
    heap += struct.pack("<QQ", _FREE_END, 16)

    # addresses
    sb_size = 96
    root_msgs = [None]                            # symbol table message, filled below
    root_size = 16 + 8 + 16
    heap_hdr = sb_size + root_size
    heap_data = heap_hdr + 32
    tree = heap_data + len(heap)
    tree_size = 24 + (2 * _INTERNAL_K + 1) * 8 + 2 * _INTERNAL_K * 8
    snod = tree + tree_size
    snod_size = 8 + 2 * leaf_k * 40
    pos = snod + snod_size
    headers = {}
    for n in names:
        shape = arrays[n].shape
        hdr_len = 16 + sum(len(m) for m in _dataset_messages(shape, 0, 0))
        headers[n] = pos
        pos += _pad8(hdr_len)
    data_addr = {}
    for n in names:
        data_addr[n] = pos
        pos += _pad8(arrays[n].nbytes)
    eof = pos

    out = bytearray(eof)
    # superblock version 0, 8-byte offsets and lengths
    sb = _SIGNATURE + struct.pack("<BBBBBBBB", 0, 0, 0, 0, 0, 8, 8, 0)
    sb += struct.pack("<HHI", leaf_k, _INTERNAL_K, 0)
    sb += struct.pack("<QQQQ", 0, _UNDEF, eof, _UNDEF)
    # root group symbol table entry: name offset 0, header, cache type 1 (B-tree and heap)
    sb += struct.pack("<QQII", 0, sb_size, 1, 0) + struct.pack("<QQ", tree, heap_hdr)
    assert len(sb) == sb_size
    out[0:sb_size] = sb
    root_msgs[0] = _message(0x0011, struct.pack("<QQ", tree, heap_hdr))
    out[sb_size:sb_size + root_size] = _object_header(root_msgs)
    out[heap_hdr:heap_data] = b"HEAP" + struct.pack("<B3xQQQ", 0, len(heap), free_off, heap_data)
    out[heap_data:heap_data + len(heap)] = heap
    # one leaf B-tree node (type 0: group) with one child, the symbol node
    last = name_off[names[-1]] if names else 0
    node = b"TREE" + struct.pack("<BBHQQ", 0, 0, 1, _UNDEF, _UNDEF) + struct.pack("<QQQ", 0, snod, last)
    out[tree:tree + len(node)] = node
    # symbol node: one entry per dataset, cache type 0
    sn = b"SNOD" + struct.pack("<BBH", 1, 0, len(names))
    for n in names:
        sn += struct.pack("<QQII", name_off[n], headers[n], 0, 0) + b"\x00" * 16
    out[snod:snod + len(sn)] = sn
    for n in names:
        a = arrays[n]
        hdr = _object_header(_dataset_messages(a.shape, data_addr[n], a.nbytes))
        out[headers[n]:headers[n] + len(hdr)] = hdr
        out[data_addr[n]:data_addr[n] + a.nbytes] = a.tobytes()
    with open(path, "wb") as fh:
        fh.write(bytes(out))


class H5File:
    """HDF5 + XDMF output of fields on the simulation grid (pylbm.H5File)."""

    def __init__(self, mpi_topo, filename, path="", timestep=None, init_xdmf=False):
        prefix = "_{}".format(timestep) if timestep is not None else ""
        self.path = path
        name, _ = os.path.splitext(filename)
        self.filename = name + prefix
        self.h5filename = name + prefix + ".h5"
        self.mpi_topo = mpi_topo
        self.origin = self.dx = self.dim = self.n = self.region = None
        self.global_size = None
        self.xdmf_file = None
        self.scalars = {}
        self.vectors = {}
        self._data = {}
        self._init_grid = True
        self._init_xdmf = init_xdmf
        if path and not os.path.exists(path):
            os.mkdir(path)

    def set_grid(self, x, y=None, z=None):
        """The grid: one coordinate array per axis (datasets x_0, x_1, x_2)."""
        coords = [np.asarray(c, dtype=np.float64) for c in (x, y, z) if c is not None]
        self.origin = [float(c[0]) for c in coords]
        self.dx = [float(c[1] - c[0]) if c.size > 1 else 0.0 for c in coords]
        self.dim = len(coords)
        self.n = [int(c.size) for c in coords]
        self.global_size = list(self.n)
        self.region = [[0, n] for n in self.n]
        for i, c in enumerate(coords):
            self._data["x_{}".format(i)] = c.copy()

    def _check(self, name, data):
        if self.global_size is None:
            raise RuntimeError("set_grid must be called before adding fields")
        if name in self._data:
            raise ValueError("dataset %r exists" % name)
        data = np.asarray(data, dtype=np.float64)
        if list(data.shape) != self.global_size:
            raise ValueError("field %r has shape %s, the grid is %s" % (name, data.shape, self.global_size))
        return data

    def add_scalar(self, name, f, *fargs):
        """A scalar field (an array on the grid, or f(*fargs) returning one), stored transposed."""
        data = self._check(name, f if isinstance(f, np.ndarray) else f(*fargs))
        self._data[name] = np.ascontiguousarray(data.T)
        self.scalars[name] = self.h5filename + ":/" + name

    def add_vector(self, name, f, *fargs):
        """A vector field (a list of component arrays, or f(*fargs) returning one), stored as
        (reversed grid shape, 3) with component i at [..., i]."""
        datas = f if isinstance(f, list) else f(*fargs)
        datas = [self._check(name, d) for d in datas]
        out = np.zeros(self.global_size[::-1] + [3])
        for i, d in enumerate(datas):
            out[..., i] = d.T
        self._data[name] = out
        self.vectors[name] = self.h5filename + ":/" + name

    def save(self):
        """Write the .h5 file and its .xdmf description."""
        folder = self.path if self.path else "."
        write_h5(os.path.join(folder, self.h5filename), self._data)
        size = " ".join(map(str, self.global_size))
        rsize = " ".join(map(str, self.global_size[::-1]))
        lines = ['<?xml version="1.0" ?>', '<!DOCTYPE Xdmf SYSTEM "Xdmf.dtd" []>', "<Xdmf>", " <Domain>",
                 '  <Grid Name="Structured Grid" GridType="Uniform">',
                 '   <Topology TopologyType="%dDRectMesh" NumberOfElements="%s"/>' % (self.dim, size),
                 '   <Geometry GeometryType="%s">' % ("VXVY" if self.dim == 2 else "VXVYVZ")]
        for i in range(self.dim):
            lines += ['    <DataItem Format="HDF" Dimensions="%d">' % self.global_size[i],
                      "     %s.h5:/x_%d" % (self.filename, i), "    </DataItem>"]
        lines.append("   </Geometry>")
        for k, v in self.scalars.items():
            lines += ['   <Attribute Name="%s" AttributeType="Scalar" Center="Node">' % k,
                      '    <DataItem Format="HDF" Dimensions="%s">' % rsize, "     %s" % v,
                      "    </DataItem>", "   </Attribute>"]
        for k, v in self.vectors.items():
            lines += ['   <Attribute Name="%s" AttributeType="Vector" Center="Node">' % k,
                      '    <DataItem Format="HDF" Dimensions="%s %d">' % (rsize, self.dim), "     %s" % v,
                      "    </DataItem>", "   </Attribute>"]
        lines += ["  </Grid>", " </Domain>", "</Xdmf>"]
        self.xdmf_file = os.path.join(folder, self.filename + ".xdmf")
        with open(self.xdmf_file, "w") as fh:
            fh.write("\n".join(lines) + "\n")
