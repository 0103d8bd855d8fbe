"""HDF5 / XDMF output (pylbm_amd.hdf5, pylbm/hdf5.py:19-325).

The writer follows the HDF5 file format specification directly.  Its files are
read back (1) by the from-spec reader that converts pylbm's own fixtures
(tests/golden/h5read.py), and (2) by the HDF5 library itself through its C API
where the machine has one (the libhdf5 shipped with this image's conda), so a
structural mistake the reader shares with the writer still fails (2).  The
layout of the datasets follows pylbm: coordinates x_i, scalars transposed,
vectors (reversed shape, 3)."""
import ctypes
import os
import xml.etree.ElementTree as ET

import numpy as np
import pytest

from pylbm_amd.hdf5 import H5File, write_h5

HERE = os.path.dirname(os.path.abspath(__file__))


def _reader():
    import sys
    sys.path.insert(0, os.path.join(HERE, "golden"))
    from h5read import H5Reader
    return H5Reader


_LIBS = ["/opt/conda/lib/libhdf5.so", "libhdf5.so", "libhdf5_serial.so"]


def _libhdf5():
    for name in _LIBS:
        try:
            lib = ctypes.CDLL(name)
        except OSError:
            continue
        lib.H5open()
        for fn in ("H5Fopen", "H5Dopen2", "H5Dget_space", "H5Gopen2"):
            getattr(lib, fn).restype = ctypes.c_int64
        lib.H5Fopen.argtypes = [ctypes.c_char_p, ctypes.c_uint, ctypes.c_int64]
        lib.H5Dopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        lib.H5Gopen2.argtypes = [ctypes.c_int64, ctypes.c_char_p, ctypes.c_int64]
        lib.H5Dget_space.argtypes = [ctypes.c_int64]
        lib.H5Sget_simple_extent_ndims.argtypes = [ctypes.c_int64]
        lib.H5Sget_simple_extent_dims.argtypes = [ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p]
        lib.H5Dread.argtypes = [ctypes.c_int64] * 5 + [ctypes.c_void_p]
        lib.H5Gget_info.argtypes = [ctypes.c_int64, ctypes.c_void_p]
        for fn in ("H5Dclose", "H5Sclose", "H5Fclose", "H5Gclose"):
            getattr(lib, fn).argtypes = [ctypes.c_int64]
        return lib
    return None


def _read_with_library(lib, path, names):
    fid = lib.H5Fopen(path.encode(), 0, 0)                   # H5F_ACC_RDONLY, H5P_DEFAULT
    assert fid >= 0, "libhdf5 refused the file"
    gid = lib.H5Gopen2(fid, b"/", 0)
    info = (ctypes.c_uint8 * 64)()
    assert lib.H5Gget_info(gid, info) >= 0
    nlinks = int.from_bytes(bytes(info[8:16]), "little")      # H5G_info_t.nlinks (after a 4-byte enum + pad)
    lib.H5Gclose(gid)
    native_double = ctypes.c_int64.in_dll(lib, "H5T_NATIVE_DOUBLE_g").value
    out = {}
    for n in names:
        did = lib.H5Dopen2(fid, ("/" + n).encode(), 0)
        assert did >= 0, n
        sid = lib.H5Dget_space(did)
        rank = lib.H5Sget_simple_extent_ndims(sid)
        dims = (ctypes.c_uint64 * rank)()
        lib.H5Sget_simple_extent_dims(sid, dims, None)
        buf = np.empty([int(d) for d in dims], dtype=np.float64)
        assert lib.H5Dread(did, native_double, 0, 0, 0, buf.ctypes.data) >= 0, n
        out[n] = buf
        lib.H5Sclose(sid)
        lib.H5Dclose(did)
    lib.H5Fclose(fid)
    return nlinks, out


def _datasets(rng):
    return {"x_0": np.linspace(0.5, 9.5, 10), "x_1": np.linspace(-1.0, 1.0, 7),
            "rho": rng.standard_normal((7, 10)), "u": rng.standard_normal((7, 10, 3)),
            "scalar1d": rng.standard_normal(5), "a_long_dataset_name_2": rng.standard_normal((2, 3, 4))}


def test_writer_round_trip_reader(tmp_path):
    rng = np.random.default_rng(0)
    data = _datasets(rng)
    path = str(tmp_path / "f.h5")
    write_h5(path, data)
    r = _reader()(path)
    got = dict(r.group_entries(r.root_header))
    assert sorted(got) == sorted(data)
    for n, a in data.items():
        assert np.array_equal(r.dataset(got[n]), a), n


@pytest.mark.skipif(_libhdf5() is None, reason="no HDF5 library on this machine")
def test_writer_read_by_hdf5_library(tmp_path):
    rng = np.random.default_rng(1)
    data = _datasets(rng)
    # more datasets than one default symbol node holds (8): the leaf K grows
    for i in range(12):
        data["extra_%02d" % i] = rng.standard_normal((3, 4))
    path = str(tmp_path / "g.h5")
    write_h5(path, data)
    nlinks, got = _read_with_library(_libhdf5(), path, list(data))
    assert nlinks == len(data)
    for n, a in data.items():
        assert np.array_equal(got[n], a), n


def test_h5file_pylbm_layout(tmp_path):
    """pylbm's H5File contents: x_i coordinates, scalars transposed, vectors with three
    components (the third zero in 2D), and an XDMF file that points at them."""
    x = np.linspace(0.5, 15.5, 16)
    y = np.linspace(0.25, 5.75, 12)
    X, Y = np.meshgrid(x, y, indexing="ij")
    rho = 1.0 + 0.1 * np.sin(X) * np.cos(Y)
    qx, qy = 0.01 * X, -0.02 * Y
    h = H5File(None, "lbm_output.h5", str(tmp_path / "out"), timestep=3)
    h.set_grid(x, y)
    h.add_scalar("rho", rho)
    h.add_scalar("qx", lambda a: a, qx)
    h.add_vector("velocity", [qx, qy])
    h.save()
    path = str(tmp_path / "out" / "lbm_output_3.h5")
    r = _reader()(path)
    ent = dict(r.group_entries(r.root_header))
    assert sorted(ent) == ["qx", "rho", "velocity", "x_0", "x_1"]
    assert np.array_equal(r.dataset(ent["x_0"]), x) and np.array_equal(r.dataset(ent["x_1"]), y)
    assert np.array_equal(r.dataset(ent["rho"]), rho.T)
    assert np.array_equal(r.dataset(ent["qx"]), qx.T)
    v = r.dataset(ent["velocity"])
    assert v.shape == (12, 16, 3)
    assert np.array_equal(v[..., 0], qx.T) and np.array_equal(v[..., 1], qy.T) and not v[..., 2].any()
    lib = _libhdf5()
    if lib is not None:
        _, got = _read_with_library(lib, path, ["rho", "velocity"])
        assert np.array_equal(got["rho"], rho.T) and np.array_equal(got["velocity"], v)
    root = ET.parse(str(tmp_path / "out" / "lbm_output_3.xdmf")).getroot()
    grid = root.find("Domain/Grid")
    assert grid.find("Topology").get("TopologyType") == "2DRectMesh"
    assert grid.find("Topology").get("NumberOfElements") == "16 12"
    items = [d.text.strip() for d in grid.find("Geometry").findall("DataItem")]
    assert items == ["lbm_output_3.h5:/x_0", "lbm_output_3.h5:/x_1"]
    attrs = {a.get("Name"): a for a in grid.findall("Attribute")}
    assert attrs["rho"].find("DataItem").get("Dimensions") == "12 16"
    assert attrs["rho"].find("DataItem").text.strip() == "lbm_output_3.h5:/rho"
    assert attrs["velocity"].get("AttributeType") == "Vector"


def test_h5file_reads_back_a_reference_fixture_layout():
    """The writer reproduces, dataset for dataset, the contents the reader finds in one
    of pylbm's own files (names, shapes, values), written from the converted .npz."""
    ref = np.load(os.path.join(HERE, "golden", "h5", "test2D_lid_driven_cavity.npz"))
    data = {k: ref[k] for k in ref.files}
    assert {"x_0", "x_1", "rho", "qx", "qy"} <= set(data)
    assert data["rho"].shape == (data["x_1"].size, data["x_0"].size)
