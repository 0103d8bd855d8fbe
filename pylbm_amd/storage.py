"""Device-resident population and moment arrays.

Replaces pylbm's ``Array`` / containers (pylbm/storage.py:24-524,
pylbm/container.py:10-106) for the hip generator.  The data live in HBM in
a structure-of-arrays layout (population-major, then the spatial axes in
reference order with the last one contiguous, as pylbm's numpy SOA storage
``(nv, nx, ny, nz)``), with every row padded to a multiple of 128 bytes and
shifted so the first interior element of each row is 128-byte aligned.  Host
copies are made only on access (``arr[k]``, ``arr._in(k)``), like the
Loo.py arrays' ``.get()`` (pylbm/storage.py:126-157).
"""
import ctypes

import numpy as np
import sympy as sp

from .runtime import DEVIATION_FORMATS, ELEM_SIZE, LBM_BF16D, LBM_F16D, LBM_F32, LBM_F32D, LBM_F64, LbmLayout

__all__ = ["make_layout", "DeviceArray", "STORAGE_BY_DTYPE"]

STORAGE_BY_DTYPE = {
    "float64": LBM_F64, "double": LBM_F64,
    "float32": LBM_F32, "single": LBM_F32,
    "float32-shifted": LBM_F32D, "f32-shifted": LBM_F32D,
    "float16": LBM_F16D, "f16-shifted": LBM_F16D, "float16-shifted": LBM_F16D,
    "bfloat16": LBM_BF16D, "bf16-shifted": LBM_BF16D, "bfloat16-shifted": LBM_BF16D,
}
_NP_STORE = {LBM_F64: np.float64, LBM_F32: np.float32, LBM_F16D: np.float16, LBM_BF16D: np.uint16,
             LBM_F32D: np.float32}


def make_layout(ns, shape_halo, halo, storage, interleaved=False):
    """lbm_layout of an array with ``ns`` fields over ``shape_halo`` (reference axes).

    interleaved=False: population planes (each field contiguous);
    interleaved=True: the ns rows of every (o, m) line contiguous, which keeps
    the streams of one workgroup within a few DRAM pages."""
    dim = len(shape_halo)
    n = [1] * (3 - dim) + [int(s) for s in shape_halo]
    h = [0] * (3 - dim) + [int(v) for v in halo]
    es = ELEM_SIZE[storage]
    align = 128 // es
    lead = (align - h[2] % align) % align
    pitch = -(-(lead + n[2]) // align) * align
    L = LbmLayout()
    L.dim, L.ns, L.storage, L.elem_size = dim, ns, storage, es
    for a in range(3):
        L.n[a] = n[a]
        L.halo[a] = h[a]
    L.pitch, L.lead = pitch, lead
    if interleaved:
        L.pop_stride = pitch
        L.row_stride = ns * pitch
    else:
        L.pop_stride = n[0] * n[1] * pitch
        L.row_stride = pitch
    return L


def layout_elements(L):
    """Number of elements the array of layout L spans."""
    return L.ns * L.n[0] * L.n[1] * L.pitch


def is_interleaved(L):
    return L.pop_stride < L.row_stride


def bf16_to_f32(bits):
    return (bits.astype(np.uint32) << 16).view(np.float32)


def f32_to_bf16(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


class DeviceArray:
    """``ns`` fields on the halo-inclusive grid, stored on the GPU."""

    def __init__(self, rt, ns, shape_halo, halo, storage=LBM_F64, wshift=None, interleaved=False):
        self.rt = rt
        self.ns = ns
        self.shape_halo = tuple(int(s) for s in shape_halo)
        self.vmax = tuple(int(h) for h in halo)
        self.storage = storage
        self.layout = make_layout(ns, shape_halo, halo, storage, interleaved)
        self.interleaved = bool(interleaved)
        self.nbytes = layout_elements(self.layout) * self.layout.elem_size
        self.ptr = rt.malloc(self.nbytes)
        self.consm = {}
        self.wshift = np.zeros(ns) if wshift is None else np.asarray(wshift, dtype=np.float64)
        self.vel = None   # the scheme's lbm_velocities, given by the container (update)

    def __del__(self):
        try:
            if self.ptr:
                self.rt.free(self.ptr)
                self.ptr = None
        except Exception:
            pass

    def update(self):
        """Array.update (pylbm/storage.py:306-367): every ghost slot of every field takes
        the periodic wrap of the interior, axis by axis (asynchronous on the stream)."""
        if self.vel is None:
            raise RuntimeError("update needs the velocities of the array's scheme")
        self.rt.check(self.rt.lib.lbm_ghost_fill_all(ctypes.c_void_p(self.ptr), ctypes.byref(self.layout),
                                                     ctypes.byref(self.vel), self.rt.stream), "update")

    # ------------------------------------------------------------ geometry
    @property
    def nspace(self):
        return self.shape_halo

    @property
    def nv(self):
        return self.ns

    @property
    def dim(self):
        return len(self.shape_halo)

    def set_conserved_moments(self, consm):
        self.consm.update(consm)

    def _index(self, key):
        if isinstance(key, (sp.Symbol, sp.IndexedBase)) or hasattr(key, "free_symbols"):
            return self.consm[key]
        return key

    def _padded_shape(self, nfields):
        L = self.layout
        return (nfields, L.n[0], L.n[1], L.pitch)

    def _unpad(self, raw):
        L = self.layout
        a = raw[..., L.lead:L.lead + L.n[2]]
        return a.reshape((raw.shape[0],) + self.shape_halo)

    # ------------------------------------------------------------ host I/O
    def _raw_all(self):
        """Every stored element as (ns, n0, n1, pitch) (storage dtype)."""
        L = self.layout
        if not self.interleaved:
            raw = np.empty(self._padded_shape(self.ns), dtype=_NP_STORE[self.storage])
            return self.rt.d2h(raw, self.ptr)
        raw = np.empty((L.n[0], L.n[1], self.ns, L.pitch), dtype=_NP_STORE[self.storage])
        self.rt.d2h(raw, self.ptr)
        return np.ascontiguousarray(np.moveaxis(raw, 2, 0))

    def _write_all(self, raw):
        if not self.interleaved:
            self.rt.h2d(self.ptr, raw)
        else:
            self.rt.h2d(self.ptr, np.ascontiguousarray(np.moveaxis(raw, 0, 2)))

    def fields_raw(self, first=0, count=None):
        """Stored values of fields [first, first + count) as a host array (storage dtype)."""
        count = self.ns - first if count is None else count
        L = self.layout
        if self.interleaved:
            return self._unpad(self._raw_all()[first:first + count])
        raw = np.empty(self._padded_shape(count), dtype=_NP_STORE[self.storage])
        off = first * L.pop_stride * L.elem_size
        self.rt.d2h(raw, self.ptr + off)
        return self._unpad(raw)

    def fields(self, first=0, count=None):
        """Field values in FP64 (deviation formats get w_k rho0 added back)."""
        raw = self.fields_raw(first, count)
        if self.storage == LBM_BF16D:
            vals = bf16_to_f32(raw).astype(np.float64)
        else:
            vals = raw.astype(np.float64)
        if self.storage in DEVIATION_FORMATS:
            count = raw.shape[0]
            vals += self.wshift[first:first + count].reshape((count,) + (1,) * self.dim)
        return vals

    def set_fields(self, values, first=0):
        """Write FP64 host values (count, *shape_halo) into fields first.. (converted to storage)."""
        values = np.asarray(values, dtype=np.float64)
        count = values.shape[0]
        L = self.layout
        if self.storage in DEVIATION_FORMATS:
            values = values - self.wshift[first:first + count].reshape((count,) + (1,) * self.dim)
        if self.storage == LBM_BF16D:
            stored = f32_to_bf16(values.astype(np.float32))
        else:
            stored = values.astype(_NP_STORE[self.storage])
        if self.interleaved:
            raw = self._raw_all() if count < self.ns else np.zeros(self._padded_shape(self.ns), stored.dtype)
            raw[first:first + count, ..., L.lead:L.lead + L.n[2]] = stored.reshape((count, L.n[0], L.n[1], L.n[2]))
            self._write_all(raw)
            return
        raw = np.zeros(self._padded_shape(count), dtype=stored.dtype)
        raw[..., L.lead:L.lead + L.n[2]] = stored.reshape((count, L.n[0], L.n[1], L.n[2]))
        self.rt.h2d(self.ptr + first * L.pop_stride * L.elem_size, raw)

    def __getitem__(self, key):
        k = self._index(key)
        if isinstance(k, (int, np.integer)):
            return self.fields(int(k), 1)[0]
        return self.fields()[k]

    def __setitem__(self, key, values):
        k = self._index(key)
        if isinstance(k, (int, np.integer)):
            cur = np.empty((1,) + self.shape_halo)
            cur[0] = values
            self.set_fields(cur, int(k))
        else:
            cur = self.fields()
            cur[k] = values
            self.set_fields(cur)

    def _in(self, key):
        sl = tuple(slice(h, -h) if h > 0 else slice(None) for h in self.vmax)
        return self[key][sl]

    def copy_from(self, other):
        assert other.nbytes == self.nbytes
        self.rt.d2d(self.ptr, other.ptr, self.nbytes)
