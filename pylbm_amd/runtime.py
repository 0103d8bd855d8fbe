"""ctypes binding of the C ABI declared in include/pylbm_hip.h.

The core library ``libpylbm_hip.so`` is built in-tree next to this file
(``python -c "import __graft_entry__ as g; g.build()"``).  There is no
fallback: if the library or a GPU is missing, :func:`get` raises
:class:`HipUnavailable` with the reason.
"""
import ctypes
import os
import threading

import numpy as np

__all__ = ["HipUnavailable", "HipError", "LbmLayout", "LbmVelocities", "LbmParams", "LbmLinkSet",
           "get", "lib_path", "LBM_F64", "LBM_F32", "LBM_F16D", "LBM_BF16D", "LBM_F32D", "MAX_Q",
           "DEVIATION_FORMATS"]

LBM_F64, LBM_F32, LBM_F16D, LBM_BF16D, LBM_F32D = 0, 1, 2, 3, 4
MAX_Q = 32
ELEM_SIZE = {LBM_F64: 8, LBM_F32: 4, LBM_F16D: 2, LBM_BF16D: 2, LBM_F32D: 4}
DEVIATION_FORMATS = (LBM_F16D, LBM_BF16D, LBM_F32D)

_HERE = os.path.dirname(os.path.abspath(__file__))


def lib_path():
    return os.path.join(_HERE, "libpylbm_hip.so")


class HipUnavailable(RuntimeError):
    """The HIP library or a GPU device is not available."""


class HipError(RuntimeError):
    """A C-ABI entry point returned a non-zero status."""


class LbmLayout(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("ns", ctypes.c_int32), ("storage", ctypes.c_int32),
                ("elem_size", ctypes.c_int32), ("n", ctypes.c_int64 * 3), ("halo", ctypes.c_int32 * 3),
                ("_pad", ctypes.c_int32), ("pitch", ctypes.c_int64), ("lead", ctypes.c_int64),
                ("pop_stride", ctypes.c_int64), ("row_stride", ctypes.c_int64)]


class LbmVelocities(ctypes.Structure):
    _fields_ = [("q", ctypes.c_int32), ("periodic_faces", ctypes.c_int32), ("v", (ctypes.c_int8 * 3) * MAX_Q)]


class LbmParams(ctypes.Structure):
    _fields_ = [("omega", ctypes.c_double * MAX_Q), ("one_minus", ctypes.c_double * MAX_Q),
                ("wshift", ctypes.c_double * MAX_Q), ("cref", ctypes.c_double * MAX_Q),
                ("dt", ctypes.c_double), ("t", ctypes.c_double), ("bbmask", ctypes.c_void_p),
                ("bbtab_bits", ctypes.c_void_p), ("bbtab_base", ctypes.c_void_p), ("bbtab_src", ctypes.c_void_p),
                ("bbtab_sign", ctypes.c_void_p), ("bbtab_rhs", ctypes.c_void_p), ("bbtab_rel", ctypes.c_void_p),
                ("bbtab_src2", ctypes.c_void_p), ("bbtab_c2", ctypes.c_void_p), ("bbtab_rel2", ctypes.c_void_p),
                ("origin", ctypes.c_double * 3), ("dx", ctypes.c_double),
                ("xcd_remap", ctypes.c_int32), ("vec_cells", ctypes.c_int32)]


class LbmLinkSet(ctypes.Structure):
    _fields_ = [("d_store", ctypes.c_void_p), ("d_load", ctypes.c_void_p), ("d_rhs", ctypes.c_void_p),
                ("d_sign", ctypes.c_void_p), ("sign", ctypes.c_double), ("n", ctypes.c_int64),
                ("two_phase", ctypes.c_int32), ("merge_prev", ctypes.c_int32), ("d_work", ctypes.c_void_p),
                ("d_load1", ctypes.c_void_p), ("d_ca", ctypes.c_void_p), ("d_cb", ctypes.c_void_p)]


_P = ctypes.c_void_p
_I = ctypes.c_int
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_D = ctypes.c_double

CORE_SIGNATURES = {
    "lbm_last_error": ([], ctypes.c_char_p),
    "lbm_device_count": ([ctypes.POINTER(_I)], _I),
    "lbm_set_device": ([_I], _I),
    "lbm_device_name": ([_I, ctypes.c_char_p, _I], _I),
    "lbm_malloc": ([ctypes.POINTER(_P), _U64], _I),
    "lbm_free": ([_P], _I),
    "lbm_memcpy_h2d": ([_P, _P, _U64, _P], _I),
    "lbm_memcpy_d2h": ([_P, _P, _U64, _P], _I),
    "lbm_memcpy_d2d": ([_P, _P, _U64, _P], _I),
    "lbm_memset": ([_P, _I, _U64, _P], _I),
    "lbm_copy": ([_P, _P, _U64, _P], _I),
    "lbm_stream_create": ([ctypes.POINTER(_P)], _I),
    "lbm_stream_destroy": ([_P], _I),
    "lbm_stream_sync": ([_P], _I),
    "lbm_device_sync": ([], _I),
    "lbm_event_create": ([ctypes.POINTER(_P)], _I),
    "lbm_event_destroy": ([_P], _I),
    "lbm_event_record": ([_P, _P], _I),
    "lbm_event_elapsed_ms": ([_P, _P, ctypes.POINTER(ctypes.c_float)], _I),
    "lbm_event_sync": ([_P], _I),
    "lbm_stream_wait_event": ([_P, _P], _I),
    "lbm_host_malloc": ([ctypes.POINTER(_P), _U64], _I),
    "lbm_host_free": ([_P], _I),
    "lbm_ghost_fill": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmVelocities), _P], _I),
    "lbm_ghost_fill_all": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmVelocities), _P], _I),
    "lbm_links_apply": ([_P, ctypes.POINTER(LbmLayout), _P, _P, _P, _P, _D, _I64, _I, _P, _P], _I),
    "lbm_links_apply_bouzidi": ([_P, ctypes.POINTER(LbmLayout), _P, _P, _P, _P, _P, _P, _I64, _I, _P, _P], _I),
    "lbm_links_bouzidi_coefficients": ([_P, _D, _P, _P, _I64, _I, _P], _I),
    "lbm_links_prepare_rhs": ([ctypes.POINTER(LbmLayout), ctypes.POINTER(_D), _P, _P, _P, _P, _D, _P, _P, _I64, _P],
                              _I),
    "lbm_links_apply_sets": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmLinkSet), _I, _P], _I),
    "lbm_reduce_blocks": ([ctypes.POINTER(LbmLayout), ctypes.POINTER(_I64)], _I),
    "lbm_reduce_populations": ([_P, ctypes.POINTER(LbmLayout), _P, _P, ctypes.POINTER(_D),
                                ctypes.POINTER(_D), _P], _I),
    "lbm_reduce_populations_noalloc": ([_P, ctypes.POINTER(LbmLayout), _P, ctypes.POINTER(_D),
                                        ctypes.POINTER(_D), _P], _I),
}

_F = ctypes.c_float
_FP = ctypes.POINTER(ctypes.c_float)

class LbmSurNet(ctypes.Structure):
    _fields_ = [("act", ctypes.c_void_p * 4), ("xalt", ctypes.c_void_p), ("wpk", ctypes.c_void_p * 4),
                ("bias", ctypes.c_void_p * 4), ("hidden", ctypes.c_int32), ("elem", ctypes.c_int32)]


SURROGATE_SIGNATURES = {
    "lbm_sur_run": ([ctypes.POINTER(LbmSurNet), _P, _P, _FP, _FP, _FP, _I64, _I, _I, _I, _P], _I),
    "lbm_sur_conv": ([_P, _I, _P, _P, _P, _P, _P, _I, _I, _I, _I, _I, _I, _P], _I),
    "lbm_sur_apply_step": ([_P, _I, _P, _P, _P, _P, _FP, _FP, _FP, _I, _I, _I, _I, _P], _I),
    "lbm_sur_wgrad_blocks": ([], _I),
    "lbm_sur_wgrad": ([_P, _I, _P, _I, _I, _I, _I, _P, _P, _P, _I, _P], _I),
    "lbm_sur_mse": ([_P, _P, _I64, _F, _P, _P, ctypes.POINTER(_D), _I, _P], _I),
    "lbm_sur_pack": ([_P, _I, _I, _I, _I, _P, _P, _I, _P], _I),
    "lbm_sur_adam": ([_P, _P, _P, _P, _P, _P, _I, _I, _I, _I, _F, _F, _F, _F, _I, _P], _I),
    "lbm_sur_make_input": ([_P, _I, _I, _I, _FP, _FP, _P, _I, _P], _I),
    "lbm_sur_make_target": ([_P, _P, _I, _I, _I, _FP, _P, _P], _I),
    "lbm_sur_moments_to_state": ([_P, ctypes.POINTER(LbmLayout), _I, _I, _I, _I, _I, _P, _P], _I),
    "lbm_sur_state_to_moments": ([_P, _I, _I, ctypes.POINTER(LbmLayout), _I, _I, _I, _P, _P], _I),
    "lbm_sur_rel_l2": ([_P, _P, _I64, _P, ctypes.POINTER(_D), _P], _I),
    "lbm_sur_perturb": ([_P, _I64, _FP, ctypes.c_uint64, _P], _I),
    "lbm_sur_rel_l2_async": ([_P, _P, _I64, _P, _P, _P], _I),
    "lbm_sur_populations_to_state": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(_D), _I, _I, _P, _P], _I),
    "lbm_sur_judge_read": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(_D), _I, _I, _P, _P, _P, _FP, _FP, _P, _I, _P,
                            _P, _P], _I),
}

GEN_SIGNATURES = {
    "lbm_gen_info": ([ctypes.POINTER(ctypes.c_int32)] * 3, _I),
    "lbm_gen_step": ([_P, _P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmParams), _I, _P], _I),
    "lbm_gen_f2m": ([_P, ctypes.POINTER(LbmLayout), _P, ctypes.POINTER(LbmLayout),
                     ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_m2f": ([_P, ctypes.POINTER(LbmLayout), _P, ctypes.POINTER(LbmLayout),
                     ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_equilibrium": ([_P, ctypes.POINTER(LbmLayout), _P], _I),
    "lbm_gen_eq_m2f": ([_P, ctypes.POINTER(LbmLayout), _P, ctypes.POINTER(LbmLayout),
                        ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_transport": ([_P, _P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_relaxation": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_source_term": ([_P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmParams), _P], _I),
    "lbm_gen_run_blocked": ([_P, _P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmParams), _I64, _I, _I, _I,
                             ctypes.POINTER(_I), _P], _I),
    "lbm_gen_run": ([_P, _P, ctypes.POINTER(LbmLayout), ctypes.POINTER(LbmVelocities),
                     ctypes.POINTER(LbmParams), ctypes.POINTER(LbmLinkSet), _I, _I64, _I, _I,
                     ctypes.POINTER(_I), _P], _I),
}


def bind(lib, signatures):
    for name, (args, res) in signatures.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    return lib


def load_core_library(path=None):
    """Load the core library without touching the GPU (used to check exports)."""
    path = path or lib_path()
    if not os.path.exists(path):
        raise HipUnavailable("%s is not built (run __graft_entry__.build())" % path)
    lib = bind(ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL), CORE_SIGNATURES)
    return bind(lib, SURROGATE_SIGNATURES)


class Runtime:
    """The loaded core library plus the device and stream the simulations use."""

    def __init__(self, device=0):
        self.lib = load_core_library()
        n = ctypes.c_int(0)
        rc = self.lib.lbm_device_count(ctypes.byref(n))
        if rc != 0 or n.value == 0:
            raise HipUnavailable("no HIP device available (%s)" % self.lib.lbm_last_error().decode())
        self.device = device
        self.check(self.lib.lbm_set_device(device), "lbm_set_device")
        buf = ctypes.create_string_buffer(256)
        self.check(self.lib.lbm_device_name(device, buf, 256), "lbm_device_name")
        self.device_name = buf.value.decode()
        s = ctypes.c_void_p()
        self.check(self.lib.lbm_stream_create(ctypes.byref(s)), "lbm_stream_create")
        self.stream = s

    def check(self, rc, what):
        if rc != 0:
            raise HipError("%s failed (%d): %s" % (what, rc, self.lib.lbm_last_error().decode()))

    # -------------------------------------------------------------- memory
    def malloc(self, nbytes):
        p = ctypes.c_void_p()
        self.check(self.lib.lbm_malloc(ctypes.byref(p), int(nbytes)), "lbm_malloc(%d)" % nbytes)
        return p.value

    def free(self, ptr):
        if ptr:
            self.check(self.lib.lbm_free(ctypes.c_void_p(ptr)), "lbm_free")

    def h2d(self, dptr, host):
        host = np.ascontiguousarray(host)
        self.check(self.lib.lbm_memcpy_h2d(ctypes.c_void_p(dptr), host.ctypes.data_as(ctypes.c_void_p),
                                           host.nbytes, self.stream), "lbm_memcpy_h2d")

    def d2h(self, host, dptr):
        assert host.flags.c_contiguous
        self.check(self.lib.lbm_memcpy_d2h(host.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(dptr),
                                           host.nbytes, self.stream), "lbm_memcpy_d2h")
        return host

    def d2d(self, dst, src, nbytes):
        self.check(self.lib.lbm_memcpy_d2d(ctypes.c_void_p(dst), ctypes.c_void_p(src), int(nbytes),
                                           self.stream), "lbm_memcpy_d2d")

    def sync(self):
        self.check(self.lib.lbm_stream_sync(self.stream), "lbm_stream_sync")

    def upload(self, array):
        """Copy a host array to a new device buffer; returns the device pointer."""
        array = np.ascontiguousarray(array)
        p = self.malloc(max(array.nbytes, 1))
        self.h2d(p, array)
        return p

    # -------------------------------------------------------------- events
    def event(self):
        e = ctypes.c_void_p()
        self.check(self.lib.lbm_event_create(ctypes.byref(e)), "lbm_event_create")
        return e

    def record(self, event):
        self.check(self.lib.lbm_event_record(event, self.stream), "lbm_event_record")

    def elapsed_ms(self, start, stop):
        ms = ctypes.c_float()
        self.check(self.lib.lbm_event_elapsed_ms(start, stop, ctypes.byref(ms)), "lbm_event_elapsed_ms")
        return float(ms.value)


_RUNTIME = None
_LOCK = threading.Lock()


def get(device=0):
    """The process-wide runtime (created on first use)."""
    global _RUNTIME
    with _LOCK:
        if _RUNTIME is None:
            _RUNTIME = Runtime(device)
        return _RUNTIME
