"""The generator module object pylbm's ``call_genfunction`` would call for backend 'hip'.

pylbm calls its generated routines by keyword: ``call_genfunction``
(pylbm/symbolic.py:288-299) reads ``inspect.getfullargspec(func).args`` and takes
each argument by name from ``BaseAlgorithm._get_args`` (pylbm/algorithm/base.py:
630-667: ``f, fnew, m, t, dt, ...``) or ``BoundaryMethod._get_args``
(pylbm/boundary.py:358-376: ``f, istore, iload0, rhs, ncond, ...``).
:class:`HipModule` exposes the routines of the hot path under those names and
argument names, over the C ABI of include/pylbm_hip.h:

* ``one_time_step(f, fnew, t, dt)``       generated one_time_step (pull.py:11-59)
* ``f2m(f, m)``, ``m2f(m, f)``, ``equilibrium(m)``   generated routines (base.py:298-402)
* ``transport(f, fnew)``, ``relaxation(m)``, ``source_term(m, t, dt)``
                                          generated split routines (base.py:273-296,
                                          404-437, 439-504)
* ``update(f)``                           Array.update: periodic wrap of every
                                          halo slot (storage.py:306-367)
* ``bounce_back``, ``anti_bounce_back``   (f, istore, iload0, rhs, ncond)
                                          (boundary.py:430-469, 646-684)
* ``neumann``, ``neumannx/y/z``           (f, istore, iload0, ncond) (boundary.py:787-895)
* ``Bouzidi_bounce_back``                 (f, fcopy, istore, iload0, iload1, rhs, dist, ncond)
                                          (boundary.py:537-623)
* ``Bouzidi_anti_bounce_back``            (f, istore, iload0, iload1, rhs, dist, ncond)
                                          (boundary.py:707-760)

Buffers are device pointers (integers); ``istore`` / ``iload0`` are flat uint32
element indices of the population layout (``pylbm_amd.boundary._flat`` converts
pylbm's ``(k, ix, iy[, iz])`` rows).  ``rhs`` is pylbm's own right-hand side:
FP64, one value per link (``BoundaryMethod.rhs``, pylbm/boundary.py:204-205), or
None.  For FP64 storage it goes to the link kernel as it is; for every other
format the module first makes the compute-type rhs the kernel takes
(``lbm_links_prepare_rhs``, one small launch on the same stream): FP32 values,
and for the deviation formats (``float32-shifted``, ``float16``, ``bfloat16``)
the rest-population term ``sign w(load) - w(store)`` (Bouzidi:
``sign s w(load0) + (1 - s) w(load1) - w(store)``) added, because their link
kernels combine stored deviations ``f - w``; that term is applied even when
``rhs`` is None (an anti-bounce-back list without a value still needs
``-2 w_k``).  A list that reads slots it writes must be given resolved
(``pylbm_amd.boundary.host_lists``) together with a workspace, the keyword-only
``work`` of ncond values (keyword-only, so that the argument names pylbm reads
are exactly its own).  For the Bouzidi routines ``dist`` is the FP64 device
array of pylbm's per-link ``s``, ``iload1`` the second loads, and a list that
loads slots it stores is given with ``work`` (pylbm's ``fcopy`` semantics: every
value is gathered before any store; ``fcopy`` itself is not read).  One call is
one launch: an anti-bounce-back list one of whose links loads a slot an earlier
link stores (pylbm's in-place loop reads that new value) is given as the lists
of ``pylbm_amd.boundary.bouzidi_levels``, one call per level in order.  The
coefficient and rhs buffers the module makes are reused across calls (grown to
the largest ncond) and their contents recomputed on every call, so new values
behind the same addresses are never served stale; they are released by
:meth:`HipModule.free` or when the module is collected.  Every routine is
asynchronous on the module's stream and raises RuntimeError with the library's
message on failure.
"""
import ctypes

from . import runtime as rtm

__all__ = ["HipModule"]


class HipModule:
    """pylbm's routine names for one generated scheme module and one population layout."""

    def __init__(self, gen, layout, mlayout, vel, params, rt=None):
        self.rt = rt or rtm.get()
        self.core = self.rt.lib
        self.gen = gen
        self.layout, self.mlayout, self.vel, self.params = layout, mlayout, vel, params
        self.stream = self.rt.stream

    @classmethod
    def from_simulation(cls, sim):
        """The module of a pylbm_amd Simulation: its generated module, layouts,
        velocities and relaxation parameters (folded link lists are not used)."""
        params = rtm.LbmParams.from_buffer_copy(sim.params)
        for name in ("bbmask", "bbtab_bits", "bbtab_base", "bbtab_src", "bbtab_sign", "bbtab_rhs", "bbtab_rel"):
            setattr(params, name, None)
        return cls(sim.module, sim.container.F.layout, sim.container.m.layout, sim.vel, params, sim.rt)

    def _ok(self, rc, what):
        if rc:
            raise RuntimeError("%s: %s" % (what, self.core.lbm_last_error().decode()))

    @staticmethod
    def _p(x):
        return ctypes.c_void_p(int(x)) if x is not None else None

    # ------------------------------------------------ algorithm (base.py:603-628)
    def one_time_step(self, f, fnew, t, dt):
        self.params.t, self.params.dt = float(t), float(dt)
        self._ok(self.gen.lbm_gen_step(self._p(f), self._p(fnew), ctypes.byref(self.layout),
                                       ctypes.byref(self.params), 0, self.stream), "one_time_step")

    def f2m(self, f, m):
        self._ok(self.gen.lbm_gen_f2m(self._p(f), ctypes.byref(self.layout), self._p(m),
                                      ctypes.byref(self.mlayout), ctypes.byref(self.params), self.stream), "f2m")

    def m2f(self, m, f):
        self._ok(self.gen.lbm_gen_m2f(self._p(m), ctypes.byref(self.mlayout), self._p(f),
                                      ctypes.byref(self.layout), ctypes.byref(self.params), self.stream), "m2f")

    def transport(self, f, fnew):
        self._ok(self.gen.lbm_gen_transport(self._p(f), self._p(fnew), ctypes.byref(self.layout),
                                            ctypes.byref(self.params), self.stream), "transport")

    def relaxation(self, m):
        self._ok(self.gen.lbm_gen_relaxation(self._p(m), ctypes.byref(self.mlayout), ctypes.byref(self.params),
                                             self.stream), "relaxation")

    def source_term(self, m, t, dt):
        self.params.t = float(t)
        self.params.dt = float(dt)
        self._ok(self.gen.lbm_gen_source_term(self._p(m), ctypes.byref(self.mlayout), ctypes.byref(self.params),
                                              self.stream), "source_term")

    def equilibrium(self, m):
        self._ok(self.gen.lbm_gen_equilibrium(self._p(m), ctypes.byref(self.mlayout), self.stream), "equilibrium")

    # ------------------------------------------- ghost update (storage.py:306-367)
    def update(self, f):
        self._ok(self.core.lbm_ghost_fill_all(self._p(f), ctypes.byref(self.layout), ctypes.byref(self.vel),
                                              self.stream), "update")

    # ------------------------------------- boundary routines (boundary.py:430-895)
    def _buffer(self, name, nvals):
        """A module-owned device buffer of at least nvals compute-type values, grown on demand."""
        cs = 8 if self.layout.storage == rtm.LBM_F64 else 4
        bufs = self.__dict__.setdefault("_bufs", {})
        ptr, cap = bufs.get(name, (None, 0))
        if cap < nvals:
            if ptr:
                self.rt.free(ptr)
            cap = max(1, nvals)
            ptr = self.rt.malloc(cap * cs)
            bufs[name] = (ptr, cap)
        return ptr

    def _rhs(self, rhs, istore, iload0, iload1, dist, sign, ncond):
        """The compute-type rhs the link kernel takes, from pylbm's FP64 rhs (or None)."""
        if self.layout.storage == rtm.LBM_F64:
            return rhs
        if rhs is None and self.layout.storage not in rtm.DEVIATION_FORMATS:
            return None
        out = self._buffer("rhs", ncond)
        w = (ctypes.c_double * rtm.MAX_Q)(*self.params.wshift)
        self._ok(self.core.lbm_links_prepare_rhs(ctypes.byref(self.layout), w, self._p(istore), self._p(iload0),
                                                 self._p(iload1), self._p(dist), float(sign), self._p(rhs),
                                                 self._p(out), int(ncond), self.stream), "rhs")
        return out

    def _links(self, f, istore, iload0, rhs, sign, ncond, work, what):
        rhs = self._rhs(rhs, istore, iload0, None, None, sign, ncond)
        self._ok(self.core.lbm_links_apply(self._p(f), ctypes.byref(self.layout), self._p(istore),
                                           self._p(iload0), self._p(rhs), None, float(sign), int(ncond),
                                           1 if work is not None else 0, self._p(work), self.stream), what)

    def bounce_back(self, f, istore, iload0, rhs, ncond, *, work=None):
        self._links(f, istore, iload0, rhs, 1.0, ncond, work, "bounce_back")

    def anti_bounce_back(self, f, istore, iload0, rhs, ncond, *, work=None):
        self._links(f, istore, iload0, rhs, -1.0, ncond, work, "anti_bounce_back")

    def neumann(self, f, istore, iload0, ncond, *, work=None):
        self._links(f, istore, iload0, None, 1.0, ncond, work, "neumann")

    neumannx = neumanny = neumannz = neumann

    # ------------------------------- interpolated routines (boundary.py:472-760)
    def _coefficients(self, dist, ncond, sign):
        # recomputed on every call (one small launch, ordered before the link launch
        # on the same stream) into buffers reused across calls
        ca, cb = self._buffer("ca%+d" % sign, ncond), self._buffer("cb%+d" % sign, ncond)
        self._ok(self.core.lbm_links_bouzidi_coefficients(self._p(dist), sign, self._p(ca), self._p(cb),
                                                          int(ncond), self.layout.storage, self.stream),
                 "Bouzidi coefficients")
        return ca, cb

    def _bouzidi(self, f, istore, iload0, iload1, rhs, dist, ncond, work, sign, what):
        ca, cb = self._coefficients(dist, ncond, sign)
        rhs = self._rhs(rhs, istore, iload0, iload1, dist, sign, ncond)
        self._ok(self.core.lbm_links_apply_bouzidi(self._p(f), ctypes.byref(self.layout), self._p(istore),
                                                   self._p(iload0), self._p(iload1), self._p(ca), self._p(cb),
                                                   self._p(rhs), int(ncond), 1 if work is not None else 0,
                                                   self._p(work), self.stream), what)

    def Bouzidi_bounce_back(self, f, fcopy, istore, iload0, iload1, rhs, dist, ncond, *, work=None):
        self._bouzidi(f, istore, iload0, iload1, rhs, dist, ncond, work, 1.0, "Bouzidi_bounce_back")

    def Bouzidi_anti_bounce_back(self, f, istore, iload0, iload1, rhs, dist, ncond, *, work=None):
        self._bouzidi(f, istore, iload0, iload1, rhs, dist, ncond, work, -1.0, "Bouzidi_anti_bounce_back")

    def free(self):
        """Release the coefficient and rhs buffers the module made (after the stream's
        work that uses them)."""
        bufs = self.__dict__.pop("_bufs", {})
        if bufs:
            self.rt.sync()
        for ptr, _ in bufs.values():
            self.rt.free(ptr)

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
