"""``Simulation(dico)`` / ``one_time_step()`` served by generated HIP kernels.

Mirrors pylbm/simulation.py:89-420 for ``generator='hip'``:

* construction: Domain, Scheme, code generation + compilation, containers,
  boundary link lists, initialisation (simulation.py:89-163);
* ``initialization``: init dict on the halo-inclusive meshgrid into the
  conserved moments, then equilibrium + m2f on the GPU (simulation.py:257-320);
* ``one_time_step``: periodic ghost fill, every link method in order, the
  fused pull step, swap F / Fnew, t += dt, nt += 1 (simulation.py:373-420);
* ``m[...]`` / ``F[...]`` read-back through a GPU f2m (simulation.py:200-243).

Extensions that keep the reference semantics: ``run(nsteps)`` performs
nsteps complete steps in one call into the library (no Python round-trip per
step); ``dtype`` selects the population storage ('float64' FP64, 'float32',
'float32-shifted' / 'float16' / 'bfloat16' = deviations f_k - w_k rho0 in
FP32 / FP16 / BF16; every non-FP64 format computes the collision in FP32 on
deviations from the rest state); ``conserved_sums()`` returns FP64 compensated global sums.
"""
import ctypes
import math
import time
import types

import numpy as np
import sympy as sp

from . import runtime as rtm
from .boundary import (Boundary, DeviceLinks, bouzidi_levels, host_lists, merge_flags, resolve_step,
                       unwritten_ghost_reads)
from .boundary import _unflat as _unflat_fn
from .domain import Domain
from .generator.build import load_module
from .monitoring import Monitor
from .scheme import Scheme
from .storage import STORAGE_BY_DTYPE, DeviceArray
from .utils import itemproperty

__all__ = ["Simulation"]

LBM_F64 = rtm.LBM_F64
# Largest table (links that are not a copy of the cell's own ksym population) the
# whole-step fold may build: table entries are scattered reads inside the step's
# waves, and 126,620 of them spread over C4's waves cost 37 us per step against a
# 4.5 us launch of the same copies (MI355X); smaller tables save their launches.
FOLD_TABLE_MAX = 1 << 16


def _canon_velocities(spec, box_label=None):
    vel = rtm.LbmVelocities()
    vel.q = spec.ns
    if box_label is not None:
        # reference faces (x-, x+, y-, y+, z-, z+) -> canonical axes (outer, mid, inner)
        shift = 2 * (3 - spec.dim)
        vel.periodic_faces = sum(1 << (shift + f) for f, lab in enumerate(box_label) if lab == -1)
    for k, v in enumerate(spec.velocities):
        c = [0] * (3 - spec.dim) + list(v)
        for a in range(3):
            vel.v[k][a] = c[a]
    return vel


def _no_extra_arguments(name, kwargs):
    """pylbm's routines take keyword arguments that override the arguments of its
    generated functions (pylbm/algorithm/base.py:669-681); the compiled routines
    here take none, so passing one is an error instead of being ignored."""
    if kwargs:
        raise NotImplementedError("%s: the hip generator's routines take no extra arguments (got %s)"
                                  % (name, ", ".join(sorted(kwargs))))


class _MomentView:
    """Moments of boundary points handed to user value functions (m[RHO] = ...)."""

    def __init__(self, arr, consm):
        self.arr = arr
        self.consm = consm

    def _k(self, key):
        return key if isinstance(key, (int, np.integer)) else self.consm[key]

    def __getitem__(self, key):
        return self.arr[self._k(key)]

    def __setitem__(self, key, value):
        self.arr[self._k(key)] = value


class _Container:
    """F, Fnew and m of one simulation (pylbm/container.py:10-106)."""

    gpu_support = True

    def __init__(self, rt, ns, shape_halo, vmax, storage, consm, wshift, interleaved=True, vel=None):
        self.F = DeviceArray(rt, ns, shape_halo, vmax, storage, wshift, interleaved)
        self.Fnew = DeviceArray(rt, ns, shape_halo, vmax, storage, wshift, interleaved)
        self._m = None
        self._rt, self._shape, self._vmax, self._consm, self._vel = rt, shape_halo, vmax, consm, vel
        for a in (self.F, self.Fnew):
            a.set_conserved_moments(consm)
            a.vel = vel
        self.nv = ns
        self.sorder = list(range(len(shape_halo) + 1))

    @property
    def m(self):
        if self._m is None:
            self._m = DeviceArray(self._rt, self.nv, self._shape, self._vmax, LBM_F64)
            self._m.set_conserved_moments(self._consm)
            self._m.vel = self._vel
        return self._m


class Simulation:
    """Lattice Boltzmann simulation on one MI355X (pylbm.Simulation, generator 'hip')."""

    def __init__(self, dico, sorder=None, dtype="float64", check_inverse=False, initialize=True,
                 rho0=1.0, device=0, layout="interleaved"):
        gen = str(dico.get("generator", "hip")).upper()
        if gen != "HIP":
            raise ValueError("pylbm_amd provides the 'hip' generator only (got %r)" % dico.get("generator"))
        if dtype not in STORAGE_BY_DTYPE:
            raise ValueError("dtype must be one of %s" % sorted(STORAGE_BY_DTYPE))
        self.rt = rtm.get(device)
        self.domain = Domain(dico, need_validation=False)
        # pylbm's storage order (the axis order of its host arrays, pylbm/storage.py:
        # 60-118) never changes a value or what the accessors return; the device
        # layout here is the backend's own, so a valid order is accepted and kept
        if sorder is not None:
            if sorted(int(i) for i in sorder) != list(range(self.domain.dim + 1)):
                raise ValueError("storage order must be a permutation of range(%d)" % (self.domain.dim + 1))
            self.sorder = [int(i) for i in sorder]
        else:
            self.sorder = list(range(self.domain.dim + 1))
        Monitor.set_size(int(np.prod(self.domain.shape_in)))
        self.scheme = Scheme(dico, check_inverse=check_inverse, need_validation=False)
        if self.domain.dim != self.scheme.dim:
            raise ValueError("the dimension of the domain and of the scheme are not the same")
        self.dim = self.domain.dim
        self._update_m = True
        self.t = 0.0
        self.nt = 0
        self.dt_ = self.domain.dx / self.scheme.la
        self.extra_parameters = {}
        self.storage = STORAGE_BY_DTYPE[dtype]
        self.dtype = dtype
        self.rho0 = float(rho0)
        self.spec = self.scheme.kernel_spec()
        self.module = load_module(self.spec)
        self.ns = self.spec.ns
        self.vel = _canon_velocities(self.spec, self.domain.box_label)
        self.params = rtm.LbmParams()
        for i, s in enumerate(self.spec.s):
            sv = float(s)
            self.params.omega[i] = sv
            self.params.one_minus[i] = 1.0 - sv
        self.params.dt = self.dt
        # coordinates of the cells (source terms that depend on X, Y, Z)
        for d, c in enumerate(self.domain.coords_halo):
            self.params.origin[d] = float(c[0])
        self.params.dx = float(self.domain.dx)
        self.params.xcd_remap = -1
        vmax = list(self.domain.stencil.vmax)
        # reference state of the FP32-compute formats: conserved moments
        # (rho0, 0, ...), its equilibrium populations w_k rho0
        for c in range(self.spec.nconsm):
            self.params.cref[c] = self.rho0 if c == 0 else 0.0
        wshift = self._rest_populations() if self.storage != rtm.LBM_F64 else np.zeros(self.ns)
        for k in range(self.ns):
            self.params.wshift[k] = wshift[k]
        self.wshift = wshift
        if layout not in ("interleaved", "planes"):
            raise ValueError("layout must be 'interleaved' or 'planes'")
        self.container = _Container(self.rt, self.ns, self.domain.shape_halo, vmax, self.storage,
                                    self.scheme.consm, wshift, interleaved=(layout == "interleaved"),
                                    vel=self.vel)
        self.bc = Boundary(self.domain, dico)
        for method in self.bc.methods:
            method.set_iload()
        self._dev_links = []
        # ghost slots: the periodic-face slabs are filled on the device and the other
        # slots any pull or Neumann-type list reads come from a host-computed copy list
        # (boundary.unwritten_ghost_reads); the full wrap (2) stays available
        self._ghost_mode = 1
        halo = np.asarray(self.domain.stencil.vmax)[:, None]
        shape = np.asarray(self.domain.shape_halo)[:, None]
        for m in self.bc.methods:
            if m.kind.bouzidi and any(np.any((np.asarray(r[1:]) < halo) | (np.asarray(r[1:]) >= shape - halo))
                                      for r in m.iload):
                # an interpolated list loads a halo slot (its second load, one link
                # beyond a fluid cell next to the box): wrap every halo slot, as pylbm
                # does (pylbm/storage.py:306-367)
                self._ghost_mode = 2
        # consecutive independent link lists share one launch in run() (False: one
        # launch per list; the results are the same)
        self.merge_links = True
        # bounce-back / anti-bounce-back lists folded into the fused step's pulls when
        # that is exact (False: every list runs as a link-list launch; same results)
        self.fold_links = True
        # small 2D grids on the wrap path advance several steps per launch in LDS
        # (run() only; same results, see _blocking_shape)
        self.temporal_blocking = True
        self.init_type = dico.get("inittype", "moments")
        self.init_data = dico.get("init", None)
        self._need_init = True
        if initialize:
            self._initialize()

    # ------------------------------------------------------------- basics
    @property
    def dt(self):
        if isinstance(self.dt_, sp.Expr):
            subs = list(self.scheme.param.items()) + list(self.extra_parameters.items())
            self.dt_ = float(self.dt_.subs(subs))
        return float(self.dt_)

    def _check(self, rc, what):
        self.rt.check(rc, what)

    def _rest_populations(self):
        """w_k rho0: m2f of the equilibrium with rho = rho0 and the other conserved moments 0."""
        m = DeviceArray(self.rt, self.ns, [1] * self.spec.dim, [0] * self.spec.dim, LBM_F64)
        host = np.zeros((self.ns,) + (1,) * self.spec.dim)
        host[0] = self.rho0
        m.set_fields(host)
        f = DeviceArray(self.rt, self.ns, [1] * self.spec.dim, [0] * self.spec.dim, LBM_F64)
        self._equilibrium(m)
        self._m2f(m, f)
        return f.fields().reshape(self.ns)

    # ---------------------------------------------------------- kernels
    def _equilibrium(self, m):
        self._check(self.module.lbm_gen_equilibrium(ctypes.c_void_p(m.ptr), ctypes.byref(m.layout),
                                                    self.rt.stream), "equilibrium")

    def _m2f(self, m, f):
        self._check(self.module.lbm_gen_m2f(ctypes.c_void_p(m.ptr), ctypes.byref(m.layout),
                                            ctypes.c_void_p(f.ptr), ctypes.byref(f.layout),
                                            ctypes.byref(self.params), self.rt.stream), "m2f")

    def _f2m(self, f, m):
        self._check(self.module.lbm_gen_f2m(ctypes.c_void_p(f.ptr), ctypes.byref(f.layout),
                                            ctypes.c_void_p(m.ptr), ctypes.byref(m.layout),
                                            ctypes.byref(self.params), self.rt.stream), "f2m")

    def equilibrium(self, m_user=None, **kwargs):
        _no_extra_arguments("equilibrium", kwargs)
        self._equilibrium(m_user if m_user is not None else self.container.m)

    def m2f(self, m_user=None, f_user=None, **kwargs):
        _no_extra_arguments("m2f", kwargs)
        self._m2f(m_user if m_user is not None else self.container.m,
                  f_user if f_user is not None else self.container.F)

    def f2m(self, **kwargs):
        _no_extra_arguments("f2m", kwargs)
        self._f2m(self.container.F, self.container.m)

    def transport(self, **kwargs):
        """pylbm's transport phase (pylbm/simulation.py:322-327, the generated routine of
        pylbm/algorithm/base.py:273-296): Fnew[k, x] = F[k, x - v_k] on the interior.
        F's ghost slots are read as they are, and F and Fnew are not exchanged."""
        _no_extra_arguments("transport", kwargs)
        c = self.container
        self._check(self.module.lbm_gen_transport(ctypes.c_void_p(c.F.ptr), ctypes.c_void_p(c.Fnew.ptr),
                                                  ctypes.byref(c.F.layout), ctypes.byref(self.params),
                                                  self.rt.stream), "transport")

    def relaxation(self, **kwargs):
        """pylbm's relaxation phase on the moment array m (pylbm/simulation.py:329-334,
        pylbm/algorithm/base.py:404-437): m_i <- (1 - s_i) m_i + s_i eq_i on every cell."""
        _no_extra_arguments("relaxation", kwargs)
        m = self.container.m
        self._check(self.module.lbm_gen_relaxation(ctypes.c_void_p(m.ptr), ctypes.byref(m.layout),
                                                   ctypes.byref(self.params), self.rt.stream), "relaxation")

    def source_term(self, fraction_of_time_step=1.0, **kwargs):
        """pylbm's source-term phase on the moment array m (pylbm/simulation.py:336-342,
        pylbm/algorithm/base.py:439-504 with ode.py:11-16): one explicit Euler half step
        m_c <- m_c + dt/2 S(m, X, Y, Z, t) on the interior at the current t.  As in
        pylbm, fraction_of_time_step is accepted and does not change the step."""
        _no_extra_arguments("source_term", kwargs)
        m = self.container.m
        self.params.t = self.t
        self._check(self.module.lbm_gen_source_term(ctypes.c_void_p(m.ptr), ctypes.byref(m.layout),
                                                    ctypes.byref(self.params), self.rt.stream), "source_term")

    # ---------------------------------------------------------- init
    def _initialize(self):
        self.initialization()
        for method in self.bc.methods:
            self._prepare_rhs(method)
            method.fix_iload()
            method.set_rhs()
        # one device list per method; an interpolated method whose loop reads values
        # stored earlier in the same loop runs as several lists (bouzidi_levels)
        layout = self.container.F.layout
        self._method_devs = []
        for m in self.bc.methods:
            nlev = len(bouzidi_levels(m, layout, self.wshift)) if getattr(m.kind, "bouzidi", False) else 1
            self._method_devs.append((m, [DeviceLinks(self.rt, m, layout, self.wshift, level=k)
                                          for k in range(nlev)]))
        self._dev_links = [d for _, devs in self._method_devs for d in devs]
        if self._ghost_mode == 1:
            st, ld = unwritten_ghost_reads(self.domain, self.scheme.stencil, self.bc.methods,
                                           self.container.F.layout)
            if st.size:
                # before the link lists: pylbm's wrap happens before its boundary methods
                self._dev_links.insert(0, DeviceLinks(self.rt, "ghost_copy", self.container.F.layout,
                                                      self.wshift, flat=(st, ld)))
        for d in self._dev_links:
            d.fused = False
        for name in ("bbmask", "bbtab_bits", "bbtab_base", "bbtab_src", "bbtab_sign", "bbtab_rhs", "bbtab_rel",
                     "bbtab_src2", "bbtab_c2", "bbtab_rel2"):
            setattr(self.params, name, None)
        if self.fold_links:
            self._fold_links()
        active = [d for d in self._dev_links if not d.fused]
        self._active_links = active
        self._link_sets = (rtm.LbmLinkSet * max(1, len(active)))(*[d.link_set() for d in active])
        # independent consecutive lists run in one launch (bit-identical results)
        flags = merge_flags([(d.h_store, d.h_load, d.conflict) for d in active])
        for ls, flag in zip(self._link_sets, flags):
            ls.merge_prev = flag if self.merge_links else 0
        # Wrap mode: when every boundary list is folded into the step, the step may read
        # every pull through the periodic wrap of the interior instead of the halo.  A
        # ghost slot the step would read holds the wrapped interior value (pylbm wraps
        # the whole halo, pylbm/storage.py:306-367; the device fill and the ghost copy
        # list write the same values), unless a link stores it -- and every link is
        # folded, so those pulls take the folded value from the mask / table instead.
        # So the step needs neither the ghost fill nor the ghost copy list: one launch
        # per step, no halo line fetched, the same bits.
        self._wrap = self.periodic_only or (bool(self.fold_links) and bool(self.bc.methods) and all(
            d.fused or isinstance(d.method, str) for d in self._dev_links))
        for d in self._dev_links:
            d.h_store = d.h_load = d.h_sign = d.h_rhs = None
        self._need_init = False

    def _fold_links(self):
        """Link lists folded into the fused step's pulls.

        Within the interior update a stored link slot (k, y) is read only by the
        pull of population k at x = y + v_k, so the step may compute the value the
        lists would have left there instead of reading it.  First the whole step
        is tried (_fold_step); failing that, single bounce-back / anti-bounce-back
        lists (_fold_lists).  Folded lists leave the step's link sets; the slots
        they would have written, in the halo and in solid cells, keep their
        previous values (the interior update never reads them)."""
        if self.ns > 31:
            return   # bit 31 of a mask word flags the cells with table entries
        if not self._fold_step():
            self._fold_lists()

    def _fold_layout(self):
        L = self.container.F.layout
        n = np.array([L.n[a] for a in range(3)], dtype=np.int64)
        h = np.array([L.halo[a] for a in range(3)], dtype=np.int64)
        return L, n, h, n - 2 * h

    def _fold_step(self):
        """Every list of the step, resolved in order into one value per stored slot
        (boundary.resolve_step): sign * f[origin] + rhs of a slot no list stores, or
        for an interpolated (Bouzidi) link ca * f[origin] + cb * f[origin2] + rhs.
        That reproduces the launches' result exactly when no chain applies two rhs
        and no link reads an interpolated value or interpolates a value with an rhs
        (a copy or a sign flip does not round); otherwise, with a time-dependent
        list, a two-phase interpolated list, or when more than FOLD_TABLE_MAX links
        would need a table entry, nothing is folded here."""
        devs = self._dev_links
        if not devs or any(not isinstance(d.method, str) and d.method._tbc for d in devs):
            return False
        L, n, h, nin = self._fold_layout()
        lists = []
        for d in devs:
            if not d.bouzidi:
                lists.append((d.h_store, d.h_load, d.h_sign, d.h_rhs))
                continue
            # an interpolated list: two sources per link, one phase only
            st, l0, l1, ca, cb, rhs, two, index, shift = d.bouzidi_form(L, self.wshift)
            if two:
                return False
            lists.append((st, l0, ca, rhs if d.d_rhs is not None else None, l1, cb))
        slot, org, sg, rh, has, bad, org2, c2 = resolve_step(lists)
        k, o, m, i = _unflat_fn(slot, L)
        vel = np.array([[self.vel.v[q][a] for a in range(3)] for q in range(self.ns)], dtype=np.int64)
        x = np.stack([o, m, i]) + vel[k].T
        inside = np.all((x >= h[:, None]) & (x < (n - h)[:, None]), axis=0)
        if np.any(bad & inside):
            return False
        k, x, org, sg, rh, has = k[inside], x[:, inside], org[inside], sg[inside], rh[inside], has[inside]
        org2, c2 = org2[inside], c2[inside]
        xi = x - h[:, None]
        cell = (xi[0] * nin[1] + xi[1]) * nin[2] + xi[2]
        ksym = np.asarray(self.scheme.stencil.get_symmetric())[k]
        own = ksym * L.pop_stride + (x[0] * L.n[1] + x[1]) * L.row_stride + L.lead + x[2]
        copy = (org == own) & (sg == 1.0) & ~has & (org2 < 0)
        if np.count_nonzero(~copy) > FOLD_TABLE_MAX:
            return False
        mask = np.zeros(int(np.prod(nin)), dtype=np.uint32)
        np.bitwise_or.at(mask, cell, np.uint32(1) << k.astype(np.uint32))
        t = ~copy
        self._upload_fold(mask, (cell[t], k[t], org[t], sg[t], np.where(has[t], rh[t], -0.0), org2[t], c2[t]),
                          devs[0])
        for d in devs:
            d.fused = True
        return True

    def _fold_lists(self):
        """Single bounce-back / anti-bounce-back lists: every link is
        f[k, x - v_k] = sign f[ksym, x] + value with x interior, no other folded link
        targets the same (k, x), and no later list of the step loads or stores the
        list's stored slots or stores its loaded slots (earlier lists run before
        the pull either way).  Interpolated (Bouzidi) lists the same way with a
        second source: f[k, x - v_k] = ca f[ksym, x] + cb f[load1] + value, when the
        list runs one-phase and no list folded before it stores a slot it loads
        (those stores are never written)."""
        devs = self._dev_links
        layout = self.container.F.layout
        sten = self.scheme.stencil
        ksym_all = np.asarray(sten.get_symmetric())
        v = np.asarray(sten.get_all_velocities())
        halo = np.asarray(self.domain.stencil.vmax)
        shape_in = tuple(int(n) for n in self.domain.shape_in)
        mask = np.zeros(int(np.prod(shape_in)), dtype=np.uint32)
        tab = []
        owner = None
        folded_stores = []
        for j, d in enumerate(devs):
            m = d.method
            if (not isinstance(m, str) and getattr(m.kind, "bouzidi", False) and not m._tbc and not d.conflict
                    and d.n > 0):
                entry = self._fold_bouzidi(d, mask, devs[:j], devs[j + 1:])
                if entry is not None:
                    tab.append(entry)
                    folded_stores.append(d.h_store)
                    d.fused = True
                    owner = owner or d
                continue
            if (isinstance(m, str) or m.kind.name not in ("bounce_back", "anti_bounce_back") or m._tbc
                    or d.conflict or d.n == 0):
                continue
            st, ld, sign, rhs, conflict, order = host_lists(m, layout, self.wshift)
            rows_s = np.asarray(m._store_rows())
            rows_l = np.asarray(m.iload[0].T if m._fixed else m.iload[0])
            if order is not None:
                rows_s, rows_l = rows_s[:, order], rows_l[:, order]
            k = rows_s[0]
            x = rows_l[1:]
            if not (np.array_equal(rows_l[0], ksym_all[k]) and np.array_equal(x, rows_s[1:] + v[k].T)):
                continue
            xi = x - halo[:, None]
            if np.any(xi < 0) or np.any(xi >= np.asarray(shape_in)[:, None]):
                continue
            cell = np.ravel_multi_index(tuple(xi), shape_in)
            bit = np.uint32(1) << k.astype(np.uint32)
            if np.any(mask[cell] & bit):
                continue
            if any(np.isin(o.h_store, d.h_store).any() or np.isin(o.h_load, d.h_store).any()
                   or np.isin(o.h_store, d.h_load).any() for o in devs[j + 1:]):
                continue
            if any(np.isin(ld, fs).any() for fs in folded_stores):
                continue
            np.bitwise_or.at(mask, cell, bit)
            if d.h_rhs is not None or d.d_sign is not None or d.sign != 1.0:
                # the link kernel adds rhs only when the list has one: -0.0 adds nothing
                tab.append((cell, k, ld, sign, rhs if d.h_rhs is not None else np.full(d.n, -0.0),
                            np.full(d.n, -1, dtype=np.int64), np.zeros(d.n)))
            folded_stores.append(d.h_store)
            d.fused = True
            owner = owner or d
        if owner is not None:
            cat = tuple(np.concatenate([t[q] for t in tab]) for q in range(7)) if tab else None
            self._upload_fold(mask, cat, owner)

    def _fold_bouzidi(self, d, mask, earlier, later):
        """The table entries (cell, k, load0, ca, rhs, load1, cb) of one interpolated
        list folded into the pulls, its mask bits set; None (mask untouched) when the
        list does not fold: a link whose first load is not the pulling cell's own
        ksym population or whose pulling cell is not interior, a direction already
        folded at a cell, a later list touching its slots, or a load of a slot that
        an earlier list stores (the ghost copy aside: it stores the periodic wrap the
        folded loads read anyway)."""
        layout = self.container.F.layout
        st, l0, l1, ca, cb, rhs, two, index, shift = d.bouzidi_form(layout, self.wshift)
        if two or st.size == 0:
            return None
        L, n, h, nin = self._fold_layout()
        k, o, m_, i = _unflat_fn(st, L)
        vel = np.array([[self.vel.v[q][a] for a in range(3)] for q in range(self.ns)], dtype=np.int64)
        x = np.stack([o, m_, i]) + vel[k].T
        if np.any(x < h[:, None]) or np.any(x >= (n - h)[:, None]):
            return None
        ksym = np.asarray(self.scheme.stencil.get_symmetric())[k]
        own = ksym * L.pop_stride + (x[0] * L.n[1] + x[1]) * L.row_stride + L.lead + x[2]
        if not np.array_equal(l0, own):
            return None
        xi = x - h[:, None]
        cell = (xi[0] * nin[1] + xi[1]) * nin[2] + xi[2]
        bit = np.uint32(1) << k.astype(np.uint32)
        if np.any(mask[cell] & bit) or np.unique(cell * 64 + k).size != cell.size:
            return None
        loads = np.concatenate([l0, l1])
        if any(np.isin(o_.h_store, st).any() or np.isin(o_.h_load, st).any() or np.isin(o_.h_store, loads).any()
               for o_ in later):
            return None
        if any(not isinstance(o_.method, str) and np.isin(o_.h_store, loads).any() for o_ in earlier):
            return None
        if np.isin(loads, st).any():
            return None
        np.bitwise_or.at(mask, cell, bit)
        value = (np.asarray(d.method.rhs, dtype=np.float64)[index] + shift) if d.d_rhs is not None \
            else np.full(st.size, -0.0)
        return cell, k, l0, ca, value, l1, cb

    def _wrap_ghost_sources(self, src):
        """Table sources that are ghost slots, moved to the fully wrapped interior slot.

        A source is a slot no link list stores (boundary.resolve_step), so a ghost
        source holds the wrapped interior value at step time (the periodic-face fill
        writes it there, as pylbm's wrap of the whole halo does): reading the
        interior slot gives the same bits and lets the step skip the ghost fill."""
        L, n, h, nin = self._fold_layout()
        k, o, m, i = _unflat_fn(np.asarray(src, dtype=np.int64), L)
        c = [np.asarray(o), np.asarray(m), np.asarray(i)]
        ghost = np.zeros(c[0].shape, dtype=bool)
        for a in range(3):
            ghost |= (c[a] < h[a]) | (c[a] >= n[a] - h[a])
            c[a] = np.where(c[a] < h[a], c[a] + nin[a], np.where(c[a] >= n[a] - h[a], c[a] - nin[a], c[a]))
        if not ghost.any():
            return src
        wrapped = k * L.pop_stride + (c[0] * L.n[1] + c[1]) * L.row_stride + L.lead + c[2]
        return np.where(ghost, wrapped, src)

    def _relative_sources(self, cell, src):
        """Table sources relative to their pulling cell, for the temporally blocked run
        (2D): population | (mid offset + 128) << 8 | (inner offset + 128) << 16, the
        offsets taken through the periodic wrap; None when a source is farther than
        one cell away or the scheme is not 2D."""
        if self.dim != 2 or src.size == 0:
            return None
        L, n, h, nin = self._fold_layout()
        k, o, m, i = _unflat_fn(np.asarray(src, dtype=np.int64), L)
        cm, ci = np.asarray(cell) // nin[2], np.asarray(cell) % nin[2]
        dm = (np.asarray(m) - h[1] - cm + nin[1] // 2) % nin[1] - nin[1] // 2
        di = (np.asarray(i) - h[2] - ci + nin[2] // 2) % nin[2] - nin[2] // 2
        if np.any(np.abs(dm) > 1) or np.any(np.abs(di) > 1) or np.any(np.asarray(o) != h[0]):
            return None
        return (np.asarray(k, np.uint32) | ((dm + 128).astype(np.uint32) << 8)
                | ((di + 128).astype(np.uint32) << 16)).astype(np.uint32)

    def _upload_fold(self, mask, tab, owner):
        """Device copies of the fold mask and table (cell, k, src, sign, rhs) -> params."""
        up = []
        if tab is not None and tab[0].size:
            cell, kk, src, sg, rh, src2, c2 = tab
            src = self._wrap_ghost_sources(src)
            two = src2 >= 0
            if two.any():
                src2 = np.where(two, self._wrap_ghost_sources(np.where(two, src2, src)), -1)
            o = np.lexsort((kk, cell))
            cell, kk, src, sg, rh, src2, c2, two = cell[o], kk[o], src[o], sg[o], rh[o], src2[o], c2[o], two[o]
            bits = np.zeros(mask.size, dtype=np.uint32)
            np.bitwise_or.at(bits, cell, np.uint32(1) << kk.astype(np.uint32))
            base = np.zeros(mask.size, dtype=np.uint32)
            first_cell, first = np.unique(cell, return_index=True)
            base[first_cell] = first.astype(np.uint32)
            mask = mask.copy()
            mask[first_cell] |= np.uint32(1 << 31)   # these cells read their table word
            ctype = np.float64 if self.storage == LBM_F64 else np.float32
            arrays = [("bbtab_bits", bits), ("bbtab_base", base), ("bbtab_src", src.astype(np.uint32)),
                      ("bbtab_sign", sg.astype(ctype)), ("bbtab_rhs", rh.astype(ctype))]
            rel = self._relative_sources(cell, src)
            if two.any():
                arrays += [("bbtab_src2", np.where(two, src2, 0xFFFFFFFF).astype(np.uint32)),
                           ("bbtab_c2", c2.astype(ctype))]
                rel2 = self._relative_sources(cell, np.where(two, src2, src))
                if rel is not None and rel2 is not None:
                    # bit 31 marks the entries with a second source (the blocked run
                    # reads no bbtab_src2); the others repeat their first source
                    rel2 = rel2 | np.where(two, np.uint32(1 << 31), np.uint32(0)).astype(np.uint32)
                    arrays.append(("bbtab_rel2", rel2))
                else:
                    rel = None
            if rel is not None:
                arrays.append(("bbtab_rel", rel))
            for name, arr in arrays:
                ptr = self.rt.upload(np.ascontiguousarray(arr))
                setattr(self.params, name, ptr)
                up.append(ptr)
        self.params.bbmask = self.rt.upload(mask)
        up.append(self.params.bbmask)
        owner._ptrs.extend(up)   # freed with the list's other device arrays

    def initialization(self):
        coords = np.meshgrid(*self.domain.coords_halo, sparse=True, indexing="ij")
        if self.init_type not in ("moments", "distributions"):
            raise ValueError("inittype must be 'moments' or 'distributions'")
        nc = self.scheme.nconsm if self.init_type == "moments" else self.ns
        host = np.zeros((nc,) + tuple(self.domain.shape_halo))
        view = _MomentView(host, self.scheme.consm)
        for k, v in (self.init_data or {}).items():
            if isinstance(v, tuple):
                args = tuple(coords) + (tuple(v[1]) if len(v) == 2 else ())
                view[k] = v[0](*args)
            elif isinstance(v, types.FunctionType) or callable(v):
                view[k] = v(*coords)
            else:
                view[k] = v
        if self.init_type == "moments":
            m = self.container.m
            m.set_fields(host, 0)
            self._equilibrium(m)
            self._m2f(m, self.container.F)
        else:
            self.container.F.set_fields(host, 0)
        self.container.Fnew.copy_from(self.container.F)
        self.rt.sync()

    def _prepare_rhs(self, method):
        dx = self.domain.dx
        v = self.scheme.stencil.get_all_velocities()
        for key, value in method.value_bc.items():
            if value is None:
                continue
            idx = np.where(method.ilabel == key)
            n = idx[0].size
            k = method.istore[0, idx]
            s = 1 - method.distance[idx]
            coords = ()
            for i in range(self.dim):
                x = self.domain.coords_halo[i][method.istore[i + 1, idx]]
                x += s * v[k, i] * dx
                x = x.ravel()
                for _ in range(1, self.dim):
                    x = x[:, np.newaxis]
                coords += (x,)
            func, args = (value[0], tuple(value[1])) if isinstance(value, tuple) else (value, ())
            shape = [n] + [1] * (self.dim - 1)
            host = np.zeros((self.ns,) + tuple(shape))
            mv = _MomentView(host, self.scheme.consm)
            if method.time_bc[key]:
                func(None, mv, 0.0, *(coords + args))
                method._tbc.append((key, func, coords + args, idx[0], shape))
            else:
                func(None, mv, *(coords + args))
            method.feq[:, idx[0]] = self._feq(host, shape)

    def _feq(self, host, shape):
        m = DeviceArray(self.rt, self.ns, shape, [0] * len(shape), LBM_F64)
        f = DeviceArray(self.rt, self.ns, shape, [0] * len(shape), LBM_F64)
        m.set_fields(host)
        self._equilibrium(m)
        self._m2f(m, f)
        return f.fields().reshape(self.ns, -1)

    def _update_feq(self, method, devs):
        """Time-dependent boundary values (pylbm/boundary.py:307-321) into the
        method's device lists."""
        if not method._tbc:
            return
        for key, func, args, cols, shape in method._tbc:
            host = np.zeros((self.ns,) + tuple(shape))
            func(None, _MomentView(host, self.scheme.consm), self.t, *args)
            method.feq[:, cols] = self._feq(host, shape)
        method.set_rhs()
        for dev in devs:
            dev.update_rhs(method.rhs, self.wshift)

    # ---------------------------------------------------------- stepping
    def boundary_condition(self, **kwargs):
        _no_extra_arguments("boundary_condition", kwargs)
        if self._wrap:
            return   # every list folded into the step, which reads through the wrap
        f = self.container.F
        fn = self.rt.lib.lbm_ghost_fill if self._ghost_mode == 1 else self.rt.lib.lbm_ghost_fill_all
        self._check(fn(ctypes.c_void_p(f.ptr), ctypes.byref(f.layout), ctypes.byref(self.vel),
                       self.rt.stream), "ghost fill")
        devs = self._dev_links
        if devs and isinstance(devs[0].method, str) and not devs[0].fused:
            devs[0].apply(f.ptr)   # the ghost copy list, unless folded into the step
        for method, mdevs in self._method_devs:
            self._update_feq(method, mdevs)
            for dev in mdevs:
                if not dev.fused:
                    dev.apply(f.ptr)

    def _fused(self):
        F, Fn = self.container.F, self.container.Fnew
        self.params.t = self.t
        self._check(self.module.lbm_gen_step(ctypes.c_void_p(F.ptr), ctypes.c_void_p(Fn.ptr),
                                             ctypes.byref(F.layout), ctypes.byref(self.params),
                                             1 if self._wrap else 0, self.rt.stream), "one_time_step")

    def one_time_step(self, **kwargs):
        _no_extra_arguments("one_time_step", kwargs)
        if self._need_init:
            self._initialize()
        t0 = time.perf_counter() if Monitor.enabled else 0.0
        self._update_m = True
        self.boundary_condition()
        self._fused()
        c = self.container
        c.F, c.Fnew = c.Fnew, c.F
        self.t += self.dt
        self.nt += 1
        if Monitor.enabled:
            self.rt.sync()
            Monitor.record("one_time_step", time.perf_counter() - t0)

    @property
    def periodic_only(self):
        return not self.bc.methods and all(lab == -1 for lab in self.domain.box_label)

    def run(self, nsteps, fold=None):
        """nsteps calls of one_time_step in one library call (same results).

        fold=None reads the pulls through the periodic wrap whenever that gives the
        same result (a periodic box, or every link list folded into the step);
        fold=False keeps the halo path (ghost fill and link launches); fold=True
        requires the wrap."""
        if self._need_init:
            self._initialize()
        if nsteps <= 0:
            return
        if any(m._tbc for m in self.bc.methods):
            for _ in range(nsteps):
                self.one_time_step()
            return
        self._update_m = True
        fold = self._wrap if fold is None else fold
        if fold and not self._wrap:
            raise ValueError("the periodic wrap needs a periodic box or every link list folded into the step")
        c = self.container
        self.params.t = self.t
        parity = ctypes.c_int(0)
        tb = self._blocking_shape() if fold else None
        if tb is not None and nsteps >= tb[0]:
            nt, tm, ti = tb
            self._check(self.module.lbm_gen_run_blocked(ctypes.c_void_p(c.F.ptr), ctypes.c_void_p(c.Fnew.ptr),
                                                        ctypes.byref(c.F.layout), ctypes.byref(self.params),
                                                        int(nsteps), nt, tm, ti, ctypes.byref(parity),
                                                        self.rt.stream), "run (blocked)")
            if parity.value:
                c.F, c.Fnew = c.Fnew, c.F
            self.nt += nsteps
            self.t = self.t + nsteps * self.dt
            return
        ghost = 0 if fold else self._ghost_mode
        nmeth = 0 if fold else len(self._active_links)
        self._check(self.module.lbm_gen_run(ctypes.c_void_p(c.F.ptr), ctypes.c_void_p(c.Fnew.ptr),
                                            ctypes.byref(c.F.layout), ctypes.byref(self.vel),
                                            ctypes.byref(self.params), self._link_sets,
                                            nmeth, int(nsteps), ghost, 1 if fold else 0,
                                            ctypes.byref(parity), self.rt.stream), "run")
        if parity.value:
            c.F, c.Fnew = c.Fnew, c.F
        self.nt += nsteps
        self.t = self.t + nsteps * self.dt

    def _blocking_shape(self):
        """(steps per launch, tile rows, tile columns) of the temporally blocked run, or
        None: 2D schemes on the wrap path (periodic, or every list folded with its
        table sources within one cell) whose grid is small enough that a step does not
        fill the GPU (at most 2^21 cells).  The tile and its halo live in LDS: two
        copies of ns populations per cell in planes of a fixed size (the most cells
        that fit the 160 KiB of a CU with the mask and table words), and the region
        has to fit one plane.  The preferred shapes were measured on MI355X
        (D2Q9, 128^2 to 1024 x 256): 6 steps of 32 x 32 tiles for the 4- and 2-byte
        formats, 4 steps of 16 x 32 tiles for FP64, 6 steps of 8 x 16 tiles for FP64
        grids of at most 2^15 cells (more tiles than CUs matter more there than the
        recomputed halo); larger schemes shrink the tile until it fits."""
        if not self.temporal_blocking or self.dim != 2 or not self._wrap:
            return None
        if np.any(np.abs(np.asarray(self.scheme.stencil.get_all_velocities())) > 1):
            return None
        if (self.params.bbtab_bits and not self.params.bbtab_rel) or (self.params.bbtab_src2 and
                                                                      not self.params.bbtab_rel2):
            return None
        nin = [int(x) for x in self.domain.shape_in]
        cells = nin[0] * nin[1]
        if cells > (1 << 21):
            return None
        es = self.container.F.layout.elem_size
        # the kernel's population planes hold this many cells (tb_plane in the
        # generated module): every region has to fit one plane
        plane = min((160 * 1024) // (2 * self.ns * es + 12),
                    65528 // ((self.ns - 1) * es) if self.ns > 1 else 1 << 30)
        if es == 8:
            nt, tm, ti = (6, 8, 16) if cells <= (1 << 15) else (4, 16, 32)
        else:
            nt, tm, ti = 6, 32, 32
        while (tm + 2 * nt) * (ti + 2 * nt) > plane:
            if ti > 16:
                ti //= 2
            elif tm > 8:
                tm //= 2
            elif nt > 2:
                nt -= 1
            else:
                return None
        nt = min(nt, min(nin))
        if nt < 2:
            return None
        return nt, min(tm, nin[0]), min(ti, nin[1])

    def sync(self):
        self.rt.sync()

    # ---------------------------------------------------------- read-back
    def _moments(self):
        if self._update_m:
            self._update_m = False
            self.f2m()

    @itemproperty
    def m_halo(self, i):
        self._moments()
        return self.container.m[i]

    @itemproperty
    def m(self, i):
        self._moments()
        return self.container.m._in(i)

    @m_halo.setter
    def m_halo(self, i, value):
        self._update_m = False
        self.container.m[i] = value

    @itemproperty
    def F_halo(self, i):
        return self.container.F[i]

    @F_halo.setter
    def F_halo(self, i, value):
        self._update_m = True
        self.container.F[i] = value

    @itemproperty
    def F(self, i):
        return self.container.F._in(i)

    def moments(self):
        """All moments on the interior, FP64, shape (ns, *shape_in)."""
        self._moments()
        full = self.container.m.fields()
        sl = (slice(None),) + tuple(slice(h, -h) if h > 0 else slice(None) for h in self.domain.stencil.vmax)
        return full[sl]

    def conserved_sums(self):
        """Global sums of the conserved moments over the interior: FP64 compensated
        per-population sums on the GPU, combined exactly with the rows of M."""
        f = self.container.F
        nb = ctypes.c_int64(0)
        self._check(self.rt.lib.lbm_reduce_blocks(ctypes.byref(f.layout), ctypes.byref(nb)), "reduce")
        part = self.rt.malloc(8 * 2 * self.ns * (nb.value + 1))
        wptr = self.rt.upload(self.wshift.astype(np.float64)) if self.storage in rtm.DEVIATION_FORMATS else None
        s = (ctypes.c_double * self.ns)()
        c = (ctypes.c_double * self.ns)()
        try:
            self._check(self.rt.lib.lbm_reduce_populations(ctypes.c_void_p(f.ptr), ctypes.byref(f.layout),
                                                           ctypes.c_void_p(wptr), ctypes.c_void_p(part),
                                                           s, c, self.rt.stream), "reduce")
        finally:
            self.rt.free(part)
            if wptr:
                self.rt.free(wptr)
        out = []
        for i in range(self.scheme.nconsm):
            terms = []
            for k in range(self.ns):
                coef = self.spec.M[i][k]
                if coef != 0:
                    terms += [float(coef) * s[k], float(coef) * c[k]]
            out.append(math.fsum(terms))
        return out

    def __str__(self):
        return ("Simulation(hip, dim=%d, shape=%s, ns=%d, storage=%s, t=%g, nt=%d, device=%s)"
                % (self.dim, self.domain.shape_in, self.ns, self.dtype, self.t, self.nt, self.rt.device_name))

    __repr__ = __str__
