"""numpy FP64 driver for one scheme -- oracle restatement of the hot path.

Follows:
  * Simulation.initialization: init dict on the halo-inclusive meshgrid into
    the conserved moments, then equilibrium + m2f      pylbm/simulation.py:257-320
  * one_time_step = ghost fill, link methods in order,
    fused pull step, swap, t += dt                     pylbm/simulation.py:373-420
  * Array.update: periodic ghost wrap, axis by axis,
    every population, full halo-inclusive extent       pylbm/storage.py:306-367
  * pull step: f_k(x - v_k) -> m = M f -> [source dt/2]
    -> relax m <- (1-s) m + s eq -> [source dt/2] ->
    fnew = M^-1 m on the interior                      pylbm/algorithm/pull.py:11-59,
                                                       pylbm/algorithm/base.py:298-600
  * source half steps m_c <- m_c + dt/2 S(m)           pylbm/algorithm/ode.py:11-16,
                                                       pylbm/algorithm/base.py:439-495
  * read-back m = M f on the full domain, interior
    view                                               pylbm/simulation.py:215-224
  * time-dependent boundary values ("time_bc"): the
    value function is called with t = 0 at set-up and
    with the current t before every step's links,
    then equilibrium + m2f -> feq -> rhs               pylbm/boundary.py:287-321,
                                                       pylbm/simulation.py:384-390

Pinned by the known answers in tests/test_oracle_lbm.py: the rest state is a
fixed point; with every non-conserved rate 0 a step is the exact permutation
f_k(x, t+1) = f_k(x - v_k, t) on a periodic box; with every rate 1 the
post-collision state is M^-1 EQ(m_c); mass and momentum are conserved to
round-off on periodic boxes (math.fsum); bounce-back on resting walls
conserves mass; a time-dependent value whose function ignores t gives the
static value's run bit for bit, and one that switches at t0 gives the static
runs' states before t0.  tests/test_oracle_pins.py adds closed forms for the
source half-steps (a body force raises the total momentum by exactly
dt g sum(rho) per step; a uniform logistic source follows two Euler half-steps),
the valued bounce-back / anti-bounce-back right-hand sides (2 w_k c_k.j / c_s^2
and the even second-order equilibrium) and plane Couette flow.  End to end, the
driver reproduces the reference's own fixtures (tests/reference/*.h5 of pylbm,
converted under tests/golden/h5) at the reference's tolerance
(tests/test_reference_fixtures.py).  The C1-C5 trajectories themselves have no
reference fixture; they rest on those checks.
"""
import math

import numpy as np

from .boundary import (ABB, BB, BZ_ABB, BZ_BB, NEUMANN, NEUMANN_X, NEUMANN_Y, NEUMANN_Z, apply_links,
                       build_links)
from .domain import Domain
from .scheme import Scheme

_KINDS = {"BounceBack": BB, "AntiBounceBack": ABB, "Neumann": NEUMANN,
          "NeumannX": NEUMANN_X, "NeumannY": NEUMANN_Y, "NeumannZ": NEUMANN_Z,
          "BouzidiBounceBack": BZ_BB, "BouzidiAntiBounceBack": BZ_ABB}


def method_kind(cls):
    name = cls if isinstance(cls, str) else cls.__name__
    return _KINDS[name]


class _Moments:
    """Minimal moment container for user value/init functions (symbol keys)."""

    def __init__(self, arr, consm):
        self.arr = arr
        self.consm = consm

    def _key(self, key):
        return self.consm[key] if not isinstance(key, (int, np.integer)) else key

    def __getitem__(self, key):
        return self.arr[self._key(key)]

    def __setitem__(self, key, value):
        self.arr[self._key(key)] = value


class Simulation:
    def __init__(self, dico):
        self.dico = dico
        self.domain = Domain(dico)
        self.scheme = Scheme(dico)
        st = self.scheme.stencil
        self.dim = self.domain.dim
        self.ns = self.scheme.ns
        self.v = st.all_velocities()
        self.h = np.asarray(st.vmax)
        self.dx = self.domain.dx
        self.dt = self.dx / self.scheme.la
        self.t = 0.0
        self.nt = 0
        shape = [self.ns] + list(self.domain.shape_halo)
        self.F = np.zeros(shape)
        self.Fnew = np.zeros(shape)
        self.methods = build_links(self.domain, st, dico.get("boundary_conditions", {}), method_kind)
        self._initialize()
        for m in self.methods:
            self._prepare_rhs(m)

    # ------------------------------------------------------------ helpers
    def _eq_full(self, m):
        nc = self.scheme.nconsm
        return np.stack(self.scheme.equilibrium([m[i] for i in range(nc)]))

    def _m2f(self, m):
        return np.tensordot(self.scheme.invMnum, m, axes=(1, 0))

    def _f2m(self, f):
        return np.tensordot(self.scheme.Mnum, f, axes=(1, 0))

    def _initialize(self):
        coords = np.meshgrid(*self.domain.coords_halo, sparse=True, indexing="ij")
        m = np.zeros_like(self.F)
        mv = _Moments(m, self.scheme.consm)
        init = self.dico.get("init", {})
        for k, v in (init.items() if self.dico.get("inittype", "moments") == "moments" else ()):
            if isinstance(v, tuple):
                mv[k] = v[0](*(tuple(coords) + (tuple(v[1]) if len(v) == 2 else ())))
            elif callable(v):
                mv[k] = v(*coords)
            else:
                mv[k] = v
        if self.dico.get("inittype", "moments") == "moments":
            m = self._eq_full(m)
            self.F[...] = self._m2f(m)
        else:
            # the populations themselves, keyed by their index (simulation.py:302-312)
            for k, v in init.items():
                if isinstance(v, tuple):
                    val = v[0](*(tuple(coords) + (tuple(v[1]) if len(v) == 2 else ())))
                elif callable(v):
                    val = v(*coords)
                else:
                    val = v
                self.F[int(k)] = val
        self.Fnew[...] = self.F

    def _prepare_rhs(self, meth, t=None):
        """feq and rhs of one link method; t=None: set-up (every value, time-dependent
        ones at t = 0), else only the time-dependent values, at time t."""
        bcd = self.dico["boundary_conditions"]
        if t is None:
            meth.feq = np.zeros((self.ns, meth.istore.shape[1]))
            meth.timed = False
        for key in np.unique(meth.ilabel):
            value = bcd[key].get("value")
            if value is None:
                continue
            timed = bool(bcd[key].get("time_bc", False))
            meth.timed = meth.timed or timed
            if t is not None and not timed:
                continue
            idx = np.where(meth.ilabel == key)
            n = idx[0].size
            k = meth.istore[0, idx]
            s = 1 - meth.distance[idx]
            coords = ()
            for i in range(self.dim):
                x = self.domain.coords_halo[i][meth.istore[i + 1, idx]]
                x += s * self.v[k, i] * self.dx
                x = x.ravel()
                for _ in range(1, self.dim):
                    x = x[:, np.newaxis]
                coords += (x,)
            shp = [self.ns, n] + [1] * (self.dim - 1)
            m = np.zeros(shp)
            f = np.zeros(shp)
            func, args = (value[0], tuple(value[1])) if isinstance(value, tuple) else (value, ())
            if timed:
                func(f, _Moments(m, self.scheme.consm), 0.0 if t is None else t, *(coords + args))
            else:
                func(f, _Moments(m, self.scheme.consm), *(coords + args))
            m = self._eq_full(m)
            f = self._m2f(m)
            meth.feq[:, idx[0]] = f.reshape(self.ns, n)
        k = meth.istore[0]
        if meth.kind in (BB, ABB, BZ_BB, BZ_ABB):
            ksym = self.scheme.stencil.symmetric()[k]
            cols = np.arange(k.size)
            sign = -1.0 if meth.kind in (BB, BZ_BB) else 1.0
            meth.rhs = meth.feq[k, cols] + sign * meth.feq[ksym, cols]

    # ------------------------------------------------------------ one step
    def ghost_fill(self, f):
        for d in range(self.dim):
            h = int(self.h[d])
            if h == 0:
                continue
            n = f.shape[d + 1]
            lo = [slice(None)] * f.ndim
            src = [slice(None)] * f.ndim
            lo[d + 1] = slice(0, h)
            src[d + 1] = slice(n - 2 * h, n - h)
            f[tuple(lo)] = f[tuple(src)]
            lo[d + 1] = slice(n - h, n)
            src[d + 1] = slice(h, 2 * h)
            f[tuple(lo)] = f[tuple(src)]

    def _interior(self, shift=None):
        sl = [slice(None)]
        for d in range(self.dim):
            h = int(self.h[d])
            n = self.domain.shape_halo[d]
            s = 0 if shift is None else int(shift[d])
            sl.append(slice(h - s, n - h - s))
        return tuple(sl)

    def pull_step(self, f, fnew):
        ns = self.ns
        shp = [ns] + list(self.domain.shape_in)
        fp = np.empty(shp)
        for k in range(ns):
            fp[k] = f[(k,) + self._interior(self.v[k])[1:]]
        m = self._f2m(fp)
        if self.scheme.rel_vel is not None:
            self._relative_velocity_collision(m)
            fnew[self._interior()] = self._m2f(np.einsum("ij...,j...->i...", self._tmu, m))
            return
        self._sources(m)
        eq = self._eq_full(m)
        for i in range(ns):
            s = self.scheme.s[i]
            if s != 0:
                m[i] = (1 - s) * m[i] + s * eq[i]
        self._sources(m)
        fnew[self._interior()] = self._m2f(m)

    def _relative_velocity_collision(self, m):
        """The collision of a relative-velocity scheme (pylbm/algorithm/pull.py:38-59 with
        base.py:255-263, 316-331, 404-428) on m = M f (all rows absolute on entry):
        u from the conserved moments, the non-conserved rows moved to the frame of u
        (rows of T(u) M f), [source dt/2], relaxation towards T(u) eq, the conserved
        rows replaced by their relative values (T(u) M f)[:nc] of the pulled f,
        [source dt/2]; leaves m in the relative frame and T(-u) in self._tmu for
        f = M^-1 T(-u) m."""
        sch = self.scheme
        nc, ns = sch.nconsm, self.ns
        u = sch.relative_velocity([m[i] for i in range(nc)])
        tu = sch.shift(u)
        self._tmu = sch.shift(u, inverse=True)
        mrel = np.einsum("ij...,j...->i...", tu, m)
        m[nc:] = mrel[nc:]
        self._sources(m)
        eq = np.einsum("ij...,j...->i...", tu, self._eq_full(m))
        for i in range(ns):
            s = sch.s[i]
            if s != 0:
                m[i] = (1 - s) * m[i] + s * eq[i]
        m[:nc] = mrel[:nc]
        self._sources(m)

    def _sources(self, m):
        if not self.scheme.sources:
            return
        coords = np.meshgrid(*self.domain.coords, sparse=True, indexing="ij")
        for idx, fn in self.scheme.sources:
            m[idx] = m[idx] + 0.5 * self.dt * fn(list(m), self.t, *coords)

    # ------------------------------------------------------------ split routines
    def transport(self, f=None, fnew=None):
        """fnew[k, x] = f[k, x - v_k] on the interior, the ghost slots of f as they are
        (pylbm/algorithm/base.py:273-296; Simulation.transport, simulation.py:322-327)."""
        f = self.F if f is None else f
        fnew = self.Fnew if fnew is None else fnew
        for k in range(self.ns):
            fnew[(k,) + self._interior()[1:]] = f[(k,) + self._interior(self.v[k])[1:]]

    def relaxation(self, m):
        """m_i <- (1 - s_i) m_i + s_i eq_i(m) on every cell of the moment array, in place
        (pylbm/algorithm/base.py:404-437; the routine relaxes towards eq itself, also for
        a relative-velocity scheme)."""
        eq = self._eq_full(m)
        for i in range(self.ns):
            s = self.scheme.s[i]
            if s != 0:
                m[i] = (1 - s) * m[i] + s * eq[i]

    def source_term(self, m):
        """One explicit Euler half step m_c <- m_c + dt/2 S(m, X, t) on the interior cells
        of the halo-inclusive moment array, in place (pylbm/algorithm/base.py:439-504,
        algorithm/ode.py:11-16)."""
        if not self.scheme.sources:
            return
        inner = m[self._interior()]
        self._sources(inner)
        m[self._interior()] = inner

    def one_time_step(self):
        self.ghost_fill(self.F)
        for meth in self.methods:
            if getattr(meth, "timed", False):
                self._prepare_rhs(meth, t=self.t)
            apply_links(self.F, meth)
        self.pull_step(self.F, self.Fnew)
        self.F, self.Fnew = self.Fnew, self.F
        self.t += self.dt
        self.nt += 1

    # ------------------------------------------------------------ read-back
    def moments(self):
        """All moments on the interior (ns, *shape_in)."""
        return self._f2m(self.F[self._interior()])

    def m(self, key):
        i = self.scheme.consm[key] if not isinstance(key, (int, np.integer)) else key
        return self.moments()[i]

    def conserved_sums(self):
        """math.fsum of every conserved moment over the interior."""
        mm = self.moments()
        return [math.fsum(mm[i].ravel()) for i in range(self.scheme.nconsm)]
