"""Scheme mathematics -- oracle restatement.

Follows pylbm/scheme.py:
  * M[i, j] = P_i(la * v_j), exact inverse          scheme.py:464-532
  * equilibrium list                                scheme.py:202-223
  * conserved-moment indices (index of the symbol
    inside the scheme's equilibrium list)           scheme.py:575-607
  * sequential row/column swaps putting conserved
    moments in front                                scheme.py:225-241
and the parameter substitution / simplification of
pylbm/algorithm/base.py:91-155 (recursive_sub + nsimplify).

Relative-velocity schemes (dict key ``relative_velocity``, scheme.py:149,
464-531): per elementary scheme Mu[i, j] = P_i(la v_j - u) with u the symbols
rel_ux, rel_uy, rel_uz, the shift matrix T(u) = Mu M^-1 (block-diagonal, the
identity elsewhere) and T(-u), both permuted by the same sequential swaps as M
(rows and columns, scheme.py:233-236); the relative velocity itself is the
dict's expression of the conserved moments (base.py:102-114, 141-143).
Pinned by T(u) T(-u) = I exactly (scheme.py:534-543), T(0) = I, and
Mu = T(u) M evaluated against P_i(la v_j - u) with numpy.

``feq`` equilibria (scheme.py:202-223): the equilibrium moments of a block are
M_block . feq(velocities, *args).

Pinned by: M @ invM == I exactly (scheme.py:545-573), rest weights of the
standard lattices (D2Q9: 4/9, 1/9, 1/36; D3Q19: 1/3, 1/18, 1/36), and the
closed form M_ij = P_i(la v_j) evaluated independently with numpy in
tests/test_oracle_scheme.py.
"""
import numpy as np
import sympy as sp

from .stencil import Stencil

X, Y, Z = sp.symbols("X, Y, Z")
T = sp.Symbol("t")
REL_U = sp.symbols("rel_ux, rel_uy, rel_uz")


def _subs_all(expr, pairs):
    expr = sp.sympify(expr)
    for _ in range(max(1, len(pairs))):
        new = expr.subs(pairs)
        if new == expr:
            return new
        expr = new
    return expr


def _dim_of(dico):
    if "dim" in dico:
        return dico["dim"]
    box = dico["box"]
    return 1 + ("y" in box) + ("z" in box and "y" in box)


class Scheme:
    def __init__(self, dico):
        self.dim = _dim_of(dico)
        schemes = dico["schemes"]
        self.stencil = Stencil(self.dim, [s["velocities"] for s in schemes])
        st = self.stencil
        self.param = dict(dico.get("parameters", {}))
        pairs = list(self.param.items())
        la = _subs_all(dico["scheme_velocity"], pairs)
        self.la = float(la)
        ns = int(st.nv_ptr[-1])
        self.ns = ns
        coords = [X, Y, Z][: self.dim]
        self.rel_vel = dico.get("relative_velocity")
        M = sp.zeros(ns, ns)
        invM = sp.zeros(ns, ns)
        Tu = sp.eye(ns)
        u = REL_U[: self.dim]
        for k, sch in enumerate(schemes):
            polys = sch["polynomials"]
            vel = [st.all_velocities()[st.nv_ptr[k] + j] for j in range(st.nv[k])]
            Mk = sp.zeros(st.nv[k], st.nv[k])
            Muk = sp.zeros(st.nv[k], st.nv[k])
            for i, p in enumerate(polys):
                for j, v in enumerate(vel):
                    e = sp.sympify(p).subs([(c, sp.Integer(int(v[d])) * la) for d, c in enumerate(coords)])
                    Mk[i, j] = sp.nsimplify(_subs_all(e, pairs))
                    if self.rel_vel is not None:
                        e = sp.sympify(p).subs([(c, sp.Integer(int(v[d])) * la - u[d])
                                                for d, c in enumerate(coords)], simultaneous=True)
                        Muk[i, j] = sp.nsimplify(_subs_all(e, pairs))
            iMk = Mk.inv()
            o = st.nv_ptr[k]
            M[o:o + st.nv[k], o:o + st.nv[k]] = Mk
            invM[o:o + st.nv[k], o:o + st.nv[k]] = iMk
            if self.rel_vel is not None:
                Tu[o:o + st.nv[k], o:o + st.nv[k]] = (Muk * iMk).applyfunc(sp.expand)
        eq = []
        s = []
        for k, sch in enumerate(schemes):
            if sch.get("feq") is not None:
                fn, args = sch["feq"][0], tuple(sch["feq"][1]) if len(sch["feq"]) > 1 else ()
                o = st.nv_ptr[k]
                blk = M[o:o + st.nv[k], o:o + st.nv[k]]
                vel = np.asarray([st.all_velocities()[o + j] for j in range(st.nv[k])])
                feq = sp.Matrix(list(fn(vel, *args)))
                eq.extend(sp.simplify(e) for e in blk * feq)
            else:
                eq.extend(sp.sympify(e) for e in sch["equilibrium"])
            s.extend(sch["relaxation_parameters"])
        Tmu = Tu.subs(list(zip(u, [-c for c in u])), simultaneous=True)
        # conserved moments: index of the symbol in the scheme's equilibrium list
        consm = {}
        for k, sch in enumerate(schemes):
            cm = sch.get("conserved_moments")
            if cm is None:
                continue
            if not isinstance(cm, (list, tuple)):
                cm = [cm]
            leq = eq[st.nv_ptr[k]:st.nv_ptr[k + 1]]
            for c in cm:
                consm[c] = int(st.nv_ptr[k]) + leq.index(c)
        # sequential swaps (scheme.py:225-241)
        s = list(s)
        for ic, c in enumerate(list(consm.values())):
            eq[ic], eq[c] = eq[c], eq[ic]
            s[ic], s[c] = s[c], s[ic]
            M.row_swap(ic, c)
            invM.col_swap(ic, c)
            Tu.row_swap(ic, c)
            Tu.col_swap(ic, c)
            Tmu.row_swap(ic, c)
            Tmu.col_swap(ic, c)
        self.consm = {c: i for i, c in enumerate(consm.keys())}
        self.nconsm = len(self.consm)
        self.M = M
        self.invM = invM
        self.Mnum = np.array(M.evalf(30).tolist(), dtype=np.float64)
        self.invMnum = np.array(invM.evalf(30).tolist(), dtype=np.float64)
        self.msyms = [sp.Symbol("m%d" % i) for i in range(ns)]
        to_m = [(c, self.msyms[i]) for c, i in self.consm.items()]
        self.eq = [sp.nsimplify(_subs_all(e, pairs + to_m)) for e in eq]
        self.s = [float(_subs_all(v, pairs)) for v in s]
        cons_syms = self.msyms[: self.nconsm]
        self._eq_fn = [sp.lambdify(cons_syms, e, "numpy") for e in self.eq]
        self.Tu, self.Tmu = Tu, Tmu
        if self.rel_vel is not None:
            rv = self.rel_vel if isinstance(self.rel_vel, (list, tuple)) else [self.rel_vel]
            self.rel_vel_exprs = [sp.nsimplify(_subs_all(sp.sympify(e), pairs + to_m)) for e in rv]
            self._rel_fn = [sp.lambdify(cons_syms, e, "numpy") for e in self.rel_vel_exprs]
            self._tu_fn = [[sp.lambdify(list(u), e, "numpy") for e in Tu.row(i)] for i in range(ns)]
            self._tmu_fn = [[sp.lambdify(list(u), e, "numpy") for e in Tmu.row(i)] for i in range(ns)]
        # source terms: (moment index, rhs function of (m, t, coords))
        self.sources = []
        self.source_exprs = []
        for sch in schemes:
            st_dico = sch.get("source_terms")
            if not st_dico:
                continue
            for lhs, rhs in st_dico.items():
                idx = self.consm[lhs] if not isinstance(lhs, int) else lhs
                r = _subs_all(sp.sympify(rhs), pairs + to_m)
                fn = sp.lambdify([self.msyms, T] + list(coords), r, "numpy")
                self.sources.append((idx, fn))
                self.source_exprs.append((idx, r))

    def relative_velocity(self, mcons):
        """The relative velocity arrays (base.py:255-257) from the conserved moments."""
        shape = np.shape(mcons[0])
        return [np.broadcast_to(np.asarray(fn(*mcons), dtype=np.float64), shape) for fn in self._rel_fn]

    def shift(self, u, inverse=False):
        """T(u) (or T(-u)) evaluated per cell: array (ns, ns, *cells)."""
        shape = np.shape(u[0])
        fns = self._tmu_fn if inverse else self._tu_fn
        return np.stack([np.stack([np.broadcast_to(np.asarray(fn(*u), dtype=np.float64), shape) for fn in r])
                         for r in fns])

    def equilibrium(self, mcons):
        """Equilibrium moments from the list of conserved-moment arrays."""
        out = []
        shape = np.shape(mcons[0])
        for fn in self._eq_fn:
            out.append(np.broadcast_to(np.asarray(fn(*mcons), dtype=np.float64), shape))
        return out
