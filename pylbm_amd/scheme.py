"""Scheme description: moment matrix, equilibrium, relaxation rates.

Mirrors pylbm/scheme.py:136-241, 464-607 and the substitutions of
pylbm/algorithm/base.py:91-155:

* ``M[i, j] = P_i(la v_j)`` per elementary scheme, block-diagonal over the
  schemes, with the exact rational inverse (scheme.py:464-532);
* the equilibrium list, conserved-moment indices = position of the symbol
  in the scheme's equilibrium list (scheme.py:202-223, 575-607);
* conserved moments moved to the front by *sequential* swaps of EQ/s rows,
  M rows and M^-1 columns (scheme.py:225-241);
* ``parameters`` substituted, equilibrium nsimplify'd (base.py:127-135);
* ``feq`` equilibria: the equilibrium moments of a block are
  M_block . feq(velocities, *args) (scheme.py:202-223);
* relative-velocity schemes (``relative_velocity``, scheme.py:149, 464-531):
  per elementary scheme Mu[i, j] = P_i(la v_j - u), the shift matrix
  T(u) = Mu M^-1 (the identity outside the blocks) and T(-u), permuted by the
  same sequential swaps as M (rows and columns, scheme.py:225-241), and the
  relative velocity as an expression of the conserved moments
  (base.py:102-114, 141-143).

``kernel_spec()`` freezes the result into the plain data the HIP code
generator consumes (exact rationals for M, M^-1; sympy expressions of the
conserved moments for EQ and sources; relaxation rates as numbers).
"""
from dataclasses import dataclass, field
from fractions import Fraction
from typing import List, Tuple

import sympy as sp

from .stencil import Stencil, _dim_from_dico

__all__ = ["Scheme", "KernelSpec"]

X, Y, Z = sp.symbols("X, Y, Z")
REL_U = sp.symbols("rel_ux, rel_uy, rel_uz")   # pylbm's relative-velocity symbols (pylbm/symbolic.py)
RU = sp.symbols("ru0, ru1, ru2")               # their names in the generated code


def _exact_or_double(x) -> Fraction:
    """An entry of M or M^-1 as a Fraction: exact when it is rational, else the
    correctly rounded double of its value (a scheme velocity such as 10 sqrt(3) makes
    the entries irrational; pylbm computes them in floating point)."""
    x = sp.sympify(x)
    if x.is_Rational:
        return Fraction(int(x.p), int(x.q))
    return Fraction(float(sp.Float(x.evalf(40), 40)))


def _fixed_point_subs(expr, pairs):
    expr = sp.sympify(expr)
    for _ in range(max(1, len(pairs))):
        nxt = expr.subs(pairs)
        if nxt == expr:
            break
        expr = nxt
    return expr


@dataclass
class KernelSpec:
    """Everything the code generator needs, detached from sympy matrices."""
    dim: int
    ns: int
    nconsm: int
    velocities: List[Tuple[int, ...]]
    M: List[List[Fraction]]
    invM: List[List[Fraction]]
    eq: List[sp.Expr]                      # in symbols m0..m{nconsm-1}
    s: List[sp.Expr]                       # numbers (or symbols left free)
    sources: List[Tuple[int, sp.Expr]] = field(default_factory=list)
    msyms: List[sp.Symbol] = field(default_factory=list)
    blocks: List[Tuple[int, int]] = field(default_factory=list)
    # relative-velocity schemes: u as expressions of m0..m{nconsm-1}, and T(u), T(-u)
    # as ns x ns lists of expressions of ru0..ru{dim-1}; None otherwise
    rel_vel: List[sp.Expr] = None
    Tu: List[List[sp.Expr]] = None
    Tmu: List[List[sp.Expr]] = None

    def key(self):
        """Canonical text used to name/cache generated code."""
        parts = [str(self.dim), str(self.ns), str(self.nconsm), repr(self.velocities),
                 repr([[str(c) for c in r] for r in self.M]),
                 repr([sp.srepr(e) for e in self.eq]),
                 repr([str(sp.Integer(0) if e == 0 else "s") for e in self.s]),
                 repr([(i, sp.srepr(e)) for i, e in self.sources])]
        if self.rel_vel is not None:
            parts += [repr([sp.srepr(e) for e in self.rel_vel]),
                      repr([[sp.srepr(e) for e in r] for r in self.Tu]),
                      repr([[sp.srepr(e) for e in r] for r in self.Tmu])]
        return "|".join(parts)


class Scheme:
    """pylbm.Scheme for one or several elementary schemes (vectorial schemes)."""

    def __init__(self, dico, check_inverse=False, need_validation=True):
        self.stencil = Stencil(dico, need_validation=False)
        self.dim = _dim_from_dico(dico)
        self.param = dict(dico.get("parameters", {}) or {})
        self.la = dico["scheme_velocity"]
        self.rel_vel = dico.get("relative_velocity")
        schemes = dico["schemes"]
        self.nschemes = len(schemes)
        st = self.stencil
        for i, sch in enumerate(schemes):
            if ("feq" in sch) == ("equilibrium" in sch):
                raise KeyError("scheme %d needs exactly one of 'equilibrium' and 'feq'" % i)
        for key in ("polynomials", "relaxation_parameters", "equilibrium"):
            for i, sch in enumerate(schemes):
                if key == "equilibrium" and key not in sch:
                    continue
                if key not in sch:
                    raise KeyError("scheme %d has no %r entry" % (i, key))
                if len(sch[key]) != st.nv[i]:
                    raise ValueError("scheme %d: %r has %d entries for %d velocities"
                                     % (i, key, len(sch[key]), st.nv[i]))
        pairs = list(self.param.items())
        la = _fixed_point_subs(self.la, pairs)
        coords = [X, Y, Z][: self.dim]
        ns = int(st.nv_ptr[-1])
        self.P = sp.Matrix([sp.sympify(p) for sch in schemes for p in sch["polynomials"]])
        M = sp.zeros(ns, ns)
        invM = sp.zeros(ns, ns)
        Tu = sp.eye(ns)
        u = REL_U[: self.dim]
        for k in range(self.nschemes):
            lo, hi = int(st.nv_ptr[k]), int(st.nv_ptr[k + 1])
            vel = st.get_all_velocities(k)
            blk = sp.Matrix(hi - lo, hi - lo, lambda i, j: sp.nsimplify(_fixed_point_subs(
                self.P[lo + i].subs({c: int(vel[j][d]) * la for d, c in enumerate(coords)}), pairs)))
            M[lo:hi, lo:hi] = blk
            iblk = blk.inv()
            invM[lo:hi, lo:hi] = iblk
            if self.rel_vel is not None:
                mu = sp.Matrix(hi - lo, hi - lo, lambda i, j: sp.nsimplify(_fixed_point_subs(
                    self.P[lo + i].subs({c: int(vel[j][d]) * la - u[d] for d, c in enumerate(coords)},
                                        simultaneous=True), pairs)))
                Tu[lo:hi, lo:hi] = (mu * iblk).applyfunc(sp.expand)
        Tmu = Tu.subs(list(zip(u, [-c for c in u])), simultaneous=True)
        EQ = []
        for k, sch in enumerate(schemes):
            if "feq" in sch:
                lo, hi = int(st.nv_ptr[k]), int(st.nv_ptr[k + 1])
                fn = sch["feq"][0]
                args = tuple(sch["feq"][1]) if len(sch["feq"]) > 1 else ()
                feq = sp.Matrix([sp.sympify(e) for e in fn(st.get_all_velocities(k), *args)])
                EQ += [sp.simplify(e) for e in M[lo:hi, lo:hi] * feq]
            else:
                EQ += [sp.sympify(e) for e in sch["equilibrium"]]
        s = [sp.sympify(r) for sch in schemes for r in sch["relaxation_parameters"]]
        consm = {}
        for k, sch in enumerate(schemes):
            cm = sch.get("conserved_moments")
            if cm is None:
                continue
            cm = list(cm) if isinstance(cm, (list, tuple)) else [cm]
            block = EQ[int(st.nv_ptr[k]):int(st.nv_ptr[k + 1])]
            for c in cm:
                consm[c] = int(st.nv_ptr[k]) + block.index(c)
        for front, pos in enumerate(list(consm.values())):
            EQ[front], EQ[pos] = EQ[pos], EQ[front]
            s[front], s[pos] = s[pos], s[front]
            M.row_swap(front, pos)
            invM.col_swap(front, pos)
            for T in (Tu, Tmu):
                T.row_swap(front, pos)
                T.col_swap(front, pos)
        self.consm = {c: i for i, c in enumerate(consm)}
        self.M, self.invM, self.EQ, self.s = M, invM, EQ, s
        self.Tu, self.Tmu = Tu, Tmu
        if check_inverse and M * invM != sp.eye(ns):
            raise ValueError("M * invM is not the identity")
        self._source_terms = [sch.get("source_terms") for sch in schemes]
        self._pairs = pairs

    @property
    def nconsm(self):
        return len(self.consm)

    def kernel_spec(self, extra_subs=()):
        ns = self.M.shape[0]
        nc = self.nconsm
        msyms = [sp.Symbol("m%d" % i) for i in range(ns)]
        to_m = [(c, msyms[i]) for c, i in self.consm.items()]
        pairs = self._pairs + list(extra_subs)
        eq = [sp.nsimplify(_fixed_point_subs(e, pairs + to_m)) for e in self.EQ]
        for i, e in enumerate(eq):
            bad = e.free_symbols - set(msyms[:nc])
            if bad:
                raise NotImplementedError("equilibrium %d depends on %s (only conserved moments are "
                                          "supported by the hip generator)" % (i, bad))
        s = [_fixed_point_subs(r, pairs) for r in self.s]
        # a rate given as 0.0 is a Float, which sympy does not compare equal to 0
        s = [sp.Integer(0) if (r.is_number and float(r) == 0.0) else r for r in s]
        for i in range(nc):
            if s[i] != 0:
                raise ValueError("conserved moment %d has a non-zero relaxation rate" % i)
        sources = []
        for sts in self._source_terms:
            if not sts:
                continue
            for lhs, rhs in sts.items():
                idx = lhs if isinstance(lhs, int) else self.consm[sp.sympify(lhs) if isinstance(lhs, str) else lhs]
                expr = rhs if isinstance(rhs, (int, float)) else _fixed_point_subs(sp.sympify(rhs), pairs + to_m)
                expr = sp.sympify(expr)
                allowed = set(msyms[:nc]) | set([X, Y, Z][: self.dim]) | {sp.Symbol("t")}
                bad = expr.free_symbols - allowed
                if bad:
                    raise NotImplementedError("source term of moment %d depends on %s (conserved moments, "
                                              "X/Y/Z and t are supported)" % (idx, bad))
                sources.append((idx, expr))
        rel = Tus = Tmus = None
        if self.rel_vel is not None:
            rv = self.rel_vel if isinstance(self.rel_vel, (list, tuple)) else [self.rel_vel]
            if len(rv) != self.dim:
                raise ValueError("relative_velocity needs %d components" % self.dim)
            rel = [sp.nsimplify(_fixed_point_subs(sp.sympify(e), pairs + to_m)) for e in rv]
            for e in rel:
                bad = e.free_symbols - set(msyms[:nc])
                if bad:
                    raise NotImplementedError("relative velocity depends on %s (only conserved moments are "
                                              "supported)" % bad)
            ren = list(zip(REL_U, RU))
            Tus = [[sp.nsimplify(self.Tu[i, j]).xreplace(dict(ren)) for j in range(ns)] for i in range(ns)]
            Tmus = [[sp.nsimplify(self.Tmu[i, j]).xreplace(dict(ren)) for j in range(ns)] for i in range(ns)]
        frac = lambda mat: [[_exact_or_double(mat[i, j]) for j in range(ns)] for i in range(ns)]
        return KernelSpec(dim=self.dim, ns=ns, nconsm=nc,
                          velocities=[tuple(int(c) for c in v) for v in self.stencil.get_all_velocities()],
                          M=frac(self.M), invM=frac(self.invM), eq=eq, s=s, sources=sources, msyms=msyms,
                          blocks=[(int(self.stencil.nv_ptr[k]), int(self.stencil.nv_ptr[k + 1]))
                                  for k in range(self.nschemes)],
                          rel_vel=rel, Tu=Tus, Tmu=Tmus)
