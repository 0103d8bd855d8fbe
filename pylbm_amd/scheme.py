"""Scheme description: moment matrix, equilibrium, relaxation rates.

Mirrors pylbm/scheme.py:136-241, 464-607 and the substitutions of
pylbm/algorithm/base.py:91-155:

* ``M[i, j] = P_i(la v_j)`` per elementary scheme, block-diagonal over the
  schemes, with the exact rational inverse (scheme.py:464-532);
* the equilibrium list, conserved-moment indices = position of the symbol
  in the scheme's equilibrium list (scheme.py:202-223, 575-607);
* conserved moments moved to the front by *sequential* swaps of EQ/s rows,
  M rows and M^-1 columns (scheme.py:225-241);
* ``parameters`` substituted, equilibrium nsimplify'd (base.py:127-135).

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

    def key(self):
        """Canonical text used to name/cache generated code."""
        parts = [str(self.dim), str(self.ns), str(self.nconsm), repr(self.velocities),
                 repr([[str(c) for c in r] for r in self.M]),
                 repr([sp.srepr(e) for e in self.eq]),
                 repr([str(sp.Integer(0) if e == 0 else "s") for e in self.s]),
                 repr([(i, sp.srepr(e)) for i, e in self.sources])]
        return "|".join(parts)


class Scheme:
    """pylbm.Scheme for one or several elementary schemes (vectorial schemes)."""

    def __init__(self, dico, check_inverse=False, need_validation=True):
        self.stencil = Stencil(dico, need_validation=False)
        self.dim = _dim_from_dico(dico)
        self.param = dict(dico.get("parameters", {}) or {})
        self.la = dico["scheme_velocity"]
        self.rel_vel = dico.get("relative_velocity")
        if self.rel_vel is not None:
            raise NotImplementedError("relative_velocity schemes are not supported by the hip generator")
        schemes = dico["schemes"]
        self.nschemes = len(schemes)
        st = self.stencil
        for key in ("polynomials", "relaxation_parameters", "equilibrium"):
            for i, sch in enumerate(schemes):
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
        for k in range(self.nschemes):
            lo, hi = int(st.nv_ptr[k]), int(st.nv_ptr[k + 1])
            vel = st.get_all_velocities(k)
            blk = sp.Matrix(hi - lo, hi - lo, lambda i, j: sp.nsimplify(_fixed_point_subs(
                self.P[lo + i].subs({c: int(vel[j][d]) * la for d, c in enumerate(coords)}), pairs)))
            M[lo:hi, lo:hi] = blk
            invM[lo:hi, lo:hi] = blk.inv()
        EQ = [sp.sympify(e) for sch in schemes for e in sch["equilibrium"]]
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
        self.consm = {c: i for i, c in enumerate(consm)}
        self.M, self.invM, self.EQ, self.s = M, invM, EQ, s
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
                sources.append((idx, sp.sympify(expr)))
        frac = lambda mat: [[Fraction(int(sp.fraction(mat[i, j])[0]), int(sp.fraction(mat[i, j])[1]))
                             for j in range(ns)] for i in range(ns)]
        return KernelSpec(dim=self.dim, ns=ns, nconsm=nc,
                          velocities=[tuple(int(c) for c in v) for v in self.stencil.get_all_velocities()],
                          M=frac(self.M), invM=frac(self.invM), eq=eq, s=s, sources=sources, msyms=msyms,
                          blocks=[(int(self.stencil.nv_ptr[k]), int(self.stencil.nv_ptr[k + 1]))
                                  for k in range(self.nschemes)])
