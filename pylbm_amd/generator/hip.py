"""HIP (gfx950) code generator for one lattice Boltzmann scheme.

Plays the role of pylbm's code printers + ``CodeGen`` (pylbm/generator/codegen.py:
1014-1178 for Cython, printing/cython.py:301-350) for the routines
``one_time_step`` (pylbm/algorithm/pull.py:11-59), ``f2m``, ``m2f`` and
``equilibrium`` (pylbm/algorithm/base.py:298-402): from the frozen
:class:`~pylbm_amd.scheme.KernelSpec` it prints one ``.hip`` translation unit
holding

* ``fused_step<S, FOLD>``: pull f_k(x - v_k), m = M f, [source dt/2],
  relax m_i <- (1 - s_i) m_i + s_i eq_i(m_c), [source dt/2], fnew = M^-1 m,
  one thread per interior cell, stored through the population codec S;
* ``f2m_kernel<S>`` / ``m2f_kernel<S>`` / ``equilibrium_kernel`` over the whole
  halo-inclusive domain with FP64 moments;
* the ``extern "C"`` entry points declared in include/pylbm_hip.h.

The q x q transforms are printed from the exact rational matrices.  When every
moment polynomial has a definite parity under v -> -v (true for the usual MRT
bases) the transforms are factored through the pair sums f_k + f_ksym and
differences f_k - f_ksym, which halves the work; terms sharing a coefficient
are grouped and identical partial sums are computed once.  Relaxation rates
are run-time arguments, so one compiled module serves every viscosity; only
the zero pattern of s (which rows are relaxed at all, pylbm drops identity
rows, printing/cython.py:347-349) is baked in.
"""
import hashlib
import os
from fractions import Fraction

import sympy as sp
from sympy.printing.c import C99CodePrinter

from ..scheme import KernelSpec

__all__ = ["generate_source", "spec_hash"]


def _source_tag() -> str:
    """Digest of the code generator, its compile flags and the device header it
    includes: a module built from other sources never matches the cache key."""
    h = hashlib.sha1()
    here = os.path.dirname(os.path.abspath(__file__))
    for path in (os.path.abspath(__file__), os.path.join(here, "build.py"),
                 os.path.join(here, "..", "csrc", "lbm_device.h"),
                 os.path.join(here, "..", "..", "include", "pylbm_hip.h")):
        with open(path, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:12]


_SOURCE_TAG = _source_tag()


def spec_hash(spec: KernelSpec) -> str:
    return hashlib.sha1((spec.key() + _SOURCE_TAG).encode()).hexdigest()[:16]


def _lit(x) -> str:
    """Correctly rounded FP64 literal of an exact rational, cast to the compute type."""
    v = float(Fraction(x)) if not isinstance(x, float) else x
    r = repr(v)
    if "e" not in r and "." not in r and "inf" not in r and "nan" not in r:
        r += ".0"
    return "C(%s)" % r


class _Printer(C99CodePrinter):
    """Prints sympy expressions in the compute type ``C`` of the kernels."""

    def _print_Symbol(self, expr):
        name = expr.name
        if name == "t":
            return "tt"
        if name in ("X", "Y", "Z"):
            return "xc%d" % "XYZ".index(name)
        if name.startswith("cr") and name[2:].isdigit():
            return "cref[%s]" % name[2:]
        return name

    def _print_Integer(self, expr):
        return _lit(int(expr))

    def _print_Rational(self, expr):
        return _lit(Fraction(int(expr.p), int(expr.q)))

    def _print_Float(self, expr):
        return _lit(float(expr))

    def _print_Pow(self, expr):
        b, e = expr.as_base_exp()
        if e.is_Integer:
            n = int(e)
            if n == 0:
                return "C(1.0)"
            base = self.parenthesize(b, 100)
            prod = "*".join([base] * abs(n))
            return "(%s)" % prod if n > 0 else "(C(1.0)/(%s))" % prod
        if e == sp.Rational(1, 2):
            return "sqrt(%s)" % self._print(b)
        if e.is_number and e.is_real:
            return "pow(%s, %s)" % (self._print(b), self._print(e))
        raise NotImplementedError("power %s not supported" % expr)


def _ccode(expr) -> str:
    return _Printer().doprint(expr)


# ----------------------------------------------------------------- structure
def _canon(v, dim):
    """Reference axes -> canonical (outer, mid, inner)."""
    v = list(v)
    return tuple([0] * (3 - dim) + v)


def _pairs(spec: KernelSpec):
    """ksym[k] inside the same block of M, or None if the stencil is not symmetric."""
    ns = spec.ns
    blocks = spec.blocks
    ksym = [None] * ns
    for lo, hi in blocks:
        index = {tuple(spec.velocities[j]): j for j in range(lo, hi)}
        for k in range(lo, hi):
            neg = tuple(-c for c in spec.velocities[k])
            if neg not in index:
                return None
            ksym[k] = index[neg]
    return ksym


def _row_parities(mat, ksym, by_rows=True):
    """Parity sigma_i of every row i (by_rows: mat[i][ksym j] == sigma_i mat[i][j]) or
    column i (mat[ksym j][i] == sigma_i mat[j][i]); None if one is indefinite."""
    ns = len(mat)
    get = (lambda i, j: mat[i][j]) if by_rows else (lambda i, j: mat[j][i])
    sig = []
    for i in range(ns):
        s = 0
        for j in range(ns):
            a = get(i, j)
            if a == 0 or ksym[j] == j:
                continue
            b = get(i, ksym[j])
            if b == a:
                s = 1
            elif b == -a:
                s = -1
            else:
                return None
            break
        s = s or 1
        for j in range(ns):
            if get(i, ksym[j]) != s * get(i, j):
                return None
        sig.append(s)
    return sig


class _SumBuilder:
    """Emits grouped linear combinations, reusing identical partial sums."""

    def __init__(self, prefix):
        self.lines = []
        self.cache = {}
        self.prefix = prefix
        self.count = 0

    def _group(self, terms):
        # terms: list of (sign, var) with sign +-1
        key = tuple(sorted(terms))
        if len(terms) == 1 and terms[0][0] == 1:
            return terms[0][1]
        if key in self.cache:
            return self.cache[key]
        pos = [v for s, v in terms if s > 0]
        neg = [v for s, v in terms if s < 0]
        if pos:
            expr = " + ".join(pos) + "".join(" - %s" % v for v in neg)
        else:
            expr = "-(" + " + ".join(neg) + ")"
        name = "%s%d" % (self.prefix, self.count)
        self.count += 1
        self.lines.append("const C %s = %s;" % (name, expr))
        self.cache[key] = name
        return name

    def combo(self, coeffs):
        """coeffs: list of (Fraction, var). Returns a C expression."""
        groups = {}
        for c, v in coeffs:
            if c == 0:
                continue
            a = abs(c)
            groups.setdefault(a, []).append((1 if c > 0 else -1, v))
        if not groups:
            return "C(0.0)"
        parts = []
        for a in sorted(groups):
            g = self._group(groups[a])
            parts.append(g if a == 1 else "%s*%s" % (_lit(a), g))
        return " + ".join(parts)


def _transform(builder, mat, invars, outvars, ksym, sig, rows_are_outputs=True):
    """Lines computing outvars = mat @ invars (f2m) using pair parity if available."""
    ns = len(mat)
    lines = []
    if ksym is not None and sig is not None:
        # pair sums/differences of the inputs
        reps = [k for k in range(ns) if k <= ksym[k]]
        P, D = {}, {}
        for k in reps:
            if ksym[k] == k:
                P[k] = invars[k]
            else:
                P[k] = builder._group([(1, invars[k]), (1, invars[ksym[k]])])
                D[k] = builder._group([(1, invars[k]), (-1, invars[ksym[k]])])
        for i in range(ns):
            coeffs = []
            for k in reps:
                c = mat[i][k]
                if c == 0:
                    continue
                if sig[i] > 0:
                    coeffs.append((c, P[k]))
                elif ksym[k] != k:
                    coeffs.append((c, D[k]))
            lines.append((outvars[i], builder.combo(coeffs)))
    else:
        for i in range(ns):
            lines.append((outvars[i], builder.combo([(mat[i][j], invars[j]) for j in range(ns)])))
    return lines


def _inverse_transform(builder, inv, mvars, fvars, ksym, sig, base=None, tag=""):
    """Lines computing fvars = [base +] inv @ mvars (entries of mvars that are None
    are zero); with parity f_k = E_k + O_k, f_ksym = E_k - O_k."""
    ns = len(inv)
    lines = []

    def plus(k, expr):
        return "%s + (%s)" % (base[k], expr) if base is not None else expr

    if ksym is not None and sig is not None:
        for k in range(ns):
            if k > ksym[k]:
                continue
            even = [(inv[k][i], mvars[i]) for i in range(ns) if sig[i] > 0 and mvars[i] is not None]
            odd = [(inv[k][i], mvars[i]) for i in range(ns) if sig[i] < 0 and mvars[i] is not None]
            if ksym[k] == k:
                lines.append((fvars[k], plus(k, builder.combo(even))))
                continue
            e = "E%s%d" % (tag, k)
            o = "O%s%d" % (tag, k)
            builder.lines.append("const C %s = %s;" % (e, builder.combo(even)))
            builder.lines.append("const C %s = %s;" % (o, builder.combo(odd)))
            lines.append((fvars[k], plus(k, "%s + %s" % (e, o))))
            lines.append((fvars[ksym[k]], plus(ksym[k], "%s - %s" % (e, o))))
    else:
        for k in range(ns):
            lines.append((fvars[k], plus(k, builder.combo([(inv[k][i], mvars[i]) for i in range(ns)
                                                           if mvars[i] is not None]))))
    return lines


def _deviation_eq(expr, full, rest):
    """eq(c_ref + m') - eq(c_ref) with the constant cancelled symbolically."""
    e = sp.expand(sp.sympify(expr).xreplace(full) - sp.sympify(expr).xreplace(rest))
    zero = {s: 0 for s in full}
    if sp.simplify(e.xreplace(zero)) != 0:
        e = sp.cancel(sp.together(e))
    return e


def _collision_lines(spec: KernelSpec, with_sources: bool):
    """Full-value collision (FP64, as pylbm): sources, relax, sources on m0..m{ns-1}."""
    ns, nc = spec.ns, spec.nconsm
    out = []

    def sources():
        for idx, expr in spec.sources:
            out.append("m%d = m%d + half_dt * (%s);" % (idx, idx, _ccode(sp.sympify(expr))))

    if with_sources and spec.sources:
        sources()
    relaxed = [i for i in range(nc, ns) if spec.s[i] != 0]
    repl, red = sp.cse([spec.eq[i] for i in relaxed], symbols=sp.numbered_symbols("q_"),
                       optimizations="basic")
    for sym, e in repl:
        out.append("const C %s = %s;" % (sym, _ccode(e)))
    for i, e in zip(relaxed, red):
        out.append("m%d = one_minus[%d] * m%d + omega[%d] * (%s);" % (i, i, i, i, _ccode(e)))
    if with_sources and spec.sources:
        sources()
    return out


def _deviation_collision_lines(spec: KernelSpec):
    """Collision of the FP32-compute formats, in increments.

    The m_i are moments of d = f - w (w the rest populations of the reference
    state c_ref = (rho0, 0, ..., 0), rho0 = cref[0] at run time), so the
    conserved moments are c_ref + m_c and the target of a relaxed row is
    eq_i(c_ref + m_c) - eq_i(c_ref).  The kernel computes the change of every
    moment, dm_i = s_i (eq_i' - m_i) for relaxed rows and the source half
    steps for source rows, and applies fnew = d + M^-1 dm: equal in exact
    arithmetic to M^-1 (m + dm), but the rounding of the M^-1 coefficients
    now multiplies the small non-equilibrium change instead of the moments,
    and the rest state (dm = 0) is reproduced exactly.

    Returns (lines, names) with names[i] the C name of dm_i or None."""
    ns, nc = spec.ns, spec.nconsm
    msyms = spec.msyms
    crefs = [sp.Symbol("cr0")] + [sp.Integer(0)] * (nc - 1) if nc else []
    full = {msyms[i]: crefs[i] + msyms[i] for i in range(nc)}
    rest = {msyms[i]: crefs[i] for i in range(nc)}
    out = []
    names = [None] * ns
    src_rows = sorted({idx for idx, _ in spec.sources})
    for idx in src_rows:
        out.append("C dm%d = C(0.0);" % idx)
        names[idx] = "dm%d" % idx

    def sources():
        for idx, expr in spec.sources:
            e = sp.sympify(expr).xreplace(full)
            out.append("{ const C inc = half_dt * (%s); m%d += inc; dm%d += inc; }" % (_ccode(e), idx, idx))

    sources()
    relaxed = [i for i in range(nc, ns) if spec.s[i] != 0]
    eqs = [_deviation_eq(spec.eq[i], full, rest) for i in relaxed]
    repl, red = sp.cse(eqs, symbols=sp.numbered_symbols("q_"), optimizations="basic")
    for sym, e in repl:
        out.append("const C %s = %s;" % (sym, _ccode(e)))
    for i, e in zip(relaxed, red):
        target = "C(0.0)" if e == 0 else "(%s)" % _ccode(e)
        if names[i] is None:
            out.append("C dm%d = omega[%d] * (%s - m%d);" % (i, i, target, i))
            names[i] = "dm%d" % i
        else:
            out.append("{ const C inc = omega[%d] * (%s - m%d); dm%d += inc; m%d += inc; }" % (i, target, i, i, i))
            continue
        out.append("m%d += dm%d;" % (i, i))
    sources()
    return out, names


def _relative_collision_lines(spec: KernelSpec):
    """Collision of a relative-velocity scheme on m = M f (pylbm/algorithm/pull.py:38-59
    with base.py:255-263, 316-331, 366-368, 404-428), in full values:

    u = rel_vel(m_c); mr = T(u) m (the rows of Mu f = T(u) M f); the non-conserved
    rows take their relative values; [source dt/2]; relaxed rows
    m_i <- (1 - s_i) m_i + s_i (T(u) eq(m_c))_i; the conserved rows take their
    relative values mr_c (restore_conserved_moments); [source dt/2]; then
    ma = T(-u) m, whose M^-1 gives the new populations.

    Returns (lines, names) with names[i] the C name of ma_i."""
    ns, nc = spec.ns, spec.nconsm
    msyms = spec.msyms
    out = []
    for d, e in enumerate(spec.rel_vel):
        out.append("const C ru%d = %s;" % (d, _ccode(e)))
    Tu = [[sp.sympify(c) for c in r] for r in spec.Tu]
    Tmu = [[sp.sympify(c) for c in r] for r in spec.Tmu]

    def rows(mat, vec, which, prefix):
        exprs = [sum((mat[i][j] * vec[j] for j in range(ns) if mat[i][j] != 0), sp.Integer(0)) for i in which]
        repl, red = sp.cse(exprs, symbols=sp.numbered_symbols(prefix), optimizations="basic")
        lines = ["const C %s = %s;" % (sym, _ccode(e)) for sym, e in repl]
        return lines, red

    # relative moments of the pulled populations
    lines, red = rows(Tu, msyms, range(ns), "tr_")
    out += lines
    out += ["const C mr%d = %s;" % (i, _ccode(e)) for i, e in zip(range(ns), red)]
    out += ["m%d = mr%d;" % (i, i) for i in range(nc, ns)]

    def sources():
        for idx, expr in spec.sources:
            out.append("m%d = m%d + half_dt * (%s);" % (idx, idx, _ccode(sp.sympify(expr))))

    sources()
    relaxed = [i for i in range(nc, ns) if spec.s[i] != 0]
    eqv = [msyms[j] if j < nc else sp.sympify(spec.eq[j]) for j in range(ns)]
    lines, red = rows(Tu, eqv, relaxed, "te_")
    out += lines
    for i, e in zip(relaxed, red):
        out.append("m%d = one_minus[%d] * m%d + omega[%d] * (%s);" % (i, i, i, i, _ccode(e)))
    out += ["m%d = mr%d;" % (i, i) for i in range(nc)]
    sources()
    lines, red = rows(Tmu, msyms, range(ns), "ta_")
    out += lines
    out += ["const C ma%d = %s;" % (i, _ccode(e)) for i, e in zip(range(ns), red)]
    return out, ["ma%d" % i for i in range(ns)]


def _equilibrium_lines(spec: KernelSpec):
    ns, nc = spec.ns, spec.nconsm
    rows = list(range(nc, ns))
    repl, red = sp.cse([spec.eq[i] for i in rows], symbols=sp.numbered_symbols("q_"),
                       optimizations="basic")
    out = ["const C %s = %s;" % (s, _ccode(e)) for s, e in repl]
    for i, e in zip(rows, red):
        out.append("m%d = %s;" % (i, _ccode(e)))
    return out


# ------------------------------------------------------------------ template
_TEMPLATE = r"""// Generated by pylbm_amd.generator.hip for scheme %(hash)s -- do not edit.
// ns=%(ns)d nconsm=%(nc)d dim=%(dim)d pair-factored=%(paired)s
// Population stores bypass the caches (each value is read once, in the next step).
#ifndef LBM_NT_STORES
#define LBM_NT_STORES 1
#endif
#include "lbm_device.h"

using namespace lbm;

#define NS %(ns)d
#define BBSUB_OK %(bbsub_ok)d  // the step can substitute fused bounce-back pulls
#define NCONSM %(nc)d
#define NCREF (NCONSM > 0 ? NCONSM : 1)
#define DIM %(dim)d
#define TB_UNIT_VELOCITIES %(unit_vel)d  // every velocity component in {-1, 0, 1}
#ifndef TB_THREADS
#define TB_THREADS 1024                // threads of a temporally blocked workgroup
#endif
#define TB_MAX_NT 16                   // steps per temporally blocked launch, at most
#define TB_MASK %(bbsub_ok)d                    // folded links possible: mask words kept in LDS

namespace {

template <typename C>
struct StepArgs {
  int64_t src_off[NS];   // pull source offset of population k relative to the cell
  int64_t dst_off[NS];   // k * pop_stride
  int32_t nin[3];        // interior extents
  int32_t halo[3];
  int32_t n1, rs, lead, remap;     // rs: elements between consecutive rows; remap: XCD-aware order
  const uint32_t* bbmask;          // NULL or folded link directions per interior cell
  const uint32_t* bbtab_bits;      // NULL or the folded links with a sign / value table entry
  const uint32_t* bbtab_base;
  const uint32_t* bbtab_src;
  const C* bbtab_sign;
  const C* bbtab_rhs;
  const uint32_t* bbtab_rel;       // NULL or the table sources relative to the pulling cell
  const uint32_t* bbtab_src2;      // NULL or second sources (0xffffffff: a one-source entry)
  const C* bbtab_c2;
  const uint32_t* bbtab_rel2;
  C omega[NS], one_minus[NS], wshift[NS];
  C cref[NCREF];
  C half_dt, t;
  C xorg[3], dx;         // cell centre of reference axis d: xorg[d] + dx * (halo-inclusive index)
};

// Launch shape: LBM_BX threads along the contiguous axis x block_rows<S>() rows
// per workgroup (LBM_BY when set above 1), each thread updating LBM_MROWS
// consecutive rows of its column.
#ifndef LBM_BX
#define LBM_BX 256
#endif
#ifndef LBM_BY
#define LBM_BY 1
#endif
#ifndef LBM_MROWS
#define LBM_MROWS 1
#endif

// Cells per thread along the contiguous axis of the vector instantiation (vector
// loads / stores): 0 = by element size: 4 cells for the 4-byte formats (16 B per
// lane) and the 2-byte formats (8 B per lane, pulled rows at odd offsets assembled
// by vload_pull), 1 for FP64 (measured no gain on MI355X).
#ifndef LBM_VEC
#define LBM_VEC 0
#endif
#ifndef LBM_FORCE_I64
#define LBM_FORCE_I64 0
#endif
// whether the launch picks the vector instantiation by itself for the 2-byte formats
#ifndef LBM_VEC16_AUTO
#define LBM_VEC16_AUTO 0
#endif
// Rows of a fused-step workgroup: LBM_BY, or by default 2 for FP64 and 1 for the
// other formats (MI355X, 256^3 D3Q19: FP64 0.913 ms per step with 2 rows against
// 0.921 with 1; FP32 0.458 against 0.449)
template <typename S>
__host__ __device__ constexpr int block_rows() {
  return LBM_BY > 1 ? LBM_BY : (sizeof(S) == 8 ? 2 : 1);
}

template <typename S>
constexpr int vec_cells() {
  return LBM_VEC > 0 ? LBM_VEC : (sizeof(S) <= 4 ? 4 : 1);
}

// Workgroup order: consecutive workgroups are dealt round-robin to the 8 XCDs of
// MI355X.  With remap the linear workgroup index is renumbered so that each XCD
// takes a contiguous range of (mid, outer) rows, whose pulled neighbour rows then
// hit the same L2 (only when the count is a multiple of 8).  A run-time choice
// (lbm_params.xcd_remap; by default on for the halo path, off for the wrap path).
__device__ __forceinline__ void block_coords(unsigned& bx, unsigned& by, unsigned& bz, int remap) {
  bx = blockIdx.x;
  by = blockIdx.y;
  bz = blockIdx.z;
  if (remap) {
    const unsigned total = gridDim.x * gridDim.y * gridDim.z;
    if (total %% 8 == 0) {
      const unsigned lin = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
      const unsigned logical = (lin %% 8) * (total / 8) + lin / 8;
      bx = logical %% gridDim.x;
      by = (logical / gridDim.x) %% gridDim.y;
      bz = logical / (gridDim.x * gridDim.y);
    }
  }
}

// I: the element index type, int32_t when every index of the array fits it (the
// usual case: 32-bit offsets from a 64-bit base), int64_t for larger arrays
template <typename S, bool FOLD, int VEC, typename I>
__global__ void __launch_bounds__(LBM_BX * block_rows<S>())
fused_step(const S* __restrict__ f, S* __restrict__ fn, const StepArgs<typename Codec<S>::C> a) {
  using C = typename Codec<S>::C;
  // full-value arithmetic for FP64 (as pylbm), deviations from the rest state otherwise
  constexpr bool DEV = Codec<S>::kind != LBM_F64;
  unsigned bkx, bky, bkz;
  block_coords(bkx, bky, bkz, a.remap);
  const int i = (bkx * blockDim.x + threadIdx.x) * VEC;
  if (i >= a.nin[2]) return;
  const int oo = bkz;
  const int mbase = (bky * blockDim.y + threadIdx.y) * LBM_MROWS;
  const C half_dt = a.half_dt;
  const C tt = a.t;
  (void)half_dt; (void)tt;
  const C* one_minus = a.one_minus;
  const C* omega = a.omega;
  const C* cref = a.cref;
  (void)cref;
#pragma unroll
  for (int r = 0; r < LBM_MROWS; ++r) {
  const int mm = mbase + r;
  if (mm >= a.nin[1]) return;
  const int o = oo + a.halo[0], m = mm + a.halo[1], ii = i + a.halo[2];
  const I cell = ((I)o * a.n1 + m) * a.rs + a.lead + ii;
%(loads)s
%(bbsub)s
%(gdecl)s
#pragma unroll
  for (int j = 0; j < VEC; ++j) {
%(coords)s
%(decode)s
%(f2m)s
  if constexpr (DEV) {
%(coll_dev)s
%(m2f_dev)s
%(stores_dev)s
  } else {
%(coll)s
%(m2f)s
%(stores)s
  }
  }
%(vstores)s
  }
}

%(tb_kernel)stemplate <typename S>
__global__ void __launch_bounds__(256)
f2m_kernel(const S* __restrict__ f, lbm_layout fl, double* __restrict__ mo, lbm_layout ml,
           const StepArgs<double> a) {
  using C = double;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= fl.n[2]) return;
  const int64_t m_ = blockIdx.y, o_ = blockIdx.z;
  const int64_t cf = cell_offset(fl, o_, m_, i);
  const int64_t cm = cell_offset(ml, o_, m_, i);
%(f2m_loads)s
%(f2m_full)s
%(f2m_stores)s
}

template <typename S>
__global__ void __launch_bounds__(256)
m2f_kernel(const double* __restrict__ mi, lbm_layout ml, S* __restrict__ f, lbm_layout fl,
           const StepArgs<double> a) {
  using C = double;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= fl.n[2]) return;
  const int64_t m_ = blockIdx.y, o_ = blockIdx.z;
  const int64_t cf = cell_offset(fl, o_, m_, i);
  const int64_t cm = cell_offset(ml, o_, m_, i);
%(m2f_loads)s
%(m2f_full)s
%(m2f_stores)s
}

__global__ void __launch_bounds__(256)
equilibrium_kernel(double* __restrict__ mo, lbm_layout ml) {
  using C = double;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ml.n[2]) return;
  const int64_t m_ = blockIdx.y, o_ = blockIdx.z;
  const int64_t cm = cell_offset(ml, o_, m_, i);
%(eq_loads)s
%(eq_body)s
%(eq_stores)s
}

// pylbm's split routines, generated from the same spec: transport
// (pylbm/algorithm/base.py:273-296: fnew[k, x] = f[k, x - v_k] on the interior, a
// copy of the stored words), relaxation (404-437: m_i <- (1 - s_i) m_i + s_i eq_i
// on the whole halo-inclusive domain, FP64 moments) and source_term (439-504 with
// algorithm/ode.py:11-16: one explicit Euler half step m_c <- m_c + dt/2 S(m) on
// the interior, FP64 moments)
template <typename S>
__global__ void __launch_bounds__(256)
transport_kernel(const S* __restrict__ f, S* __restrict__ fn, const StepArgs<typename Codec<S>::C> a) {
  using B = typename Bits<S>::T;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nin[2]) return;
  const int64_t cell = ((int64_t)(blockIdx.z + a.halo[0]) * a.n1 + blockIdx.y + a.halo[1]) * a.rs + a.lead + i +
                       a.halo[2];
  const B* fb = reinterpret_cast<const B*>(f);
  B* gb = reinterpret_cast<B*>(fn);
#pragma unroll
  for (int k = 0; k < NS; ++k) gb[a.dst_off[k] + cell] = fb[a.src_off[k] + cell];
}

__global__ void __launch_bounds__(256)
relaxation_kernel(double* __restrict__ mo, lbm_layout ml, const StepArgs<double> a) {
  using C = double;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= ml.n[2]) return;
  const int64_t m_ = blockIdx.y, o_ = blockIdx.z;
  const int64_t cm = cell_offset(ml, o_, m_, i);
  const C* one_minus = a.one_minus;
  const C* omega = a.omega;
  (void)one_minus; (void)omega;
%(rx_loads)s
%(rx_body)s
%(rx_stores)s
}

__global__ void __launch_bounds__(256)
source_term_kernel(double* __restrict__ mo, lbm_layout ml, const StepArgs<double> a) {
  using C = double;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.nin[2]) return;
  const int o = (int)blockIdx.z + a.halo[0], m = (int)blockIdx.y + a.halo[1], ii = i + a.halo[2];
  (void)o; (void)m;
  constexpr int j = 0;
  (void)j;
  const int64_t cm = cell_offset(ml, o, m, ii);
  const C half_dt = a.half_dt;
  const C tt = a.t;
  (void)half_dt; (void)tt;
%(st_coords)s
%(st_loads)s
%(st_body)s
%(st_stores)s
}

// A value the compiler must treat as opaque: an equilibrium moment used as if it
// had been stored and loaded again, so that none of its products is contracted
// into an FMA of m2f's sums (which would round differently from the two-kernel path).
__device__ __forceinline__ double stored_moment(double x) {
  asm("" : "+v"(x));
  return x;
}

// equilibrium from the conserved moments followed by m2f, the moments kept in
// registers: the same expressions as equilibrium_kernel then m2f_kernel, so the
// same populations (the non-conserved rows of mo are not written)
template <typename S>
__global__ void __launch_bounds__(256)
eq_m2f_kernel(const double* __restrict__ mo, lbm_layout ml, S* __restrict__ f, lbm_layout fl,
              const StepArgs<double> a) {
  using C = double;
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= fl.n[2]) return;
  const int64_t m_ = blockIdx.y, o_ = blockIdx.z;
  const int64_t cf = cell_offset(fl, o_, m_, i);
  const int64_t cm = cell_offset(ml, o_, m_, i);
%(eqm2f_loads)s
%(eq_body)s
%(eqm2f_opaque)s
%(m2f_full)s
%(m2f_stores)s
}

template <typename C>
static int make_args(const lbm_layout* L, const lbm_params* p, StepArgs<C>* a) {
  int rc = check_layout(L);
  if (rc) return rc;
  LBM_REQUIRE(L->ns == NS, "layout has ns=%%d, module was generated for %%d", L->ns, NS);
  static const int hv[NS][3] = {%(vel)s};
  for (int k = 0; k < NS; ++k) {
    a->dst_off[k] = k * L->pop_stride;
    a->src_off[k] = k * L->pop_stride - ((int64_t)hv[k][0] * L->n[1] + hv[k][1]) * L->row_stride - hv[k][2];
  }
  for (int d = 0; d < 3; ++d) {
    a->nin[d] = (int32_t)(L->n[d] - 2 * L->halo[d]);
    a->halo[d] = L->halo[d];
  }
  a->n1 = (int32_t)L->n[1];
  a->rs = (int32_t)L->row_stride;
  a->lead = (int32_t)L->lead;
  a->remap = 0;
  for (int k = 0; k < NS; ++k) {
    a->omega[k] = p ? (C)p->omega[k] : (C)0;
    a->one_minus[k] = p ? (C)p->one_minus[k] : (C)1;
    a->wshift[k] = p ? (C)p->wshift[k] : (C)0;
  }
  for (int c = 0; c < NCREF; ++c) a->cref[c] = p ? (C)p->cref[c] : (C)0;
  a->half_dt = p ? (C)(0.5 * p->dt) : (C)0;
  a->t = p ? (C)p->t : (C)0;
  for (int d = 0; d < 3; ++d) a->xorg[d] = p ? (C)p->origin[d] : (C)0;
  a->dx = p ? (C)p->dx : (C)0;
  a->bbmask = p ? p->bbmask : nullptr;
  a->bbtab_bits = p ? p->bbtab_bits : nullptr;
  a->bbtab_base = p ? p->bbtab_base : nullptr;
  a->bbtab_src = p ? p->bbtab_src : nullptr;
  a->bbtab_sign = p ? (const C*)p->bbtab_sign : nullptr;
  a->bbtab_rhs = p ? (const C*)p->bbtab_rhs : nullptr;
  a->bbtab_rel = p ? p->bbtab_rel : nullptr;
  a->bbtab_src2 = p ? p->bbtab_src2 : nullptr;
  a->bbtab_c2 = p ? (const C*)p->bbtab_c2 : nullptr;
  a->bbtab_rel2 = p ? p->bbtab_rel2 : nullptr;
  LBM_REQUIRE(a->bbmask == nullptr || BBSUB_OK, "this stencil has no symmetric pairs: no fused bounce-back");
  return 0;
}

static unsigned block_x(int64_t n) {
  return n >= 256 ? 256u : (n >= 128 ? 128u : 64u);
}

template <typename S>
static int step_launch(const void* f, void* fn, const lbm_layout* L, const lbm_params* p, int fold,
                       hipStream_t s) {
  using C = typename Codec<S>::C;
  StepArgs<C> a;
  int rc = make_args<C>(L, p, &a);
  if (rc) return rc;
  constexpr int V = vec_cells<S>();
  // V cells per thread only where that still leaves >= 2^20 threads: small grids
  // (a 512^2 plane) need the waves of one cell per thread to hide memory latency.
  // Only the wrap path without folded links takes V cells: with a mask word per
  // cell the per-cell mask branches of a V-cell thread cost more than the wider
  // loads save, and the halo path's pulls straddle more cache lines (MI355X, 512 x
  // 256 x 256 D3Q19 sphere packing FP32: 0.90 ms per step with one cell per thread,
  // 0.93 with 2, 0.98 with 4; 256^3 halo path FP32 with its ghost fill: 0.514 ms
  // with one, 0.534 with 4).
  const int64_t cells = (int64_t)a.nin[0] * a.nin[1] * a.nin[2];
  const int want = p ? p->vec_cells : 0;
  LBM_REQUIRE(want == 0 || want == 1 || (want == V && a.nin[2] %% V == 0),
              "vec_cells %%d not available (%%d cells per thread, contiguous extent %%d)", want, V, a.nin[2]);
  const bool auto_vec = (sizeof(S) != 2 || LBM_VEC16_AUTO) && fold && a.bbmask == nullptr;
  const int vec = want ? want : ((auto_vec && a.nin[2] %% V == 0 && cells / V >= ((int64_t)1 << 20)) ? V : 1);
  a.remap = (p && p->xcd_remap >= 0) ? (p->xcd_remap != 0) : !fold;
  const int threads = a.nin[2] / vec;
  unsigned bx = LBM_BX;
  while (bx > 64 && (int)(bx / 2) >= threads) bx /= 2;
  const unsigned rows = block_rows<S>() * LBM_MROWS;
  dim3 block(bx, block_rows<S>());
  dim3 grid((unsigned)((threads + bx - 1) / bx), (unsigned)((a.nin[1] + rows - 1) / rows),
            (unsigned)a.nin[0]);
  if (fold) {
    for (int d = 0; d < 3; ++d)
      LBM_REQUIRE(a.nin[d] >= a.halo[d], "periodic fold needs extents >= halo");
  }
  if (layout_fits_int32(*L) && !LBM_FORCE_I64) {
    if (vec == V) {
      if (fold) fused_step<S, true, V, int32_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
      else fused_step<S, false, V, int32_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
    } else {
      if (fold) fused_step<S, true, 1, int32_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
      else fused_step<S, false, 1, int32_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
    }
  } else {
    if (vec == V) {
      if (fold) fused_step<S, true, V, int64_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
      else fused_step<S, false, V, int64_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
    } else {
      if (fold) fused_step<S, true, 1, int64_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
      else fused_step<S, false, 1, int64_t><<<grid, block, 0, s>>>((const S*)f, (S*)fn, a);
    }
  }
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

template <typename S>
static int f2m_launch(const void* f, const lbm_layout* fl, double* mo, const lbm_layout* ml,
                      const lbm_params* p, hipStream_t s) {
  StepArgs<double> a;
  int rc = make_args<double>(fl, p, &a);
  if (rc) return rc;
  const unsigned bx = block_x(fl->n[2]);
  dim3 grid((unsigned)((fl->n[2] + bx - 1) / bx), (unsigned)fl->n[1], (unsigned)fl->n[0]);
  f2m_kernel<S><<<grid, bx, 0, s>>>((const S*)f, *fl, mo, *ml, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

template <typename S>
static int m2f_launch(const double* mi, const lbm_layout* ml, void* f, const lbm_layout* fl,
                      const lbm_params* p, hipStream_t s) {
  StepArgs<double> a;
  int rc = make_args<double>(fl, p, &a);
  if (rc) return rc;
  const unsigned bx = block_x(fl->n[2]);
  dim3 grid((unsigned)((fl->n[2] + bx - 1) / bx), (unsigned)fl->n[1], (unsigned)fl->n[0]);
  m2f_kernel<S><<<grid, bx, 0, s>>>(mi, *ml, (S*)f, *fl, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

template <typename S>
static int transport_launch(const void* f, void* fn, const lbm_layout* fl, const lbm_params* p, hipStream_t s) {
  StepArgs<typename Codec<S>::C> a;
  int rc = make_args<typename Codec<S>::C>(fl, p, &a);
  if (rc) return rc;
  const unsigned bx = block_x(a.nin[2]);
  dim3 grid((unsigned)((a.nin[2] + bx - 1) / bx), (unsigned)a.nin[1], (unsigned)a.nin[0]);
  transport_kernel<S><<<grid, bx, 0, s>>>((const S*)f, (S*)fn, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

template <typename S>
static int eq_m2f_launch(const double* mi, const lbm_layout* ml, void* f, const lbm_layout* fl,
                         const lbm_params* p, hipStream_t s) {
  StepArgs<double> a;
  int rc = make_args<double>(fl, p, &a);
  if (rc) return rc;
  const unsigned bx = block_x(fl->n[2]);
  dim3 grid((unsigned)((fl->n[2] + bx - 1) / bx), (unsigned)fl->n[1], (unsigned)fl->n[0]);
  eq_m2f_kernel<S><<<grid, bx, 0, s>>>(mi, *ml, (S*)f, *fl, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

%(tb_launch)sstatic int same_extents(const lbm_layout* a, const lbm_layout* b) {
  LBM_REQUIRE(a && b, "layout is NULL");
  LBM_REQUIRE(b->storage == LBM_F64, "moment layout must be FP64");
  for (int d = 0; d < 3; ++d)
    LBM_REQUIRE(a->n[d] == b->n[d], "moment / population extents differ on axis %%d", d);
  return check_layout(b);
}

}  // namespace

extern "C" {

int lbm_gen_info(int32_t* ns, int32_t* nconsm, int32_t* dim) {
  *ns = NS;
  *nconsm = NCONSM;
  *dim = DIM;
  return 0;
}

int lbm_gen_step(const void* d_f, void* d_fnew, const lbm_layout* layout, const lbm_params* params,
                 int periodic_fold, void* stream) {
  LBM_REQUIRE(layout, "layout is NULL");
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(layout->storage, step_launch, d_f, d_fnew, layout, params, periodic_fold, s);
}

int lbm_gen_f2m(const void* d_f, const lbm_layout* fl, double* d_m, const lbm_layout* ml,
                const lbm_params* params, void* stream) {
  int rc = same_extents(fl, ml);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(fl->storage, f2m_launch, d_f, fl, d_m, ml, params, s);
}

int lbm_gen_m2f(const double* d_m, const lbm_layout* ml, void* d_f, const lbm_layout* fl,
                const lbm_params* params, void* stream) {
  int rc = same_extents(fl, ml);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(fl->storage, m2f_launch, d_m, ml, d_f, fl, params, s);
}

int lbm_gen_eq_m2f(const double* d_m, const lbm_layout* ml, void* d_f, const lbm_layout* fl,
                   const lbm_params* params, void* stream) {
  int rc = same_extents(fl, ml);
  if (rc) return rc;
  LBM_REQUIRE(ml->storage == LBM_F64 && ml->ns == NS, "eq_m2f needs FP64 moments of ns=%%d", NS);
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(fl->storage, eq_m2f_launch, d_m, ml, d_f, fl, params, s);
}

int lbm_gen_equilibrium(double* d_m, const lbm_layout* ml, void* stream) {
  int rc = check_layout(ml);
  if (rc) return rc;
  LBM_REQUIRE(ml->storage == LBM_F64 && ml->ns == NS, "equilibrium needs FP64 moments of ns=%%d", NS);
  const unsigned bx = block_x(ml->n[2]);
  dim3 grid((unsigned)((ml->n[2] + bx - 1) / bx), (unsigned)ml->n[1], (unsigned)ml->n[0]);
  equilibrium_kernel<<<grid, bx, 0, (hipStream_t)stream>>>(d_m, *ml);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_gen_transport(const void* d_f, void* d_fnew, const lbm_layout* fl, const lbm_params* params,
                      void* stream) {
  LBM_REQUIRE(d_f != d_fnew, "transport needs distinct source and destination arrays");
  hipStream_t s = (hipStream_t)stream;
  LBM_DISPATCH_STORAGE(fl ? fl->storage : -1, transport_launch, d_f, d_fnew, fl, params, s);
}

int lbm_gen_relaxation(double* d_m, const lbm_layout* ml, const lbm_params* params, void* stream) {
  LBM_REQUIRE(params != nullptr, "NULL params");
  StepArgs<double> a;
  int rc = make_args<double>(ml, params, &a);
  if (rc) return rc;
  LBM_REQUIRE(ml->storage == LBM_F64, "relaxation needs FP64 moments");
  const unsigned bx = block_x(ml->n[2]);
  dim3 grid((unsigned)((ml->n[2] + bx - 1) / bx), (unsigned)ml->n[1], (unsigned)ml->n[0]);
  relaxation_kernel<<<grid, bx, 0, (hipStream_t)stream>>>(d_m, *ml, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_gen_source_term(double* d_m, const lbm_layout* ml, const lbm_params* params, void* stream) {
  LBM_REQUIRE(params != nullptr, "NULL params");
  StepArgs<double> a;
  int rc = make_args<double>(ml, params, &a);
  if (rc) return rc;
  LBM_REQUIRE(ml->storage == LBM_F64, "source_term needs FP64 moments");
  const unsigned bx = block_x(a.nin[2]);
  dim3 grid((unsigned)((a.nin[2] + bx - 1) / bx), (unsigned)a.nin[1], (unsigned)a.nin[0]);
  source_term_kernel<<<grid, bx, 0, (hipStream_t)stream>>>(d_m, *ml, a);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

int lbm_gen_run_blocked(void* d_f, void* d_fnew, const lbm_layout* layout, lbm_params* params,
                        int64_t nsteps, int steps_per_launch, int tile_m, int tile_i, int* parity,
                        void* stream) {
  LBM_REQUIRE(params && layout && parity, "NULL argument");
  LBM_REQUIRE(DIM == 2, "temporal blocking is generated for 2D schemes only (this one is %%dD)", DIM);
  LBM_REQUIRE(TB_UNIT_VELOCITIES, "temporal blocking needs velocity components in {-1, 0, 1}");
  LBM_REQUIRE(steps_per_launch >= 1 && steps_per_launch <= TB_MAX_NT, "steps per launch out of [1, %%d]",
              TB_MAX_NT);
  LBM_REQUIRE(tile_m >= 1 && tile_i >= 1, "tile sizes must be positive");
  LBM_REQUIRE(params->bbtab_bits == nullptr || params->bbtab_rel != nullptr,
              "folded links with table entries need their relative sources (bbtab_rel)");
  LBM_REQUIRE(params->bbtab_src2 == nullptr || params->bbtab_rel2 != nullptr,
              "folded interpolated links need their relative second sources (bbtab_rel2)");
  hipStream_t s = (hipStream_t)stream;
  void* a = d_f;
  void* b = d_fnew;
  int64_t done = 0;
  const int64_t blocks = nsteps / steps_per_launch;
  for (int64_t it = 0; it < blocks; ++it) {
    int rc = 0;
    LBM_DISPATCH_BLOCKED(layout->storage, tb_launch, a, b, layout, params, steps_per_launch, tile_m, tile_i, s, rc);
    if (rc) return rc;
    void* tmp = a;
    a = b;
    b = tmp;
    done += steps_per_launch;
  }
  for (; done < nsteps; ++done) {
    int rc = lbm_gen_step(a, b, layout, params, 1, stream);
    if (rc) return rc;
    void* tmp = a;
    a = b;
    b = tmp;
    params->t += params->dt;
  }
  *parity = (int)(blocks & 1) ^ (int)((nsteps - blocks * steps_per_launch) & 1);
  return 0;
}

int lbm_gen_run(void* d_f, void* d_fnew, const lbm_layout* layout, const lbm_velocities* vel,
                lbm_params* params, const lbm_link_set* links, int nmeth, int64_t nsteps,
                int ghost, int periodic_fold, int* parity, void* stream) {
  LBM_REQUIRE(params && layout && parity, "NULL argument");
  LBM_REQUIRE(!(periodic_fold && nmeth > 0), "periodic fold is only valid without link lists");
  void* a = d_f;
  void* b = d_fnew;
  for (int64_t it = 0; it < nsteps; ++it) {
    int rc = 0;
    if (ghost == 1) rc = lbm_ghost_fill(a, layout, vel, stream);
    else if (ghost == 2) rc = lbm_ghost_fill_all(a, layout, vel, stream);
    if (rc) return rc;
    if (nmeth > 0) rc = lbm_links_apply_sets(a, layout, links, nmeth, stream);
    if (rc) return rc;
    rc = lbm_gen_step(a, b, layout, params, periodic_fold, stream);
    if (rc) return rc;
    void* tmp = a;
    a = b;
    b = tmp;
    params->t += params->dt;
  }
  *parity = (int)(nsteps & 1);
  return 0;
}

}  // extern "C"
"""


_TB_KERNEL = r"""// ------------------------------------------------ temporally blocked steps (2D)
// One workgroup advances a tm x ti tile of the interior (mid x inner axes) by nt
// steps in LDS: it loads the (tm + 2 nt) x (ti + 2 nt) region around the tile
// through the periodic wrap of the interior (valid where the wrap path is: a
// periodic box, or every link list folded into the mask / table), then for
// s = 1 .. nt computes the cells of the region shrunk by s from the values of
// step s - 1 with the per-cell arithmetic of fused_step (the same generated
// lines, so the same bits), and finally stores the tile.  Folded links read their
// sources at the offsets bbtab_rel gives relative to the pulling cell.
template <typename C>
struct TbArgs {
  int32_t nt, tm, ti, tiles_i;
  C times[TB_MAX_NT];   // t of each step (the host's sequence of t += dt)
};

// x in [-n, 3n) -> x mod n, without a division (tile regions reach at most nt < n
// cells before the interior and less than 2n + nt after its start)
__device__ __forceinline__ int wrap_near(int x, int n) {
  x += (x < 0) ? n : 0;
  x -= (x >= n) ? n : 0;
  x -= (x >= n) ? n : 0;
  return x;
}

// Population planes of the region lie TB_PLANE<S> cells apart: a compile-time
// stride, so that the plane offsets of the pulls and stores become immediates of
// the LDS instructions.  It is the most cells the two copies of the populations
// plus the mask / table words fit in 160 KiB with, and keeps the last plane within
// the 16-bit immediate offset.
template <typename S>
__host__ __device__ constexpr int tb_plane() {
  constexpr int a = (160 * 1024) / (2 * NS * (int)sizeof(S) + 12);
  constexpr int b = NS > 1 ? 65528 / ((NS - 1) * (int)sizeof(S)) : a;
  return a < b ? a : b;
}

template <typename S>
__global__ void __launch_bounds__(TB_THREADS)
tb_step(const S* __restrict__ f, S* __restrict__ fn, const StepArgs<typename Codec<S>::C> a,
        const TbArgs<typename Codec<S>::C> t) {
  using C = typename Codec<S>::C;
  constexpr bool DEV = Codec<S>::kind != LBM_F64;
  extern __shared__ uint4 tb_smem[];
  const int RW = t.tm + 2 * t.nt, RI = t.ti + 2 * t.nt, RN = RW * RI;
  constexpr int RC = tb_plane<S>();   // plane stride (cells), RN <= RC
  S* cur = reinterpret_cast<S*>(tb_smem);
  S* nxt = cur + NS * RC;
  // constant over the launch: the region's mask words (folded links) and, with
  // table links, every region cell's table bits and first entry (loaded with the
  // populations, so a step reads its entries after one round trip instead of two)
  uint32_t* msk = reinterpret_cast<uint32_t*>(cur + 2 * NS * RC);
  uint32_t* tbits = msk + RN;
  uint32_t* tbase_l = tbits + RN;
  const bool table = TB_MASK && a.bbmask && a.bbtab_bits;
  (void)msk; (void)tbits; (void)tbase_l; (void)table;
  const int m0 = (int)(blockIdx.x / t.tiles_i) * t.tm - t.nt;
  const int i0 = (int)(blockIdx.x %% t.tiles_i) * t.ti - t.nt;
  const C half_dt = a.half_dt;
  (void)half_dt;
  const C* one_minus = a.one_minus;
  const C* omega = a.omega;
  const C* cref = a.cref;
  (void)cref;
  for (int c = threadIdx.x; c < RN; c += blockDim.x) {
    const int lr = c / RI;
    const int gm = wrap_near(m0 + lr, a.nin[1]), gi = wrap_near(i0 + c - lr * RI, a.nin[2]);
    const int64_t src = (int64_t)(a.halo[0] * a.n1 + gm + a.halo[1]) * a.rs + a.lead + gi + a.halo[2];
#pragma unroll
    for (int k = 0; k < NS; ++k) cur[k * RC + c] = f[a.dst_off[k] + src];
    if (TB_MASK && a.bbmask) {
      const int64_t mc = (int64_t)gm * a.nin[2] + gi;
      msk[c] = a.bbmask[mc];
      if (table) {
        tbits[c] = a.bbtab_bits[mc];
        tbase_l[c] = a.bbtab_base[mc];
      }
    }
  }
  __syncthreads();
  for (int s = 1; s <= t.nt; ++s) {
    const C tt = t.times[s - 1];
    (void)tt;
    const int WC = RI - 2 * s, HC = RW - 2 * s;
    // the cells of the shrunk region in row-major order, blockDim.x apart: the
    // (row, column) of the next one follow by adding (dq, dr) with a carry, and its
    // region index by adding dl, plus dc on a carry (no multiply per cell)
    const int dq = (int)blockDim.x / WC, dr = (int)blockDim.x - dq * WC;
    int rr = (int)threadIdx.x / WC, qq = (int)threadIdx.x - rr * WC;
    const int dl = dq * RI + dr, dc = RI - WC;
    int lnext = (s + rr) * RI + s + qq;
    for (int c = threadIdx.x; c < WC * HC; c += blockDim.x) {
      const int lr = s + rr, lc = s + qq;
      const int lcell = lnext;
      qq += dr;
      rr += dq;
      lnext += dl;
      if (qq >= WC) {
        qq -= WC;
        ++rr;
        lnext += dc;
      }
      const int gm = wrap_near(m0 + lr, a.nin[1]), gi = wrap_near(i0 + lc, a.nin[2]);
      const int o = a.halo[0], m = gm + a.halo[1], ii = gi + a.halo[2];
      (void)o; (void)m; (void)ii;
      constexpr int j = 0;
%(tb_loads)s
%(tb_sub)s
%(tb_gdecl)s
      {
%(tb_coords)s
%(tb_decode)s
%(tb_f2m)s
      if constexpr (DEV) {
%(tb_coll_dev)s
%(tb_m2f_dev)s
%(tb_stores_dev)s
      } else {
%(tb_coll)s
%(tb_m2f)s
%(tb_stores)s
      }
      }
%(tb_lds_stores)s
    }
    __syncthreads();
    S* tmp = cur;
    cur = nxt;
    nxt = tmp;
  }
  for (int c = threadIdx.x; c < t.tm * t.ti; c += blockDim.x) {
    const int lr = t.nt + c / t.ti, lc = t.nt + c %% t.ti;
    const int gm = m0 + lr, gi = i0 + lc;
    if (gm >= a.nin[1] || gi >= a.nin[2]) continue;
    const int64_t dst = (int64_t)(a.halo[0] * a.n1 + gm + a.halo[1]) * a.rs + a.lead + gi + a.halo[2];
#pragma unroll
    for (int k = 0; k < NS; ++k) fn[a.dst_off[k] + dst] = cur[k * RC + lr * RI + lc];
  }
}

"""

_TB_LAUNCH = r"""// one launch of tb_step: nt steps of every tile; advances params->t by nt dt
template <typename S>
static int tb_launch(void* f, void* fn, const lbm_layout* L, lbm_params* p, int nt, int tm, int ti,
                     hipStream_t s) {
  using C = typename Codec<S>::C;
  StepArgs<C> a;
  int rc = make_args<C>(L, p, &a);
  if (rc) return rc;
  LBM_REQUIRE(a.nin[1] >= tm && a.nin[2] >= ti && a.nin[1] >= nt && a.nin[2] >= nt,
              "tile %%d x %%d / %%d steps larger than the interior", tm, ti, nt);
  TbArgs<C> t{};
  t.nt = nt;
  t.tm = tm;
  t.ti = ti;
  t.tiles_i = (a.nin[2] + ti - 1) / ti;
  for (int q = 0; q < nt; ++q) {
    t.times[q] = (C)p->t;
    p->t += p->dt;
  }
  const int tiles_m = (a.nin[1] + tm - 1) / tm;
  const size_t region = (size_t)(tm + 2 * nt) * (ti + 2 * nt);
  const bool mask = TB_MASK && p->bbmask, table = mask && p->bbtab_bits;
  LBM_REQUIRE(region <= (size_t)tb_plane<S>(), "tile region of %%zu cells exceeds the %%d cells of a plane",
              region, tb_plane<S>());
  const size_t lds = 2 * NS * sizeof(S) * (size_t)tb_plane<S>() + region * ((mask ? 4 : 0) + (table ? 8 : 0));
  LBM_REQUIRE(lds <= 160 * 1024, "tile needs %%zu bytes of LDS", lds);
  LBM_CHECK_HIP(hipFuncSetAttribute((const void*)tb_step<S>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  tb_step<S><<<dim3((unsigned)(tiles_m * t.tiles_i)), dim3(TB_THREADS), lds, s>>>((const S*)f, (S*)fn, a, t);
  LBM_CHECK_HIP(hipGetLastError());
  return 0;
}

#define LBM_DISPATCH_BLOCKED(kind, FN, f, fn, L, p, nt, tm, ti, s, rc)          \
  do {                                                                          \
    switch (kind) {                                                             \
      case LBM_F64: rc = FN<double>(f, fn, L, p, nt, tm, ti, s); break;         \
      case LBM_F32: rc = FN<float>(f, fn, L, p, nt, tm, ti, s); break;          \
      case LBM_F16D: rc = FN<_Float16>(f, fn, L, p, nt, tm, ti, s); break;      \
      case LBM_BF16D: rc = FN<::lbm::bf16>(f, fn, L, p, nt, tm, ti, s); break;  \
      case LBM_F32D: rc = FN<::lbm::f32dev>(f, fn, L, p, nt, tm, ti, s); break; \
      default: ::lbm::set_error("unknown storage kind %%d", (int)(kind)); rc = -1; \
    }                                                                           \
  } while (0)

"""

# 3D / 1D schemes: no temporally blocked kernel
_TB_LAUNCH_NONE = r"""#define LBM_DISPATCH_BLOCKED(kind, FN, f, fn, L, p, nt, tm, ti, s, rc) \
  do {                                                                \
    (void)(kind);                                                     \
    rc = 0;                                                           \
  } while (0)

"""


def _indent(lines, n=2):
    return "\n".join(" " * n + ln for ln in lines)


def generate_source(spec: KernelSpec) -> str:
    ns, dim = spec.ns, spec.dim
    vel = [_canon(v, dim) for v in spec.velocities]
    ksym = _pairs(spec)
    sig_f2m = _row_parities(spec.M, ksym, by_rows=True) if ksym else None
    sig_m2f = None
    if ksym and sig_f2m:
        sig_cols = _row_parities(spec.invM, ksym, by_rows=False)
        sig_m2f = sig_cols if sig_cols == sig_f2m else None
        if sig_m2f is None:
            sig_f2m = None
    M = [[Fraction(c) for c in r] for r in spec.M]
    iM = [[Fraction(c) for c in r] for r in spec.invM]
    fvars = ["f%d" % k for k in range(ns)]
    mvars = ["m%d" % i for i in range(ns)]

    # ---------------------------------------------------------- fused step
    # folded link lists: where bit k of the cell's mask word is set, the pulled
    # value of population k is what the link f[k, x - v_k] = sign f[ksym, x] + value
    # would have stored, computed as links_kernel computes it: a copy of f[ksym, x]
    # for sign 1 without a value, else decode_dev of the table entry's source,
    # r = sign v, r += value, encode (w = 0); bit 31 of the mask word marks the
    # cells that have table entries (ns <= 31)
    bbsub = []
    if ksym is not None:
        bbsub = ["{",
                 "  if (a.bbmask) {",
                 "    const int64_t mc = ((int64_t)oo * a.nin[1] + mm) * a.nin[2] + i;",
                 "#pragma unroll",
                 "    for (int j = 0; j < VEC; ++j) {",
                 "      const uint32_t bits = a.bbmask[mc + j];",
                 "      if (bits) {",
                 "        const uint32_t tb = (bits & 0x80000000u) ? a.bbtab_bits[mc + j] : 0u;",
                 "        const uint32_t tbase = tb ? a.bbtab_base[mc + j] : 0u;",
                 "        (void)tbase;"]
        for k in range(ns):
            if ksym[k] != k:
                bit = 1 << k
                bbsub += ["        if (bits & %du) {" % bit,
                          "          L%d[j] = f[(I)a.dst_off[%d] + cell + j];" % (k, ksym[k]),
                          "          if (tb & %du) {" % bit,
                          "            using CC = typename Codec<S>::C;",
                          "            const uint32_t e_ = tbase + (uint32_t)__builtin_popcount(tb & %du);" % (bit - 1),
                          "            CC r_ = a.bbtab_sign[e_] * Codec<S>::decode_dev(f[a.bbtab_src[e_]], (CC)0);",
                          "            const uint32_t s2_ = a.bbtab_src2 ? a.bbtab_src2[e_] : 0xffffffffu;",
                          "            if (s2_ != 0xffffffffu) r_ = r_ + a.bbtab_c2[e_] * Codec<S>::decode_dev(f[s2_], (CC)0);",
                          "            r_ += a.bbtab_rhs[e_];",
                          "            L%d[j] = Codec<S>::encode(r_, (CC)0);" % k,
                          "          }",
                          "        }"]
        bbsub += ["      }", "    }", "  }", "}"]
    loads = ["S %s;" % ", ".join("L%d[VEC]" % k for k in range(ns))]
    loads.append("if (!FOLD) {")
    for k in range(ns):
        loads.append("  vload_pull<S, VEC, %d>(L%d, f + ((I)a.src_off[%d] + (%d) + cell));" % (-vel[k][2], k, k, vel[k][2]))
    loads.append("} else {")
    # wrapped outer / mid coordinates per distinct component value; the contiguous
    # axis is wrapped per element, only for the vectors that straddle the row ends
    comps = [sorted({v[d] for v in vel}) for d in range(2)]
    names = {}
    for d, axis in enumerate(("o", "m")):
        for c in comps[d]:
            nm = "w%d_%s" % (d, ("p%d" % c) if c >= 0 else ("n%d" % -c))
            names[(d, c)] = nm
            loads.append("  int %s = %s - (%d);" % (nm, axis, c))
            loads.append("  %s += (%s < a.halo[%d]) ? a.nin[%d] : 0;" % (nm, nm, d, d))
            loads.append("  %s -= (%s >= a.halo[%d] + a.nin[%d]) ? a.nin[%d] : 0;" % (nm, nm, d, d, d))
    for k, (v0, v1, v2) in enumerate(vel):
        loads.append("  {")
        loads.append("    const I row = (I)a.dst_off[%d] + ((I)%s * a.n1 + %s) * a.rs + a.lead;"
                     % (k, names[(0, v0)], names[(1, v1)]))
        loads.append("    const int s0 = ii - (%d);" % v2)
        if v2 == 0:
            loads.append("    vload<S, VEC>(L%d, f + (row + s0));" % k)
        else:
            loads.append("    if (s0 >= a.halo[2] && s0 + VEC <= a.halo[2] + a.nin[2]) {")
            loads.append("      vload_pull<S, VEC, %d>(L%d, f + (row + ii));" % (-v2, k))
            loads.append("    } else {")
            loads.append("      for (int j = 0; j < VEC; ++j) {")
            loads.append("        int c = s0 + j;")
            loads.append("        c += (c < a.halo[2]) ? a.nin[2] : 0;")
            loads.append("        c -= (c >= a.halo[2] + a.nin[2]) ? a.nin[2] : 0;")
            loads.append("        L%d[j] = f[row + c];" % k)
            loads.append("      }")
            loads.append("    }")
        loads.append("  }")
    loads.append("}")
    gdecl = ["S %s;" % ", ".join("G%d[VEC]" % k for k in range(ns))]
    rel = spec.rel_vel is not None
    if rel:
        # relative-velocity schemes collide full values in every format
        decode = ["C f%d = Codec<S>::decode(L%d[j], a.wshift[%d]);" % (k, k, k) for k in range(ns)]
    else:
        decode = ["C f%d = DEV ? Codec<S>::decode_dev(L%d[j], a.wshift[%d]) : Codec<S>::decode(L%d[j], a.wshift[%d]);"
                  % (k, k, k, k, k) for k in range(ns)]
    vstores = ["vstore<S, VEC>(fn + ((I)a.dst_off[%d] + cell), G%d);" % (k, k) for k in range(ns)]

    b1 = _SumBuilder("p")
    f2m_lines = _transform(b1, M, fvars, mvars, ksym, sig_f2m)
    f2m_block = b1.lines + ["C %s = %s;" % (o, e) for o, e in f2m_lines]
    if rel:
        coll, ma_names = _relative_collision_lines(spec)
    else:
        coll, ma_names = _collision_lines(spec, with_sources=True), mvars
    b2 = _SumBuilder("r")
    m2f_lines = _inverse_transform(b2, iM, ma_names, ["g%d" % k for k in range(ns)], ksym, sig_m2f)
    m2f_block = b2.lines + ["const C %s = %s;" % (o, e) for o, e in m2f_lines]
    stores = ["G%d[j] = Codec<S>::encode(g%d, a.wshift[%d]);" % (k, k, k) for k in range(ns)]
    if rel:
        coll_dev, m2f_dev_block, stores_dev = coll, m2f_block, stores
    else:
        coll_dev, dm_names = _deviation_collision_lines(spec)
        bd = _SumBuilder("rd")
        m2f_dev_lines = _inverse_transform(bd, iM, dm_names, ["g%d" % k for k in range(ns)], ksym, sig_m2f,
                                           base=fvars, tag="d")
        m2f_dev_block = bd.lines + ["const C %s = %s;" % (o, e) for o, e in m2f_dev_lines]
        stores_dev = ["G%d[j] = Codec<S>::encode_dev(g%d, a.wshift[%d]);" % (k, k, k) for k in range(ns)]

    # ---------------------------------------------------------- f2m / m2f / eq
    b3 = _SumBuilder("p")
    f2m_full_lines = _transform(b3, M, fvars, mvars, ksym, sig_f2m)
    f2m_loads = ["const C f%d = Codec<S>::widen(f[%d * fl.pop_stride + cf], (double)a.wshift[%d]);" % (k, k, k)
                 for k in range(ns)]
    f2m_full = b3.lines + ["const C %s = %s;" % (o, e) for o, e in f2m_full_lines]
    f2m_stores = ["mo[%d * ml.pop_stride + cm] = m%d;" % (i, i) for i in range(ns)]

    b4 = _SumBuilder("r")
    m2f_full_lines = _inverse_transform(b4, iM, mvars, ["g%d" % k for k in range(ns)], ksym, sig_m2f)
    m2f_loads = ["const C m%d = mi[%d * ml.pop_stride + cm];" % (i, i) for i in range(ns)]
    m2f_full = b4.lines + ["const C %s = %s;" % (o, e) for o, e in m2f_full_lines]
    m2f_stores = ["Codec<S>::store(f, %d * fl.pop_stride + cf, (typename Codec<S>::C)g%d, (typename Codec<S>::C)a.wshift[%d]);" % (k, k, k)
                  for k in range(ns)]

    eq_loads = ["C m%d = mo[%d * ml.pop_stride + cm];" % (i, i) for i in range(ns)]
    eq_body = _equilibrium_lines(spec)
    eq_stores = ["mo[%d * ml.pop_stride + cm] = m%d;" % (i, i) for i in range(spec.nconsm, ns)]
    eq_loads += ["(void)m%d;" % i for i in range(spec.nconsm, ns)]
    # the fused kernel loads only the conserved rows; the equilibrium assigns the others
    eqm2f_loads = (["C m%d = mo[%d * ml.pop_stride + cm];" % (i, i) for i in range(spec.nconsm)]
                   + ["C m%d;" % i for i in range(spec.nconsm, ns)])
    eqm2f_opaque = ["m%d = stored_moment(m%d);" % (i, i) for i in range(spec.nconsm, ns)]

    # split routines: relaxation (every relaxed row) and source_term (the source rows)
    relaxed_rows = [i for i in range(spec.nconsm, ns) if spec.s[i] != 0]
    rx_loads = ["C m%d = mo[%d * ml.pop_stride + cm];" % (i, i) for i in range(ns)]
    rx_loads += ["(void)m%d;" % i for i in range(ns)]
    rx_body = _collision_lines(spec, with_sources=False)
    rx_stores = ["mo[%d * ml.pop_stride + cm] = m%d;" % (i, i) for i in relaxed_rows]
    st_loads = list(rx_loads)
    st_body = ["m%d = m%d + half_dt * (%s);" % (idx, idx, _ccode(sp.sympify(expr))) for idx, expr in spec.sources]
    st_stores = ["mo[%d * ml.pop_stride + cm] = m%d;" % (idx, idx) for idx in sorted({i for i, _ in spec.sources})]

    # cell-centre coordinates of the reference axes the source terms use (pylbm's X, Y, Z)
    used = set()
    for _, e in spec.sources:
        used |= {str(x) for x in sp.sympify(e).free_symbols}
    canon_index = ["o", "m", "(ii + j)"]
    coords = ["const C xc%d = a.xorg[%d] + a.dx * (C)%s;" % (d, d, canon_index[3 - dim + d])
              for d in range(dim) if "XYZ"[d] in used]
    vel_txt = ", ".join("{%d, %d, %d}" % v for v in vel)

    # ------------------------------------------- temporally blocked kernel (2D)
    if dim == 2:
        tb_loads = ["S %s;" % ", ".join("L%d[1]" % k for k in range(ns))]
        tb_loads += ["L%d[0] = cur[%d * RC + lcell - (%d * RI + %d)];" % (k, k, vel[k][1], vel[k][2])
                     for k in range(ns)]
        tb_sub = []
        if ksym is not None:
            # bounce-back copies as selects of unconditional LDS reads (one branch per
            # cell instead of one per direction), then the table links: bits and
            # first entry from LDS, the entries from memory
            tb_sub = ["if (a.bbmask) {",
                      "  const uint32_t bits = msk[lcell];",
                      "  if (bits) {"]
            for k in range(ns):
                if ksym[k] != k:
                    tb_sub += ["    {",
                               "      const S b_ = cur[%d * RC + lcell];" % ksym[k],
                               "      L%d[0] = (bits & %du) ? b_ : L%d[0];" % (k, 1 << k, k),
                               "    }"]
            # Per entry: the one-source tables (no bbtab_rel2) read source, sign and
            # value; tables with interpolated entries read the second relative
            # source and coefficient with them, all five loads issued together (bit
            # 31 of bbtab_rel2 flags the entries that have a second source; the
            # others repeat their first), so no load waits on another.
            tb_sub += ["    if (bits & 0x80000000u) {",
                       "      using CC = typename Codec<S>::C;",
                       "      const uint32_t tb = tbits[lcell], tbase = tbase_l[lcell];"]
            for two in (True, False):
                tb_sub += ["      if (%sa.bbtab_rel2) {" % ("" if two else "!")]
                for k in range(ns):
                    if ksym[k] == k:
                        continue
                    bit = 1 << k
                    tb_sub += ["        if (tb & %du) {" % bit,
                               "          const uint32_t e_ = tbase + (uint32_t)__builtin_popcount(tb & %du);" % (bit - 1),
                               "          const uint32_t rel = a.bbtab_rel[e_];"]
                    if two:
                        tb_sub += ["          const uint32_t rel2 = a.bbtab_rel2[e_];",
                                   "          const CC c2_ = a.bbtab_c2[e_];"]
                    tb_sub += ["          const CC sg_ = a.bbtab_sign[e_];",
                               "          const CC rh_ = a.bbtab_rhs[e_];",
                               "          const int ks = (int)(rel & 0xffu), dm = (int)((rel >> 8) & 0xffu) - 128,",
                               "                    di = (int)((rel >> 16) & 0xffu) - 128;",
                               "          CC r_ = sg_ * Codec<S>::decode_dev(cur[ks * RC + lcell + dm * RI + di], (CC)0);"]
                    if two:
                        tb_sub += ["          if (rel2 & 0x80000000u) {",
                                   "            const int ks2 = (int)(rel2 & 0xffu), dm2 = (int)((rel2 >> 8) & 0xffu) - 128,",
                                   "                      di2 = (int)((rel2 >> 16) & 0xffu) - 128;",
                                   "            r_ = r_ + c2_ * Codec<S>::decode_dev(cur[ks2 * RC + lcell + dm2 * RI + di2], (CC)0);",
                                   "          }"]
                    tb_sub += ["          r_ += rh_;",
                               "          L%d[0] = Codec<S>::encode(r_, (CC)0);" % k,
                               "        }"]
                tb_sub += ["      }"]
            tb_sub += ["    }", "  }", "}"]
        tb_kernel = _TB_KERNEL % dict(
            tb_loads=_indent(tb_loads, 6), tb_sub=_indent(tb_sub, 6),
            tb_gdecl=_indent(["S %s;" % ", ".join("G%d[1]" % k for k in range(ns))], 6),
            tb_coords=_indent(coords, 6), tb_decode=_indent(decode, 6), tb_f2m=_indent(f2m_block, 6),
            tb_coll_dev=_indent(coll_dev, 8), tb_m2f_dev=_indent(m2f_dev_block, 8),
            tb_stores_dev=_indent(stores_dev, 8), tb_coll=_indent(coll, 8), tb_m2f=_indent(m2f_block, 8),
            tb_stores=_indent(stores, 8),
            tb_lds_stores=_indent(["nxt[%d * RC + lcell] = G%d[0];" % (k, k) for k in range(ns)], 6))
        tb_launch = _TB_LAUNCH % {}
    else:
        tb_kernel = ""
        tb_launch = _TB_LAUNCH_NONE
    return _TEMPLATE % dict(
        tb_kernel=tb_kernel, tb_launch=tb_launch, unit_vel=int(all(abs(c) <= 1 for v in vel for c in v)),
        hash=spec_hash(spec), ns=ns, nc=spec.nconsm, dim=dim, paired=bool(sig_f2m),
        vel=vel_txt,
        loads=_indent(loads), bbsub=_indent(bbsub), bbsub_ok=int(ksym is not None),
        gdecl=_indent(gdecl), coords=_indent(coords, 4), decode=_indent(decode, 4), vstores=_indent(vstores),
        f2m=_indent(f2m_block, 4), coll=_indent(coll, 4), coll_dev=_indent(coll_dev, 4),
        m2f=_indent(m2f_block, 4), m2f_dev=_indent(m2f_dev_block, 4), stores_dev=_indent(stores_dev, 4),
        stores=_indent(stores, 4),
        f2m_loads=_indent(f2m_loads), f2m_full=_indent(f2m_full), f2m_stores=_indent(f2m_stores),
        m2f_loads=_indent(m2f_loads), m2f_full=_indent(m2f_full), m2f_stores=_indent(m2f_stores),
        eq_loads=_indent(eq_loads), eq_body=_indent(eq_body), eq_stores=_indent(eq_stores),
        rx_loads=_indent(rx_loads), rx_body=_indent(rx_body), rx_stores=_indent(rx_stores),
        st_coords=_indent(coords), st_loads=_indent(st_loads), st_body=_indent(st_body),
        st_stores=_indent(st_stores),
        eqm2f_loads=_indent(eqm2f_loads), eqm2f_opaque=_indent(eqm2f_opaque),
    )
