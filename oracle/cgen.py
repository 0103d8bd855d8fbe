"""C restatement of pylbm's generated Cython ``one_time_step`` -- oracle / CPU baseline.

pylbm's Cython backend prints, per scheme, a loop nest over the interior
(pylbm/generator/printing/cython.py:301-314) whose body is the sequence of
pylbm/algorithm/pull.py:11-59 with a stack array ``double m[ns]``
(pylbm/generator/codegen.py:1118-1126): dense per-row assignments
m[i] = sum_j M_ij f[x - v_j, j], source half step, relaxation rows
m[i] = (1 - s_i) m[i] + s_i eq_i (identity rows dropped), source half step,
fnew[x, k] = sum_i invM_ki m[i], with constants printed to 16 significant
digits (cython.py:209-210), on the AoS layout [x][y][z][k]
(pylbm/container.py:75-90), compiled -O3 without fast-math
(pylbm/generator/autowrap.py:103).

This module prints the same C for an oracle ``Scheme`` and compiles it with
gcc (-O3 -ffp-contract=off, optional OpenMP over the outer axis, standing in
for pylbm's MPI slab decomposition).  Used only as the timed CPU baseline and
checked against the numpy oracle in tests/test_oracle_cgen.py.
"""
import ctypes
import hashlib
import os
import subprocess

import numpy as np
import sympy as sp
from sympy.printing.c import C99CodePrinter

HERE = os.path.dirname(os.path.abspath(__file__))
GEN = os.path.join(HERE, "_gen")


class _P(C99CodePrinter):
    def _print_Rational(self, e):
        return "%.16g" % float(e)

    def _print_Float(self, e):
        return "%.16g" % float(e)

    def _print_Integer(self, e):
        return "%d.0" % int(e)

    def _print_Pow(self, e):
        b, x = e.as_base_exp()
        if x.is_Integer and int(x) > 0:
            return "(" + "*".join([self.parenthesize(b, 100)] * int(x)) + ")"
        return super()._print_Pow(e)

    def _print_Symbol(self, e):
        if e.name.startswith("m") and e.name[1:].isdigit():
            return "m[%s]" % e.name[1:]
        return e.name


def _c(e):
    return _P().doprint(sp.sympify(e))


def source(scheme):
    ns, dim = scheme.ns, scheme.dim
    v = scheme.stencil.all_velocities()
    vv = np.zeros((ns, 3), dtype=int)
    vv[:, 3 - dim:] = v
    lines = []
    for i in range(ns):
        terms = ["%.16g*f[%d]" % (scheme.Mnum[i, j], j) for j in range(ns) if scheme.Mnum[i, j] != 0]
        lines.append("m[%d] = %s;" % (i, " + ".join(terms) if terms else "0.0"))
    src_lines = ["m[%d] = m[%d] + 0.5*dt*(%s);" % (i, i, _c(e)) for i, e in scheme.source_exprs]
    lines += src_lines
    for i in range(ns):
        s = scheme.s[i]
        if s != 0:
            lines.append("m[%d] = %.16g*m[%d] + %.16g*(%s);" % (i, 1.0 - s, i, s, _c(scheme.eq[i])))
    lines += src_lines
    body = "\n        ".join(lines)
    stores = "\n        ".join(
        "fnew[c + %d] = %s;" % (k, " + ".join("%.16g*m[%d]" % (scheme.invMnum[k, i], i)
                                              for i in range(ns) if scheme.invMnum[k, i] != 0) or "0.0")
        for k in range(ns))
    loads = "\n        ".join(
        "f[%d] = fin[c - (((long)%d*n1 + %d)*n2 + %d)*NS + %d];" % (k, vv[k, 0], vv[k, 1], vv[k, 2], k)
        for k in range(ns))
    return r"""
#include <math.h>
#define NS %(ns)d
/* AoS [o][m][i][k], halo h0,h1,h2 on the canonical (outer, mid, inner) axes */
void one_time_step(const double* fin, double* fnew, long n0, long n1, long n2,
                   int h0, int h1, int h2, double t, double dt) {
  #pragma omp parallel for schedule(static)
  for (long o = h0; o < n0 - h0; ++o)
    for (long mm = h1; mm < n1 - h1; ++mm)
      for (long i = h2; i < n2 - h2; ++i) {
        const long c = ((o*n1 + mm)*n2 + i)*NS;
        double f[NS], m[NS];
        %(loads)s
        %(body)s
        %(stores)s
      }
}

/* periodic ghost fill of every population, axis by axis (pylbm/storage.py:306-367) */
void ghost_fill(double* f, long n0, long n1, long n2, int h0, int h1, int h2) {
  long n[3] = {n0, n1, n2};
  int h[3] = {h0, h1, h2};
  for (int a = 0; a < 3; ++a) {
    if (h[a] == 0) continue;
    long st[3] = {n1*n2*NS, n2*NS, NS};
    long outer = 1, inner = 1;
    for (int b = 0; b < a; ++b) outer *= n[b];
    for (int b = a + 1; b < 3; ++b) inner *= n[b];
    #pragma omp parallel for schedule(static)
    for (long p = 0; p < outer; ++p)
      for (int g = 0; g < h[a]; ++g)
        for (long q = 0; q < inner*NS; ++q) {
          const long base = p*n[a]*inner*NS + q;
          f[base + g*st[a]] = f[base + (n[a] - 2*h[a] + g)*st[a]];
          f[base + (n[a] - h[a] + g)*st[a]] = f[base + (h[a] + g)*st[a]];
        }
  }
}
""" % dict(ns=ns, loads=loads, body=body, stores=stores)


class CStepper:
    """Compiled C step of one oracle scheme; arrays are AoS float64 (n0, n1, n2, ns)."""

    def __init__(self, scheme, openmp=True):
        src = source(scheme)
        key = hashlib.sha1((src + str(openmp)).encode()).hexdigest()[:16]
        os.makedirs(GEN, exist_ok=True)
        so = os.path.join(GEN, "cstep_%s.so" % key)
        if not os.path.exists(so):
            cfile = so[:-3] + ".c"
            with open(cfile, "w") as fh:
                fh.write(src)
            flags = ["-O3", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math"]
            if openmp:
                flags.append("-fopenmp")
            tmp = so + ".tmp%d" % os.getpid()
            subprocess.run(["gcc"] + flags + [cfile, "-o", tmp, "-lm"], check=True)
            os.replace(tmp, so)
        self.lib = ctypes.CDLL(so)
        D, L, I = ctypes.c_double, ctypes.c_long, ctypes.c_int
        P = ctypes.c_void_p
        self.lib.one_time_step.argtypes = [P, P, L, L, L, I, I, I, D, D]
        self.lib.ghost_fill.argtypes = [P, L, L, L, I, I, I]
        self.scheme = scheme

    def step(self, f, fnew, halo, t, dt):
        n0, n1, n2 = f.shape[:3]
        self.lib.one_time_step(f.ctypes.data, fnew.ctypes.data, n0, n1, n2, *halo, t, dt)

    def ghost(self, f, halo):
        n0, n1, n2 = f.shape[:3]
        self.lib.ghost_fill(f.ctypes.data, n0, n1, n2, *halo)


def set_threads(n):
    """Number of OpenMP threads for subsequent calls."""
    try:
        lib = ctypes.CDLL("libgomp.so.1")
        lib.omp_set_num_threads(int(n))
    except OSError:
        pass
