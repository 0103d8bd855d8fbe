"""Relative-velocity schemes (pylbm/scheme.py:464-543, algorithm/base.py:102-114,
255-263): the shift matrices of the oracle and of the product.

* T(u) T(-u) = I exactly and T(0) = I (the reference's own check,
  scheme.py:534-543);
* T(u) M equals Mu[i, j] = P_i(la v_j - u) evaluated with numpy from the
  polynomials at random u, rows in the order of the conserved-moment swaps;
* the product's kernel spec carries the oracle's T(u), T(-u) and relative
  velocity exactly;
* end to end the oracle reproduces pylbm's own fixtures of the two
  relative-velocity demos (tests/test_reference_fixtures.py).
"""
import numpy as np
import pytest
import sympy as sp

from oracle.scheme import REL_U, Scheme as OScheme
from pylbm_amd import demos
from pylbm_amd.scheme import RU, Scheme

DEMOS = {"lid": demos.lid_driven_cavity, "karman": demos.karman_vortex_street}


@pytest.mark.parametrize("name", sorted(DEMOS))
def test_shift_inverse_and_identity(name):
    d, _ = DEMOS[name](1.0 / 16)
    sch = OScheme(d)
    prod = (sch.Tu * sch.Tmu).applyfunc(sp.expand)
    assert prod == sp.eye(sch.ns)
    zero = {u: 0 for u in REL_U}
    assert sch.Tu.subs(zero) == sp.eye(sch.ns)
    assert sch.Tmu.subs(zero) == sp.eye(sch.ns)


def _mu_numpy(d, u):
    """P_i(la v_j - u) for the dictionary's polynomials, unswapped, by numpy."""
    X, Y = sp.symbols("X, Y")
    sch = d["schemes"][0]
    la = 1.0
    vel = OScheme(d).stencil.all_velocities()
    out = np.empty((len(vel), len(vel)))
    for i, p in enumerate(sch["polynomials"]):
        f = sp.lambdify((X, Y), sp.sympify(p).subs(d.get("parameters", {})), "numpy")
        for j, v in enumerate(vel):
            out[i, j] = f(la * v[0] - u[0], la * v[1] - u[1])
    return out


@pytest.mark.parametrize("name", sorted(DEMOS))
def test_shift_times_m_is_shifted_polynomials(name):
    d, _ = DEMOS[name](1.0 / 16)
    sch = OScheme(d)
    rng = np.random.default_rng(3)
    perm = list(range(sch.ns))
    # the sequential swaps of the conserved moments (scheme.py:225-241), on row labels
    eq = [sp.sympify(e) for e in d["schemes"][0]["equilibrium"]]
    for front, c in enumerate(d["schemes"][0]["conserved_moments"]):
        pos = eq.index(c)
        eq[front], eq[pos] = eq[pos], eq[front]
        perm[front], perm[pos] = perm[pos], perm[front]
    for _ in range(3):
        u = rng.uniform(-0.2, 0.2, size=2)
        tu = sch.shift([np.asarray(u[0]), np.asarray(u[1])])
        got = tu @ sch.Mnum
        want = _mu_numpy(d, u)[perm]
        assert np.allclose(got, want, rtol=1e-13, atol=1e-13)


@pytest.mark.parametrize("name", sorted(DEMOS))
def test_product_spec_equals_oracle(name):
    d, _ = DEMOS[name](1.0 / 16)
    o = OScheme(d)
    spec = Scheme(d).kernel_spec()
    ren = dict(zip(RU, REL_U))
    ns = spec.ns
    for i in range(ns):
        for j in range(ns):
            assert sp.simplify(sp.sympify(spec.Tu[i][j]).xreplace(ren) - o.Tu[i, j]) == 0
            assert sp.simplify(sp.sympify(spec.Tmu[i][j]).xreplace(ren) - o.Tmu[i, j]) == 0
    for a, b in zip(spec.rel_vel, o.rel_vel_exprs):
        assert sp.simplify(a - b) == 0


def test_feq_equilibrium_equals_moment_equilibrium():
    """An 'feq' entry (scheme.py:202-223: M_block . feq(v, *args)) gives the moments of
    the distribution it returns: a D2Q9 BGK feq against the moments computed by hand."""
    from pylbm_amd import cases
    d = cases.lid_cavity_d2q9(8)
    sch = d["schemes"][0]
    rho, qx, qy = cases.RHO, cases.QX, cases.QY
    w = {0: sp.Rational(4, 9), 1: sp.Rational(1, 9), 2: sp.Rational(1, 36)}

    def feq(v, c2):
        out = []
        for vx, vy in v:
            cu = vx * qx + vy * qy
            out.append(w[abs(int(vx)) + abs(int(vy))] * (rho + 3 * cu + sp.Rational(9, 2) * cu ** 2 / rho
                                                         - sp.Rational(3, 2) * (qx ** 2 + qy ** 2) / rho) * c2)
        return out

    d2 = dict(d)
    s2 = dict(sch)
    s2.pop("equilibrium")
    s2["feq"] = (feq, (1,))
    d2["schemes"] = [s2]
    a = Scheme(d2)
    o = OScheme(d2)
    vel = a.stencil.get_all_velocities(0)
    fe = sp.Matrix(feq(vel, 1))
    for i in range(a.M.shape[0]):
        want = sp.simplify((a.M.row(i) * fe)[0])
        assert sp.simplify(a.EQ[i] - want) == 0
        assert sp.simplify(o.eq[i] - sp.sympify(want).subs({rho: o.msyms[0], qx: o.msyms[1], qy: o.msyms[2]})) == 0
