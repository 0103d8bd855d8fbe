"""Moment matrices, equilibrium and conserved-moment permutation.

Pins (pylbm/scheme.py:464-607): M_ij = P_i(la v_j) evaluated independently
with numpy polynomials; M invM = I exactly; rest weights of D2Q9 / D3Q19;
sequential row swaps when the conserved moments are not at the front."""
from fractions import Fraction

import numpy as np
import pytest
import sympy as sp

from oracle.scheme import Scheme as OScheme
from pylbm_amd import cases
from pylbm_amd.scheme import Scheme

X, Y, Z = sp.symbols("X, Y, Z")
DICTS = {
    "d2q9": lambda: cases.lid_cavity_d2q9(8),
    "d3q19": lambda: cases.taylor_green_d3q19(4),
    "d3q19src": lambda: cases.sphere_packing_d3q19(8, 8, 8),
    "d1q3": lambda: cases.d1q3_advection(8),
    "vectorial": lambda: cases.d2q5q9_vectorial(8),
}


@pytest.mark.parametrize("name", sorted(DICTS))
def test_matrices(name):
    d = DICTS[name]()
    s = Scheme(d)
    o = OScheme(d)
    M = np.array(s.M.evalf().tolist(), dtype=float)
    assert np.array_equal(M, o.Mnum)
    assert s.M * s.invM == sp.eye(s.M.shape[0])
    spec = s.kernel_spec()
    prod = [[sum(spec.M[i][k] * spec.invM[k][j] for k in range(spec.ns)) for j in range(spec.ns)]
            for i in range(spec.ns)]
    assert all(prod[i][j] == (1 if i == j else 0) for i in range(spec.ns) for j in range(spec.ns))
    assert [str(e) for e in spec.eq] == [str(e).replace("m", "m") for e in o.eq]


@pytest.mark.parametrize("name", ["d2q9", "d3q19"])
def test_closed_form_rows(name):
    """Unpermuted rows are the polynomials evaluated at la*v (numpy evaluation)."""
    d = DICTS[name]()
    sch = d["schemes"][0]
    s = Scheme(d)
    vel = s.stencil.get_all_velocities().astype(float)
    la = 1.0
    coords = [X, Y, Z][: vel.shape[1]]
    lam = sp.Symbol("lambda", constants=True)
    for i, p in enumerate(sch["polynomials"]):
        f = sp.lambdify(coords, sp.sympify(p).subs(lam, 1), "numpy")
        row = np.array([float(f(*(la * v))) for v in vel])
        # conserved moments were moved to the front by sequential swaps
        eq = [sp.sympify(e) for e in sch["equilibrium"]]
        consm = sch["conserved_moments"]
        perm = list(range(len(eq)))
        for front, c in enumerate(consm):
            pos = eq.index(c)
            perm[front], perm[pos] = perm[pos], perm[front]
        target = perm.index(i)
        assert np.allclose(np.array(s.M.evalf().tolist(), float)[target], row, rtol=0, atol=1e-12)


@pytest.mark.parametrize("name,weights", [("d2q9", [4 / 9] + [1 / 9] * 4 + [1 / 36] * 4),
                                          ("d3q19", [1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)])
def test_rest_weights(name, weights):
    o = OScheme(DICTS[name]())
    mc = [np.array(1.0)] + [np.array(0.0)] * (o.nconsm - 1)
    meq = np.array([float(e) for e in o.equilibrium(mc)])
    assert np.allclose(o.invMnum @ meq, weights, rtol=0, atol=1e-15)


def test_sequential_swaps_out_of_order():
    """Conserved moments listed out of order: the permutation is the sequence of swaps."""
    d = cases.lid_cavity_d2q9(8)
    sch = d["schemes"][0]
    sch["conserved_moments"] = [cases.QY, cases.RHO, cases.QX]
    s, o = Scheme(d), OScheme(d)
    assert list(s.consm.values()) == [0, 1, 2]
    assert np.array_equal(np.array(s.M.evalf().tolist(), float), o.Mnum)
    # swaps (0,2), (1,0), (2,1) applied in sequence to [rho, qx, qy]:
    # [qy, qx, rho] -> [qx, qy, rho] -> [qx, rho, qy]  (scheme.py:225-241 does exactly this)
    assert [str(e) for e in s.EQ[:3]] == ["qx", "rho", "qy"]


def test_orthogonal_d3q19():
    s = Scheme(cases.taylor_green_d3q19(4))
    M = np.array(s.M.evalf().tolist(), float)
    G = M @ M.T
    assert np.allclose(G, np.diag(np.diag(G)))
    assert int(np.count_nonzero(M)) == 213


def test_irrational_scheme_velocity_kept():
    """A scheme velocity that sympy recognises as irrational (pylbm's Kelvin-Helmholtz
    demo: la = sqrt(3) / 0.1 = 10 sqrt(3)) keeps its value in the device matrices: M
    and M^-1 of the kernel spec equal the oracle's floating-point matrices (a
    truncation to the integer 17 would leave them 1.8 % off), and their product is the
    identity to rounding."""
    from pylbm_amd import demos
    d, _ = demos.kelvin_helmholtz(1.0 / 64)
    spec = Scheme(d).kernel_spec()
    o = OScheme(d)
    M = np.array([[float(c) for c in r] for r in spec.M])
    iM = np.array([[float(c) for c in r] for r in spec.invM])
    assert np.allclose(M, np.asarray(o.Mnum, float), rtol=1e-15, atol=0)
    assert np.allclose(iM, np.asarray(o.invMnum, float), rtol=1e-13, atol=1e-18)
    assert M[1, 1] == pytest.approx(np.sqrt(3) / 0.1, rel=1e-15)
    assert np.abs(M @ iM - np.eye(spec.ns)).max() < 1e-13
