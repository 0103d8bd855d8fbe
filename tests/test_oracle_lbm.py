"""Known answers that pin the numpy oracle (oracle/lbm.py) and its C
restatement (oracle/cgen.py), independently of any GPU code."""
import math

import numpy as np
import pytest

from oracle.boundary import apply_links, LinkMethod, BB
from oracle.cgen import CStepper
from oracle.lbm import Simulation as O
from pylbm_amd import cases


def _tgv(n=8, **kw):
    return cases.taylor_green_d3q19(n, **kw)


def test_rest_state_fixed_point():
    d = _tgv(6)
    d["init"] = {cases.RHO: 1.0, cases.QX: 0.0, cases.QY: 0.0, cases.QZ: 0.0}
    o = O(d)
    f0 = o.F.copy()
    for _ in range(3):
        o.one_time_step()
    assert np.max(np.abs(o.F - f0)) < 1e-15


def test_zero_rates_is_pure_streaming():
    """With every rate 0 a step is f_k(x, t+1) = f_k(x - v_k, t) on a periodic box."""
    d = _tgv(6)
    d["schemes"][0]["relaxation_parameters"] = [0] * 19
    o = O(d)
    inner = o._interior()
    before = o.F[inner].copy()
    o.one_time_step()
    for k, v in enumerate(o.v):
        want = np.roll(before[k], shift=tuple(v), axis=(0, 1, 2))
        assert np.max(np.abs(o.F[(k,) + inner[1:]] - want)) < 1e-15


def test_unit_rates_project_on_equilibrium():
    d = _tgv(6)
    d["schemes"][0]["relaxation_parameters"] = [0] + [1] * 18
    o = O(d)
    o.ghost_fill(o.F)
    o.pull_step(o.F, o.Fnew)
    inner = o._interior()
    m = o._f2m(o.Fnew[inner])
    eq = o._eq_full(m)
    assert np.max(np.abs(o._m2f(eq) - o.Fnew[inner])) < 1e-15


def test_periodic_conservation():
    o = O(_tgv(8))
    s0 = o.conserved_sums()
    for _ in range(10):
        o.one_time_step()
    s1 = o.conserved_sums()
    assert abs(s1[0] - s0[0]) < 1e-12 * abs(s0[0])
    for a, b in zip(s0[1:], s1[1:]):
        assert abs(a - b) < 1e-13


def test_bounce_back_resting_walls_conserve_mass():
    d = cases.lid_cavity_d2q9(12)
    d["boundary_conditions"][1].pop("value")
    d["init"] = {cases.RHO: 1.0, cases.QX: lambda x, y: 0.01 * np.sin(x) * np.cos(y), cases.QY: 0.0}
    o = O(d)
    s0 = o.conserved_sums()[0]
    for _ in range(20):
        o.one_time_step()
    assert abs(o.conserved_sums()[0] - s0) < 1e-12


def test_apply_links_sequential_semantics():
    """A conflicting list is applied in loop order (later links see earlier stores)."""
    f = np.arange(12, dtype=float).reshape(1, 12)
    st = np.array([[0, 0, 0], [3, 5, 7]])
    ld = np.array([[0, 0, 0], [1, 3, 5]])
    m = LinkMethod(BB, st, np.zeros(3), np.zeros(3))
    m.iload = ld
    m.rhs = np.array([10.0, 20.0, 30.0])
    apply_links(f, m)
    assert f[0, 3] == 1 + 10 and f[0, 5] == 11 + 20 and f[0, 7] == 31 + 30


@pytest.mark.parametrize("make", [lambda: _tgv(10), lambda: cases.sphere_packing_d3q19(12, 10, 8, seed=1),
                                  lambda: cases.lid_cavity_d2q9(16), lambda: cases.karman_d2q9(40, 20, 4.0, (10., 10.))])
def test_c_restatement_matches_numpy(make):
    from oracle.boundary import apply_links as ap
    o = O(make())
    for _ in range(2):
        o.one_time_step()
    cs = CStepper(o.scheme)
    dim = o.dim
    h = tuple([0] * (3 - dim) + list(o.h))
    o.ghost_fill(o.F)
    for meth in o.methods:
        ap(o.F, meth)
    shp = (o.ns,) + (1,) * (3 - dim) + o.F.shape[1:]
    Fa = np.ascontiguousarray(np.moveaxis(o.F.reshape(shp), 0, -1))
    G = np.zeros_like(Fa)
    cs.step(Fa, G, h, o.t, o.dt)
    o.pull_step(o.F, o.Fnew)
    ref = np.moveaxis(o.Fnew.reshape(shp), 0, -1)
    sl = tuple(slice(x, -x) if x else slice(None) for x in h)
    assert np.max(np.abs(G[sl] - ref[sl])) < 1e-15


def test_c_ghost_fill_matches_numpy():
    o = O(_tgv(6))
    cs = CStepper(o.scheme)
    rng = np.random.default_rng(0)
    o.F[...] = rng.random(o.F.shape)
    A = np.ascontiguousarray(np.moveaxis(o.F, 0, -1))
    cs.ghost(A, (1, 1, 1))
    o.ghost_fill(o.F)
    assert np.array_equal(A, np.moveaxis(o.F, 0, -1))


def test_fsum_oracle():
    o = O(_tgv(6))
    m = o.moments()
    assert o.conserved_sums()[0] == math.fsum(m[0].ravel())


def _timed_lid(n, u_of_t):
    """Lid cavity whose lid value is time-dependent ("time_bc"): m[QX] = u_of_t(t)."""
    d = cases.lid_cavity_d2q9(n)

    def lid(f, m, t, x, y):
        m[cases.RHO] = 1.0
        m[cases.QX] = u_of_t(t)
        m[cases.QY] = 0.0

    d["boundary_conditions"][1] = {"method": {0: d["boundary_conditions"][1]["method"][0]}, "value": lid,
                                   "time_bc": True}
    return d


def test_time_dependent_value_ignoring_t_equals_static():
    """A time-dependent lid whose function ignores t runs exactly like the static lid."""
    a = O(cases.lid_cavity_d2q9(12))
    b = O(_timed_lid(12, lambda t: 0.05))
    for _ in range(15):
        a.one_time_step()
        b.one_time_step()
    assert np.array_equal(a.F, b.F)


def test_time_dependent_value_switch():
    """A lid that switches from 0.05 to 0.02 at t0 = 6 follows the static 0.05 run
    up to t0 and then the static 0.02 run started from its state at t0: the value
    is taken at the step's start time (pylbm/simulation.py:384-390)."""
    t0 = 6
    timed = O(_timed_lid(12, lambda t: 0.05 if t < t0 else 0.02))
    fast = O(cases.lid_cavity_d2q9(12, u_lid=0.05))
    for _ in range(t0):
        timed.one_time_step()
        fast.one_time_step()
    assert np.array_equal(timed.F, fast.F)
    # same viscosity (nu depends on u_lid): the 0.02 lid on the 0.05 case's scheme
    d = cases.lid_cavity_d2q9(12, u_lid=0.05)
    d["boundary_conditions"][1]["value"] = _timed_lid(12, lambda t: 0.02)["boundary_conditions"][1]["value"]
    d["boundary_conditions"][1]["time_bc"] = True
    slow = O(d)
    slow.F[...] = timed.F
    slow.t, slow.nt = timed.t, timed.nt
    for _ in range(5):
        timed.one_time_step()
        slow.one_time_step()
    assert np.array_equal(timed.F, slow.F)
    assert not np.array_equal(timed.F, fast.F)
