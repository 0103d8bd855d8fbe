"""Closed-form checks of the oracle's source half-steps, valued boundary
right-hand sides and Adam, each chosen so that a plausible mistake (one
half-step instead of two, dt instead of dt/2, a dropped feq term, a wrong
sign, no bias correction, a misplaced eps) fails it.

* Source terms (pylbm/algorithm/ode.py:11-16, pylbm/algorithm/pull.py:46-55):
  - a body force S(qx) = rho g on a periodic box raises the total x-momentum by
    exactly dt g sum(rho) per step (two half-steps of dt/2; relaxation keeps
    the conserved qx);
  - a uniform field under the logistic source mu u (1 - u) follows two
    explicit Euler half-steps per step, u <- e(e(u)), e(u) = u + dt/2 mu u (1 - u)
    (transport does not act on a uniform field).
* Valued bounce-back (pylbm/boundary.py:421-427) and anti-bounce-back
  (boundary.py:637-643) at a wall with prescribed (rho, j):
  - feq_k - feq_ksym = 2 w_k (c_k . j) / c_s^2, c_s^2 = la^2 / 3, for the D2Q9
    and D3Q19 MRT equilibria (their odd moments are the lattice-BGK ones);
  - feq_k + feq_ksym = 2 w_k (rho + (c_k . j)^2 / (2 c_s^4) - |j|^2 / (2 c_s^2)),
    the textbook second-order equilibrium with rho0 = 1;
  - plane Couette flow (resting floor, lid moving at U, periodic in x) reaches
    the exact linear momentum profile q_x(y) = U y / H with half-way walls.
* Adam: the oracle's update equals torch.optim.Adam (CPU) over several steps,
  and its first step is -lr sign(g) for |g| >> eps.
"""
import math

import numpy as np
import pytest
import sympy as sp

from oracle.boundary import ABB, BB
from oracle.lbm import Simulation
from pylbm_amd import bc, cases

W_D2Q9 = {0: 4 / 9, 1: 1 / 9, 2: 1 / 36}
W_D3Q19 = {0: 1 / 3, 1: 1 / 18, 2: 1 / 36}


def _weights(v):
    table = W_D2Q9 if v.shape[1] == 2 else W_D3Q19
    return np.array([table[int(np.sum(np.abs(c)))] for c in v])


# ------------------------------------------------------------------ sources

def _periodic_d3q19_force(g, dx):
    rng = np.random.default_rng(7)
    n = (6, 5, 4)
    drho = rng.uniform(-0.05, 0.05, size=n)

    def rho(x, y, z):
        i = np.rint(x / dx - 0.5).astype(int) % n[0]
        j = np.rint(y / dx - 0.5).astype(int) % n[1]
        k = np.rint(z / dx - 0.5).astype(int) % n[2]
        return 1.0 + drho[i, j, k]

    entry = cases.d3q19_mrt(1.3, source_qx=cases.RHO * g)
    return {
        "box": {"x": [0.0, n[0] * dx], "y": [0.0, n[1] * dx], "z": [0.0, n[2] * dx], "label": -1},
        "space_step": dx,
        "scheme_velocity": 1.0,
        "schemes": [entry],
        "init": {cases.RHO: rho, cases.QX: 0.0, cases.QY: 0.0, cases.QZ: 0.0},
        "generator": "numpy",
    }


@pytest.mark.parametrize("dx", [1.0, 0.5])
def test_body_force_momentum_growth(dx):
    g = 1e-3
    sim = Simulation(_periodic_d3q19_force(g, dx))
    mass0 = sim.conserved_sums()[0]
    assert sim.conserved_sums()[1] == 0.0
    n = 7
    for _ in range(n):
        sim.one_time_step()
    sums = sim.conserved_sums()
    assert abs(sums[0] - mass0) <= 1e-13 * mass0
    expect = n * sim.dt * g * mass0
    assert abs(sums[1] - expect) <= 1e-12 * expect, (sums[1], expect)
    assert abs(sums[2]) <= 1e-15 * mass0 and abs(sums[3]) <= 1e-15 * mass0


def test_logistic_source_two_euler_half_steps():
    U = sp.Symbol("u")
    C, MU, LA = sp.symbols("c, mu, lambda", constants=True)
    dx, mu, u0 = 1.0 / 64, 3.0, 0.2
    d = {
        "box": {"x": [0.0, 1.0], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{"velocities": [1, 2], "conserved_moments": U, "polynomials": [1, sp.Symbol("X")],
                     "relaxation_parameters": [0.0, 1.7], "equilibrium": [U, C * U],
                     "source_terms": {U: MU * U * (1 - U)}}],
        "init": {U: u0},
        "parameters": {LA: 1.0, C: 0.25, MU: mu},
        "generator": "numpy",
    }
    sim = Simulation(d)
    u = u0
    half = 0.5 * sim.dt * mu
    for _ in range(20):
        sim.one_time_step()
        u = u + half * u * (1 - u)
        u = u + half * u * (1 - u)
    got = sim.m(U)
    assert np.max(np.abs(got - u)) <= 1e-14, (got[0], u)
    # one full Euler step of dt per step would be off by O(dt^2 mu^2)
    assert abs(u - u0) > 1e-2


def test_space_time_source_momentum_growth():
    """A force depending on the cell centres and on t, S(qx) = rho g0 (1 + Y)(1 + t),
    S(qy) = rho g0 X / 4, on a periodic box: streaming permutes the populations and
    relaxation keeps the conserved moments, so over the step from t_n the total
    momentum grows by exactly dt sum_x rho^{n+1}(x) S(x, t_n) / rho -- the two half
    steps at the same time t_n and the density after streaming (the new state's).
    Swapping X and Y, taking t_{n+1}, or one half step fails it."""
    g0 = 1e-3
    d = cases.forced_periodic_d2q9(g0=g0, generator="numpy")
    sim = Simulation(d)
    X, Y = np.meshgrid(*sim.domain.coords, indexing="ij")
    for _ in range(5):
        t_n = sim.t
        before = sim.conserved_sums()
        sim.one_time_step()
        after = sim.conserved_sums()
        rho = sim.m(cases.RHO)
        want_x = sim.dt * math.fsum((rho * g0 * (1 + Y) * (1 + t_n)).ravel())
        want_y = sim.dt * math.fsum((rho * g0 * X / 4).ravel())
        assert abs((after[1] - before[1]) - want_x) <= 1e-12 * abs(want_x), (after[1] - before[1], want_x)
        assert abs((after[2] - before[2]) - want_y) <= 1e-12 * abs(want_y), (after[2] - before[2], want_y)
        assert abs(after[0] - before[0]) <= 1e-13 * abs(before[0])
    assert sim.t > 0


# -------------------------------------------------------- valued boundaries

def _lid_d2q9(u_lid, n=8, s_nu=1.2):
    d = cases.lid_cavity_d2q9(n, u_lid=u_lid, generator="numpy")
    d["schemes"] = [cases.d2q9_mrt(1.5, 1.4, 1.3, s_nu)]
    return d


def _lid_d3q19(u_lid, n=6):
    def lid(f, m, x, y, z):
        m[cases.RHO] = 1.0
        m[cases.QX] = u_lid[0]
        m[cases.QY] = u_lid[1]
        m[cases.QZ] = u_lid[2]

    return {
        "box": {"x": [0.0, n], "y": [0.0, n], "z": [0.0, n], "label": [0, 0, 0, 0, 0, 1]},
        "space_step": 1.0,
        "scheme_velocity": 1.0,
        "schemes": [cases.d3q19_mrt(1.3)],
        "init": {cases.RHO: 1.0, cases.QX: 0.0, cases.QY: 0.0, cases.QZ: 0.0},
        "boundary_conditions": {0: {"method": {0: bc.BounceBack}},
                                1: {"method": {0: bc.BounceBack}, "value": lid}},
        "generator": "numpy",
    }


def _check_odd_rhs(sim, j_of_label):
    v = sim.v.astype(float)
    w = _weights(sim.v)
    la = sim.scheme.la
    ndone = 0
    for meth in sim.methods:
        assert meth.kind == BB
        k = meth.istore[0]
        for lab, j in j_of_label.items():
            sel = meth.ilabel == lab
            expect = 2 * w[k[sel]] * (v[k[sel]] @ np.asarray(j, float)) * la / (la ** 2 / 3)
            assert np.allclose(meth.rhs[sel], expect, rtol=0, atol=1e-16), lab
            ndone += int(sel.sum())
    assert ndone > 0


def test_bounce_back_value_rhs_d2q9():
    sim = Simulation(_lid_d2q9(0.05))
    _check_odd_rhs(sim, {0: (0.0, 0.0), 1: (0.05, 0.0)})


def test_bounce_back_value_rhs_d3q19():
    u = (0.03, -0.02, 0.0)
    sim = Simulation(_lid_d3q19(u))
    _check_odd_rhs(sim, {0: (0.0, 0.0, 0.0), 1: u})


@pytest.mark.parametrize("flavour", ["d2q9", "d3q19"])
def test_anti_bounce_back_value_rhs(flavour):
    j = (0.04, 0.01) if flavour == "d2q9" else (0.04, 0.01, -0.02)
    rho = 1.02
    if flavour == "d2q9":
        d = _lid_d2q9(0.0)
        entry = d["schemes"][0]
    else:
        d = _lid_d3q19((0.0, 0.0, 0.0))
        entry = d["schemes"][0]

    def value(f, m, *xs):
        m[cases.RHO] = rho
        for sym, c in zip((cases.QX, cases.QY, cases.QZ), j):
            m[sym] = c

    d["boundary_conditions"] = {0: {"method": {0: bc.BounceBack}},
                                1: {"method": {0: bc.AntiBounceBack}, "value": value}}
    d["schemes"] = [entry]
    sim = Simulation(d)
    v = sim.v.astype(float)
    w = _weights(sim.v)
    cs2 = 1.0 / 3.0
    checked = 0
    for meth in sim.methods:
        if meth.kind != ABB:
            continue
        k = meth.istore[0]
        cj = v[k] @ np.asarray(j)
        j2 = float(np.dot(j, j))
        expect = 2 * w[k] * (rho + cj ** 2 / (2 * cs2 ** 2) - j2 / (2 * cs2))
        assert np.allclose(meth.rhs, expect, rtol=0, atol=2e-16)
        checked += k.size
    assert checked > 0


def test_couette_linear_profile():
    """Fluid rows y in (2, 10) between two solid slabs of a fully periodic box;
    the slab faces cut the wall links at q = 0.4, plain bounce-back puts the
    walls half-way (y = 2 and y = 10), the upper slab moves at U.  (Walls on the
    box's own faces next to a periodic face would not do: pylbm flags the
    corner links that cross both faces as periodic, domain.py:492-519.)"""
    from pylbm_amd.elements import Parallelogram

    U = 0.02

    def lid(f, m, x, y):
        m[cases.RHO] = 1.0
        m[cases.QX] = U
        m[cases.QY] = 0.0

    d = {"parameters": {cases.LA: 1}, "box": {"x": [0.0, 4.0], "y": [0.0, 12.0], "label": -1},
         "space_step": 1.0, "scheme_velocity": cases.LA, "schemes": [cases.d2q9_mrt(1.5, 1.4, 1.3, 1.1)],
         "elements": [Parallelogram((-2.0, -1.0), (8.0, 0.0), (0.0, 3.1), label=0),
                      Parallelogram((-2.0, 9.9), (8.0, 0.0), (0.0, 3.0), label=1)],
         "init": {cases.RHO: 1.0, cases.QX: 0.0, cases.QY: 0.0},
         "boundary_conditions": {0: {"method": {0: bc.BounceBack}},
                                 1: {"method": {0: bc.BounceBack}, "value": lid}},
         "generator": "numpy"}
    sim = Simulation(d)
    for _ in range(2500):
        sim.one_time_step()
    y = sim.domain.coords[1]
    fluid = (y > 2) & (y < 10)
    qx = sim.m(cases.QX)[:, fluid]
    expect = U * (y[fluid] - 2.0) / 8.0
    assert np.max(np.abs(qx - expect[None, :])) <= 1e-12 * U, np.max(np.abs(qx - expect[None, :]))
    assert np.max(np.abs(sim.m(cases.QY))) <= 1e-14


# -------------------------------------------------------------------- Adam

def test_adam_matches_torch():
    torch = pytest.importorskip("torch")
    from oracle.surrogate import Adam

    rng = np.random.default_rng(3)
    shapes = [(3, 4), (5,)]
    params = [[rng.standard_normal(s) for s in shapes]]
    tparams = [torch.tensor(p.copy(), dtype=torch.float64, requires_grad=True) for p in params[0]]
    lr, eps = 2e-3, 1e-8
    opt = Adam(params, lr=lr, eps=eps)
    topt = torch.optim.Adam(tparams, lr=lr, betas=(0.9, 0.999), eps=eps)
    for _ in range(6):
        grads = [[rng.standard_normal(s) * 10 ** rng.uniform(-4, 1) for s in shapes]]
        opt.step(params, grads)
        for tp, g in zip(tparams, grads[0]):
            tp.grad = torch.tensor(g, dtype=torch.float64)
        topt.step()
    for p, tp in zip(params[0], tparams):
        assert np.allclose(p, tp.detach().numpy(), rtol=1e-12, atol=1e-15)


def test_adam_first_step_is_lr_sign():
    from oracle.surrogate import Adam

    g = np.array([3.0, -0.5, 2e-3, -40.0])
    p = [[np.zeros(4)]]
    Adam(p, lr=1e-2, eps=1e-12).step(p, [[g]])
    assert np.allclose(p[0][0], -1e-2 * np.sign(g), rtol=1e-9, atol=0)


# -------------------------------------------------------- split routines

def test_split_transport_is_periodic_shift():
    """transport on a state whose ghost slots hold the periodic wrap is, per population,
    the periodic shift of the interior by v_k (np.roll); a sign error in the pull
    (x + v_k) or a shifted interior window fails it."""
    d = cases.taylor_green_d3q19(6, generator="numpy")
    sim = Simulation(d)
    rng = np.random.default_rng(3)
    sim.F[...] = rng.standard_normal(sim.F.shape)
    sim.ghost_fill(sim.F)
    sim.Fnew[...] = 0.0
    sim.transport()
    inner = sim.F[sim._interior()]
    got = sim.Fnew[sim._interior()]
    for k in range(sim.ns):
        want = np.roll(inner[k], shift=tuple(int(c) for c in sim.v[k]), axis=tuple(range(sim.dim)))
        assert np.array_equal(got[k], want), k


def test_split_relaxation_contracts_towards_equilibrium():
    """relaxation: every relaxed row's distance to the equilibrium shrinks by exactly
    (1 - s_i) (the equilibrium depends only on the conserved moments, which are not
    relaxed), the conserved rows are untouched, and the equilibrium state is a fixed
    point; s_i and 1 - s_i swapped, or eq taken after the update, fail it."""
    d = cases.lid_cavity_d2q9(8, generator="numpy")
    sim = Simulation(d)
    rng = np.random.default_rng(4)
    m = sim._f2m(sim.F) + 1e-3 * rng.standard_normal(sim.F.shape)
    m0 = m.copy()
    eq0 = sim._eq_full(m0)
    sim.relaxation(m)
    nc = sim.scheme.nconsm
    assert np.array_equal(m[:nc], m0[:nc])
    for i in range(nc, sim.ns):
        s = sim.scheme.s[i]
        np.testing.assert_allclose(m[i] - eq0[i], (1 - s) * (m0[i] - eq0[i]), rtol=1e-12, atol=1e-15)
    fixed = eq0.copy()
    sim.relaxation(fixed)
    np.testing.assert_allclose(fixed, eq0, rtol=0, atol=1e-15)


def test_split_source_term_half_step():
    """source_term: one explicit Euler half step at the current t on the interior only:
    for S(qx) = rho g0 (1 + Y)(1 + t), S(qy) = rho g0 X / 4 the momentum rows grow by
    dt/2 times those values at the cell centres, rho and the ghost cells are unchanged;
    a full step, swapped coordinates or t + dt fail it."""
    g0 = 1e-3
    d = cases.forced_periodic_d2q9(g0=g0, generator="numpy")
    sim = Simulation(d)
    sim.one_time_step()
    sim.one_time_step()
    m = sim._f2m(sim.F)
    m0 = m.copy()
    sim.source_term(m)
    X, Y = np.meshgrid(*sim.domain.coords, indexing="ij")
    c = sim.scheme.consm
    irho, iqx, iqy = c[cases.RHO], c[cases.QX], c[cases.QY]
    inner = sim._interior()[1:]
    rho = m0[irho][inner]
    np.testing.assert_allclose(m[iqx][inner] - m0[iqx][inner], 0.5 * sim.dt * rho * g0 * (1 + Y) * (1 + sim.t),
                               rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(m[iqy][inner] - m0[iqy][inner], 0.5 * sim.dt * rho * g0 * X / 4,
                               rtol=1e-12, atol=1e-18)
    assert np.array_equal(m[irho], m0[irho])
    ghost = np.ones(m.shape[1:], dtype=bool)
    ghost[inner] = False
    assert np.array_equal(m[:, ghost], m0[:, ghost])
