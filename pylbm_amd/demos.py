"""pylbm's end-to-end demos as scheme dictionaries.

pylbm's h5diff tests (tests/test_demo_1d.py, test_demo_2d.py, test_demo_3d.py)
run ``run(dx, Tf)`` of a demo script and compare the conserved moments, with
solid cells set to 0, against the fixtures in tests/reference/*.h5 at
``rel=1e-14, abs=1e-7`` (tests/conftest.py:233-262, 306-307).  Each function
here returns the dictionary that demo builds (same keys, symbols, values and
boundary functions) for a given ``dx``; :data:`DEMOS` lists the fixture name,
``dx``, ``Tf`` and the conserved-moment symbols.  The fixtures are converted to
.npz by tests/golden/make_h5_fixtures.py and checked by
tests/test_reference_fixtures.py.
"""
import math

import numpy as np
import sympy as sp

from . import bc
from .elements import Parallelogram

# ----------------------------------------------------------------- 1D


def advection_reaction(dx, generator="hip"):
    """D1Q2 advection with the logistic source mu u (1 - u) on the 1D torus
    (demo/1D/advection_reaction.py:26-100)."""
    X, U = sp.symbols("X, u")
    C, MU, LA = sp.symbols("c, mu, lambda", constants=True)
    xmin, xmax = 0.0, 1.0

    def u_init(x, xmin, xmax):
        middle = 0.5 * (xmin + xmax)
        width = 0.1 * (xmax - xmin)
        x_centered = xmin + x % (xmax - xmin)
        middle = 0.5 * (xmin + xmax)
        return 0.25 + 0.125 / width ** 10 * (x_centered - middle - width) ** 5 * (
            middle - x_centered - width) ** 5 * (abs(x_centered - middle) <= width)

    return {
        "box": {"x": [xmin, xmax], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": [1, 2],
            "conserved_moments": U,
            "polynomials": [1, X],
            "relaxation_parameters": [0.0, 2.0],
            "equilibrium": [U, C * U],
            "source_terms": {U: MU * U * (1 - U)},
        }],
        "init": {U: (u_init, (xmin, xmax))},
        "generator": generator,
        "parameters": {LA: 1.0, C: 0.25, MU: 1.0},
    }, [U]


def _riemann_pb(x, xmid, left, right):
    """Riemann initial data: ``left`` for x < xmid, ``right`` from xmid on
    (demo/1D/riemann_problems/exact_solvers/riemann_solvers.py:21-49)."""
    out = np.empty(x.shape)
    out[x < xmid] = left
    out[x >= xmid] = right
    return out


def _riemann_1d(dx, generator, xmin, xmax, la, conserved, eq1, sigmas, states, params):
    """The dictionary shared by pylbm's 1D Riemann problems: one D1Q2 scheme per
    conserved moment with relaxation 1/(1/2 + sigma), Neumann ends and Riemann initial
    data (demo/1D/riemann_problems/*.py, run())."""
    X = sp.symbols("X")
    LA = sp.symbols("lambda", constants=True)
    xmid = states.pop("xmid", 0.5 * (xmin + xmax))
    schemes = [{"velocities": [1, 2], "conserved_moments": w, "polynomials": [1, X],
                "relaxation_parameters": [0, 1 / (0.5 + sg)], "equilibrium": [w, e]}
               for w, e, sg in zip(conserved, eq1, sigmas)]
    return {
        "box": {"x": [xmin, xmax], "label": 0},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": schemes,
        "init": {w: (_riemann_pb, (xmid,) + tuple(states[w])) for w in conserved},
        "boundary_conditions": {0: {"method": {k: bc.Neumann for k in range(len(conserved))}}},
        "generator": generator,
        "parameters": {**params, LA: la},
    }, list(conserved)


def riemann_advection(dx, generator="hip"):
    """Riemann problem of the advection equation at velocity 1/4, s = 1.9
    (demo/1D/riemann_problems/advection.py:54-115)."""
    U = sp.symbols("u")
    C, SIGMA = sp.symbols("c, sigma", constants=True)
    return _riemann_1d(dx, generator, 0.0, 1.0, 1.0, [U], [C * U], [SIGMA],
                       {"xmid": 0.25, U: (1.0, 0.0)}, {C: 0.25, SIGMA: 1 / 1.9 - 0.5})


def riemann_burgers(dx, generator="hip"):
    """Riemann problem of Burgers' equation (a rarefaction from 0 to 0.3), s = 1.8
    (demo/1D/riemann_problems/burgers.py:53-107)."""
    U = sp.symbols("u")
    SIGMA = sp.symbols("sigma", constants=True)
    return _riemann_1d(dx, generator, -1.0, 1.0, 1.0, [U], [U ** 2 / 2], [SIGMA],
                       {U: (0.0, 0.3)}, {SIGMA: 1 / 1.8 - 0.5})


def riemann_euler(dx, generator="hip"):
    """Sod's shock tube at lambda = 3 with s = (1.9, 1.5, 1.4)
    (demo/1D/riemann_problems/euler.py:56-158)."""
    RHO, Q, E = sp.symbols("rho, q, E")
    GAMMA = sp.symbols("gamma", constants=True)
    S1, S2, S3 = sp.symbols("sigma_1, sigma_2, sigma_3", constants=True)
    gamma = 1.4
    rho_l, p_l, u_l = 1, 1, 0
    rho_r, p_r, u_r = 1 / 8, 0.1, 0
    states = {RHO: (rho_l, rho_r), Q: (rho_l * u_l, rho_r * u_r),
              E: (rho_l * u_l ** 2 + p_l / (gamma - 1.0), rho_r * u_r ** 2 + p_r / (gamma - 1.0))}
    eq1 = [Q, (GAMMA - 1) * E + (3 - GAMMA) / 2 * Q ** 2 / RHO,
           GAMMA * E * Q / RHO - (GAMMA - 1) / 2 * Q ** 3 / RHO ** 2]
    return _riemann_1d(dx, generator, 0.0, 1.0, 3.0, [RHO, Q, E], eq1, [S1, S2, S3], states,
                       {S1: 1 / 1.9 - 0.5, S2: 1 / 1.5 - 0.5, S3: 1 / 1.4 - 0.5, GAMMA: gamma})


def riemann_euler_isothermal(dx, generator="hip"):
    """Isothermal Euler Riemann problem, sound speed lambda / sqrt(3), s = (2, 1.85)
    (demo/1D/riemann_problems/euler_isothermal.py:59-145)."""
    RHO, Q = sp.symbols("rho, q")
    CO = sp.symbols("c_0", constants=True)
    S1, S2 = sp.symbols("sigma_1, sigma_2", constants=True)
    rhoo, drho = 1, 1 / 3
    states = {RHO: (rhoo - drho, rhoo + drho), Q: (0.0, 0.0)}
    return _riemann_1d(dx, generator, -1, 1, 1.0, [RHO, Q], [Q, Q ** 2 / RHO + CO ** 2 * RHO], [S1, S2],
                       states, {S1: 1 / 2.0 - 0.5, S2: 1 / 1.85 - 0.5, CO: 1.0 / math.sqrt(3)})


def riemann_p_system(dx, generator="hip"):
    """p-system with p(u1) = u1^(-2/3) at lambda = 2, s = 1.9
    (demo/1D/riemann_problems/p_system.py:52-128)."""
    U1, U2 = sp.symbols("u_1, u_2")
    GAMMA = sp.symbols("gamma", constants=True)
    S1, S2 = sp.symbols("sigma_1, sigma_2", constants=True)
    return _riemann_1d(dx, generator, 0.0, 1.0, 2.0, [U1, U2], [-U2, U1 ** (-GAMMA)], [S1, S2],
                       {U1: (1.50, 0.50), U2: (1.25, 1.50)},
                       {GAMMA: 2.0 / 3.0, S1: 1 / 1.9 - 0.5, S2: 1 / 1.9 - 0.5})


def riemann_shallow_water(dx, generator="hip"):
    """Shallow-water dam break (h 2 -> 1, q 0.1 -> 0) at lambda = 2, s = 1.5
    (demo/1D/riemann_problems/shallow_water.py:52-128)."""
    H, Q = sp.symbols("h, q")
    G = sp.symbols("g", constants=True)
    S1, S2 = sp.symbols("sigma_1, sigma_2", constants=True)
    return _riemann_1d(dx, generator, -1, 1.0, 2.0, [H, Q], [Q, Q ** 2 / H + 0.5 * G * H ** 2], [S1, S2],
                       {H: (2, 1), Q: (0.1, 0.0)}, {G: 1, S1: 1 / 1.5 - 0.5, S2: 1 / 1.5 - 0.5})


# ----------------------------------------------------------------- 2D


def poiseuille(dx, generator="hip"):
    """D2Q9 MRT channel with Bouzidi walls, a Bouzidi inflow carrying the
    parabolic profile and a NeumannX outflow (demo/2D/Poiseuille.py:12-103)."""
    X, Y, LA = sp.symbols("X, Y, lambda")
    rho, qx, qy = sp.symbols("rho, qx, qy")

    def bc_in(f, m, x, y, width, height, max_velocity, grad_pressure):
        m[rho] = (x - 0.5 * width) * grad_pressure
        m[qx] = max_velocity * (1.0 - 4.0 * y ** 2 / height ** 2)

    la = 1.0
    width, height = 2, 1
    max_velocity = 0.1
    rhoo = 1.0
    mu = 0.00185
    zeta = 1.0e-2
    xmin, xmax, ymin, ymax = 0.0, width, -0.5 * height, 0.5 * height
    grad_pressure = -max_velocity * 8.0 / height ** 2 * 3.0 / (la ** 2 * rhoo) * mu
    dummy = 3.0 / (la * rhoo * dx)
    s1 = 1.0 / (0.5 + zeta * dummy)
    s2 = 1.0 / (0.5 + mu * dummy)
    dummy = 1.0 / (LA ** 2 * rhoo)
    qx2 = dummy * qx ** 2
    qy2 = dummy * qy ** 2
    q2 = qx2 + qy2
    qxy = dummy * qx * qy
    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "label": [2, 1, 0, 0]},
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [{
            "velocities": list(range(9)),
            "polynomials": [
                1, LA * X, LA * Y,
                3 * (X ** 2 + Y ** 2) - 4,
                0.5 * (9 * (X ** 2 + Y ** 2) ** 2 - 21 * (X ** 2 + Y ** 2) + 8),
                3 * X * (X ** 2 + Y ** 2) - 5 * X,
                3 * Y * (X ** 2 + Y ** 2) - 5 * Y,
                X ** 2 - Y ** 2,
                X * Y,
            ],
            "relaxation_parameters": [0.0, 0.0, 0.0, s1, s1, s1, s1, s2, s2],
            "equilibrium": [rho, qx, qy, -2 * rho + 3 * q2, rho - 3 * q2, -qx / LA, -qy / LA,
                            qx2 - qy2, qxy],
            "conserved_moments": [rho, qx, qy],
        }],
        "parameters": {LA: la},
        "init": {rho: 1.0, qx: 0.0, qy: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack}},
            1: {"method": {0: bc.NeumannX}},
            2: {"method": {0: bc.BouzidiBounceBack},
                "value": (bc_in, (width, height, max_velocity, grad_pressure))},
        },
        "generator": generator,
    }, [rho, qx, qy]


def _smooth_bump(x, xmin, xmax, reg):
    """The initial bump of the 1D advection / Burgers demos (demo/1D/advection.py:26-66,
    demo/1D/burgers.py:26-66): a plateau of half-width w/2 around the point at a
    quarter of the box whose sides are the C^reg polynomial step
    sum_k (-1)^k C(reg, k) s^(2k+1) / (2k+1), normalised to 1 on the plateau."""
    middle, width = 0.75 * xmin + 0.25 * xmax, 0.125 * (xmax - xmin)
    left, right = middle - width, middle + width
    out = np.zeros(x.shape)
    il = np.where(np.logical_and(x > left, x <= middle))
    ir = np.where(np.logical_and(x < right, x > middle))
    sl = (x[il] - left - 0.5 * width) / (0.5 * width)
    sr = (x[ir] - middle - 0.5 * width) / (0.5 * width)
    pl, pr = np.copy(sl), np.copy(sr)
    sl2, sr2 = sl * sl, sr * sr
    cte = 0.0
    for k in range(reg + 1):
        coeff = (-1) ** k * math.comb(reg, k) / (2 * k + 1)
        out[il] += coeff * pl
        out[ir] -= coeff * pr
        cte += coeff
        pl *= sl2
        pr *= sr2
    out[il] += cte
    out[ir] += cte
    return out / (2 * cte)


def advection_1d(dx, generator="hip"):
    """D1Q2 advection at velocity 1/4 on the 1D torus, relaxation 1/(1/2 + sigma)
    (demo/1D/advection.py:69-118, the D1Q2 solution run() returns)."""
    U, X = sp.symbols("u, X")
    C, LA = sp.symbols("c, lambda", constants=True)
    SIGMA_0 = sp.symbols("sigma_0", constants=True)
    xmin, xmax = 0.0, 1.0
    return {
        "box": {"x": [xmin, xmax], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": [1, 2],
            "conserved_moments": U,
            "polynomials": [1, X],
            "relaxation_parameters": [0.0, 1 / (0.5 + SIGMA_0)],
            "equilibrium": [U, C * U],
        }],
        "init": {U: (_smooth_bump, (xmin, xmax, 0))},
        "generator": generator,
        "parameters": {LA: 1.0, C: 0.25, SIGMA_0: 1 / 2.0 - 0.5},
    }, [U]


def burgers_1d(dx, generator="hip"):
    """D1Q2 Burgers equation, equilibrium u^2 / 2, Neumann ends
    (demo/1D/burgers.py:69-121, the D1Q2 solution run() returns)."""
    U, X = sp.symbols("u, X")
    LA = sp.symbols("lambda", constants=True)
    SIGMA_0 = sp.symbols("sigma_0", constants=True)
    xmin, xmax = -1.0, 1.0
    return {
        "box": {"x": [xmin, xmax], "label": 0},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": [1, 2],
            "conserved_moments": U,
            "polynomials": [1, X],
            "relaxation_parameters": [0.0, 1 / (0.5 + SIGMA_0)],
            "equilibrium": [U, U ** 2 / 2],
        }],
        "init": {U: (_smooth_bump, (xmin, xmax, 1))},
        "boundary_conditions": {0: {"method": {0: bc.Neumann}}},
        "generator": generator,
        "parameters": {LA: 1.0, SIGMA_0: 1 / 1.8 - 0.5},
    }, [U]


def euler_1d(dx, generator="hip"):
    """Sod's shock tube: three D1Q2 schemes for (rho, q, E), Neumann ends
    (demo/1D/euler.py:35-127)."""
    X = sp.symbols("X")
    rho, q, E = sp.symbols("rho, q, E")
    LA = sp.symbols("lambda", constants=True)
    gamma = 1.4
    xmin, xmax = 0.0, 1.0
    la = 3.0
    rho_L, rho_R, p_L, p_R, u_L, u_R = 1.0, 1.0 / 8.0, 1.0, 0.1, 0.0, 0.0
    q_L, q_R = rho_L * u_L, rho_R * u_R
    E_L = rho_L * u_L ** 2 + p_L / (gamma - 1.0)
    E_R = rho_R * u_R ** 2 + p_R / (gamma - 1.0)

    def riemann(x, xmin, xmax, uL, uR):
        xm = 0.5 * (xmin + xmax)
        u = np.empty(x.shape)
        u[x < xm] = uL
        u[x == xm] = 0.5 * (uL + uR)
        u[x > xm] = uR
        return u

    def d1q2(c, s, eq1):
        return {"velocities": [1, 2], "conserved_moments": c, "polynomials": [1, LA * X],
                "relaxation_parameters": [0, s], "equilibrium": [c, eq1]}

    return {
        "box": {"x": [xmin, xmax], "label": 0},
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [
            d1q2(rho, 1.9, q),
            d1q2(q, 1.5, (gamma - 1.0) * E + 0.5 * (3.0 - gamma) * q ** 2 / rho),
            d1q2(E, 1.4, gamma * E * q / rho - 0.5 * (gamma - 1.0) * q ** 3 / rho ** 2),
        ],
        "init": {rho: (riemann, (xmin, xmax, rho_L, rho_R)), q: (riemann, (xmin, xmax, q_L, q_R)),
                 E: (riemann, (xmin, xmax, E_L, E_R))},
        "boundary_conditions": {0: {"method": {0: bc.Neumann, 1: bc.Neumann, 2: bc.Neumann}}},
        "parameters": {LA: la},
        "generator": generator,
    }, [rho, q, E]


def advection_3d(dx, generator="hip"):
    """D3Q6 advection at velocity (1/2, 1/5, 1/10) on the 3D torus from a ball of
    value 1 on a background of 1/2 (demo/3D/advection.py:40-113)."""
    u, X, Y, Z, LA = sp.symbols("u, X, Y, Z, lambda")
    ux, uy, uz = 0.5, 0.2, 0.1

    def u0(x, y, z):
        return 0.5 * np.ones((x.size, y.size, z.size)) + 0.5 * (
            ((x - 0.5) ** 2 + (y - 0.5) ** 2 + (z - 0.5) ** 2) < 0.25 ** 2)

    return {
        "box": {"x": [0.0, 1.0], "y": [0.0, 1.0], "z": [0.0, 1.0], "label": -1},
        "space_step": dx,
        "scheme_velocity": 1.0,
        "schemes": [{
            "velocities": list(range(1, 7)),
            "conserved_moments": [u],
            "polynomials": [1, LA * X, LA * Y, LA * Z, X ** 2 - Y ** 2, X ** 2 - Z ** 2],
            "equilibrium": [u, ux * u, uy * u, uz * u, 0.0, 0.0],
            "relaxation_parameters": [0.0, 1.0, 1.0, 1.0, 1.0, 1.0],
        }],
        "init": {u: u0},
        "parameters": {LA: 1.0},
        "generator": generator,
    }, [u]


def poiseuille_3d(dx, generator="hip"):
    """Poiseuille tunnel: four D3Q6 schemes for (p, ux, uy, uz), periodic in z,
    Bouzidi bounce-back / anti-bounce-back walls, a valued NeumannX end and a valued
    Bouzidi anti-bounce-back inflow (demo/3D/poiseuille.py:47-160)."""
    X, Y, Z, LA = sp.symbols("X,Y,Z,lambda")
    p, ux, uy, uz = sp.symbols("p,ux,uy,uz")
    width, height = 1.0, 0.5
    xmin, xmax, ymin, ymax = 0.0, width, -0.5 * height, 0.5 * height
    zmin, zmax = -2 * dx, 2 * dx
    la, max_velocity, mu, cte = 1.0, 0.1, 1.0e-3, 10.0
    grad_pressure = -mu * max_velocity * 8.0 / height ** 2
    s = 1.0 / (0.5 + 1.0 / np.sqrt(12))
    vs = [0.0, s, s, s, s, s]
    vel = list(range(1, 7))
    pol = [1, LA * X, LA * Y, LA * Z, X ** 2 - Y ** 2, X ** 2 - Z ** 2]

    def bc_in(f, m, x, y, z):
        m[p] = (x - 0.5 * width) * grad_pressure * cte
        m[ux] = max_velocity * (1.0 - 4.0 * y ** 2 / height ** 2)

    def bc_out(f, m, x, y, z):
        m[p] = (x - 0.5 * width) * grad_pressure * cte

    def scheme(w, eq):
        return {"velocities": vel, "conserved_moments": w, "polynomials": pol,
                "relaxation_parameters": vs, "equilibrium": eq + [0.0, 0.0]}

    abb = bc.BouzidiAntiBounceBack
    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "z": [zmin, zmax], "label": [1, 2, 0, 0, -1, -1]},
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [
            scheme(p, [p, ux, uy, uz]),
            scheme(ux, [ux, ux ** 2 + p / cte, ux * uy, ux * uz]),
            scheme(uy, [uy, uy * ux, uy ** 2 + p / cte, uy * uz]),
            scheme(uz, [uz, uz * ux, uz * uy, uz ** 2 + p / cte]),
        ],
        "init": {p: 0.0, ux: 0.0, uy: 0.0, uz: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack, 1: abb, 2: abb, 3: abb}},
            1: {"method": {0: abb, 1: bc.NeumannX, 2: bc.NeumannX, 3: bc.NeumannX}, "value": bc_out},
            2: {"method": {0: abb, 1: abb, 2: abb, 3: abb}, "value": bc_in},
        },
        "parameters": {LA: la},
        "generator": generator,
    }, [p, ux, uy, uz]


def _advection_2d_common(dx, generator):
    U, X, Y = sp.symbols("U, X, Y")
    CX, CY, LA = sp.symbols("c_0, c_1, lambda", constants=True)
    SIGMA_0, SIGMA_1 = sp.symbols("sigma_0, sigma_1", constants=True)
    s_q, s_xy = 1 / (0.5 + SIGMA_0), 1 / (0.5 + SIGMA_1)
    return U, CX, CY, LA, {
        "box": {"x": [0.0, 1.0], "y": [0.0, 1.0], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": list(range(1, 5)),
            "conserved_moments": U,
            "polynomials": [1, X, Y, X ** 2 - Y ** 2],
            "relaxation_parameters": [0.0, s_q, s_q, s_xy],
            "equilibrium": [U, CX * U, CY * U, 0],
        }],
        "generator": generator,
        "parameters": {LA: 2.0, CX: 0.2, CY: 0.5, SIGMA_0: 1.0e-2, SIGMA_1: 0.5},
        "relative_velocity": [CX, CY],
    }


def _disc(x, y, base, top):
    """base, raised to top on the disc of radius 0.1 around (1/4, 1/2)."""
    return base + (top - base) * ((x - 0.25) ** 2 + (y - 0.5) ** 2 < 0.1 ** 2)


def advection_2d(dx, generator="hip"):
    """D2Q4 advection at velocity (0.2, 0.5) with the relative velocity (0.2, 0.5)
    (a constant), periodic (demo/2D/advection.py:30-118)."""
    U, _, _, _, d = _advection_2d_common(dx, generator)
    d["init"] = {U: (lambda x, y: _disc(x, y, 1.0, 1.5), ())}
    return d, [U]


def advection_2d_init_f(dx, generator="hip"):
    """The same scheme initialised by its populations, f_k = (1/4 + c_k . c / (2 la)) u0
    (inittype 'distributions', demo/2D/advection_init_f.py:30-129)."""
    U, _, _, _, d = _advection_2d_common(dx, generator)
    cx, cy = 0.2 / 2.0, 0.5 / 2.0
    coeff = [0.25 + 0.5 * cx, 0.25 + 0.5 * cy, 0.25 - 0.5 * cx, 0.25 - 0.5 * cy]
    d["inittype"] = "distributions"
    d["init"] = {k: ((lambda x, y, c: c * _disc(x, y, 1.0, 1.5)), (coeff[k],)) for k in range(4)}
    return d, [U]


def _geier_d2q9(dx, la, rho_o, mu, zeta):
    """The 'Geier' D2Q9 moments of the lid-driven cavity and Karman demos (used with the
    relative velocity q / rho): polynomials, equilibrium, relaxation rates
    (demo/2D/lid_driven_cavity.py:137-165, demo/2D/Karman_vortex_street.py:138-166)."""
    X, Y = sp.symbols("X, Y")
    RHO, QX, QY = sp.symbols("rho, qx, qy")
    LA = sp.symbols("lambda", constants=True)
    UX, UY = QX / RHO, QY / RHO
    RHOU2 = RHO * (UX ** 2 + UY ** 2)
    polynomials = [1, X, Y, X ** 2 + Y ** 2, X * Y ** 2, Y * X ** 2, X ** 2 * Y ** 2, X ** 2 - Y ** 2, X * Y]
    equilibrium = [
        RHO, QX, QY,
        RHOU2 + 2 / 3 * RHO * LA ** 2,
        QX * (LA ** 2 / 3 + UY ** 2),
        QY * (LA ** 2 / 3 + UX ** 2),
        RHO * (LA ** 2 / 3 + UX ** 2) * (LA ** 2 / 3 + UY ** 2),
        RHO * (UX ** 2 - UY ** 2),
        RHO * UX * UY,
    ]
    dummy = 3.0 / (la * rho_o * dx)
    s_1 = 1 / (0.5 + dummy * (zeta - 2 * mu / 3))
    s_2 = 1 / (0.5 + dummy * mu)
    return polynomials, equilibrium, [0.0, 0.0, 0.0, s_1, s_1, s_1, s_1, s_2, s_2], (RHO, QX, QY, LA)


def lid_driven_cavity(dx, generator="hip"):
    """D2Q9 'Geier' scheme with the relative velocity q / rho, Bouzidi bounce-back walls
    and a Bouzidi lid carrying (rho_o, rho_o u_lid, 0) (demo/2D/lid_driven_cavity.py:27-246)."""
    la, rho_o, driven_velocity = 1.0, 1.0, 0.05
    mu = 5.0e-6
    zeta = 100 * mu
    polynomials, equilibrium, s, (RHO, QX, QY, LA) = _geier_d2q9(dx, la, rho_o, mu, zeta)

    def bc_up(f, m, x, y, rho_o, driven_velocity):
        m[RHO] = rho_o
        m[QX] = rho_o * driven_velocity
        m[QY] = 0.0

    return {
        "parameters": {LA: la},
        "box": {"x": [0.0, 1.0], "y": [0.0, 1.0], "label": [0, 0, 0, 1]},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": list(range(9)),
            "polynomials": polynomials,
            "relaxation_parameters": s,
            "equilibrium": equilibrium,
            "conserved_moments": [RHO, QX, QY],
        }],
        "init": {RHO: rho_o, QX: 0.0, QY: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack}},
            1: {"method": {0: bc.BouzidiBounceBack}, "value": (bc_up, (rho_o, driven_velocity))},
        },
        "generator": generator,
        "relative_velocity": [QX / RHO, QY / RHO],
    }, [RHO, QX, QY]


def karman_vortex_street(dx, generator="hip"):
    """D2Q9 'Geier' scheme with the relative velocity q / rho around a cylinder: Bouzidi
    bounce-back inflow and walls carrying (rho_o, rho_o u_o, 0), NeumannX outflow,
    Bouzidi bounce-back on the circle (demo/2D/Karman_vortex_street.py:23-233)."""
    from .domain import Circle
    la, rho_o, u_o = 1.0, 1.0, 0.05
    radius = 1.0 / 32
    mu = 5.0e-6
    zeta = 10 * mu
    xmin, xmax, ymin, ymax = 0.0, 2.0, 0.0, 1.0
    polynomials, equilibrium, s, (RHO, QX, QY, LA) = _geier_d2q9(dx, la, rho_o, mu, zeta)

    def bc_in(f, m, x, y, rho_o, u_o):
        m[RHO] = rho_o
        m[QX] = rho_o * u_o
        m[QY] = 0.0

    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "label": [0, 1, 0, 0]},
        "elements": [Circle([0.3, 0.5 * (ymin + ymax) + 2 * dx], radius, label=2)],
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [{
            "velocities": list(range(9)),
            "polynomials": polynomials,
            "relaxation_parameters": s,
            "equilibrium": equilibrium,
            "conserved_moments": [RHO, QX, QY],
        }],
        "parameters": {LA: la},
        "init": {RHO: rho_o, QX: rho_o * u_o, QY: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack}, "value": (bc_in, (rho_o, u_o))},
            1: {"method": {0: bc.NeumannX}},
            2: {"method": {0: bc.BouzidiBounceBack}},
        },
        "generator": generator,
        "relative_velocity": [QX / RHO, QY / RHO],
    }, [RHO, QX, QY]


def bend(dx, generator="hip"):
    """Flow through a bend: D2Q9 'Geier' scheme with the relative velocity q / rho in
    the unit box minus a Parallelogram block, parabolic Bouzidi inflow on the upper
    left, NeumannY outflow at the bottom (demo/2D/bend.py:24-275, scheme 'Geier')."""
    la, rho_o, u_o = 1.0, 1.0, 0.05
    radius = 0.25
    mu = 2.5e-6
    zeta = 100 * mu
    xmin, xmax, ymin, ymax = 0.0, 1.0, 0.0, 1.0
    polynomials, equilibrium, s, (RHO, QX, QY, LA) = _geier_d2q9(dx, la, rho_o, mu, zeta)
    xc, yc = xmax - radius, ymax - radius

    def bc_in(f, m, x, y, rho_o, u_o, ymin, ymax, radius):
        y_bound = ymax - radius
        m[RHO] = rho_o
        m[QX] = rho_o * u_o * 4 * (ymax - y) * (y - y_bound) / (ymax - y_bound) ** 2
        m[QY] = 0.0

    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "label": [2, 0, 1, 0]},
        "elements": [Parallelogram((xmin, ymin), (xc, ymin), (xmin, yc), label=0)],
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [{
            "velocities": list(range(9)),
            "polynomials": polynomials,
            "relaxation_parameters": s,
            "equilibrium": equilibrium,
            "conserved_moments": [RHO, QX, QY],
        }],
        "parameters": {LA: la},
        "init": {RHO: rho_o, QX: rho_o * u_o, QY: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack}},
            1: {"method": {0: bc.NeumannY}},
            2: {"method": {0: bc.BouzidiBounceBack}, "value": (bc_in, (rho_o, u_o, ymin, ymax, radius))},
        },
        "generator": generator,
        "relative_velocity": [QX / RHO, QY / RHO],
    }, [RHO, QX, QY]


def kelvin_helmholtz(dx, generator="hip"):
    """D2Q9 'Geier' scheme with the relative velocity q / rho on the torus: two
    tanh shear layers and a sinusoidal transverse perturbation, Mach 0.1
    (demo/2D/Kelvin_Helmoltz.py:33-230)."""
    la = math.sqrt(3) / 0.1
    rho_o, u_o, zeta, mu, k, delta = 1.0, 0.5, 0.0366, 1.0e-6, 80, 0.05
    polynomials, equilibrium, s, (RHO, QX, QY, LA) = _geier_d2q9(dx, la, rho_o, mu, zeta)

    def qx_init(x, y, U, k):
        return np.zeros_like(x) + U * np.tanh(k * (y - 0.25)) * (y <= 0.5) + U * np.tanh(k * (0.75 - y)) * (y > 0.5)

    def qy_init(x, y, U, delta):
        return np.zeros_like(y) + U * delta * np.sin(2 * np.pi * (x + 0.25))

    return {
        "parameters": {LA: la},
        "box": {"x": [0.0, 1.0], "y": [0.0, 1.0], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [{
            "velocities": list(range(9)),
            "polynomials": polynomials,
            "relaxation_parameters": s,
            "equilibrium": equilibrium,
            "conserved_moments": [RHO, QX, QY],
        }],
        "init": {RHO: rho_o, QX: (qx_init, (u_o, k)), QY: (qy_init, (u_o, delta))},
        "relative_velocity": [QX / RHO, QY / RHO],
        "generator": generator,
    }, [RHO, QX, QY]


def shallow_water(dx, generator="hip"):
    """Shallow water: three D2Q4 schemes for (h, qx, qy) sharing the relative
    velocity q / h (a vectorial relative-velocity scheme), on the torus from a raised
    disc (demo/2D/shallow_water.py:28-125)."""
    X, Y = sp.symbols("X, Y")
    H, QX, QY = sp.symbols("h, qx, qy")
    LA, G = sp.symbols("lambda, g", constants=True)
    SHX, SHXY, SQX, SQXY = sp.symbols("sigma_0, sigma_1, sigma_2, sigma_3", constants=True)
    s_h = [0.0, 1 / (0.5 + SHX), 1 / (0.5 + SHX), 1 / (0.5 + SHXY)]
    s_q = [0.0, 1 / (0.5 + SQX), 1 / (0.5 + SQX), 1 / (0.5 + SQXY)]
    vel, pol = list(range(1, 5)), [1, X, Y, X ** 2 - Y ** 2]

    def h_init(x, y):
        return 1 + 0.5 * (x ** 2 + y ** 2 < 0.1 ** 2)

    return {
        "parameters": {LA: 4, G: 1.0, SHX: 1.0e-3, SHXY: 0.5, SQX: 1.0e-1, SQXY: 0.5},
        "box": {"x": [-1.0, 1.0], "y": [-1.0, 1.0], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [
            {"velocities": vel, "conserved_moments": H, "polynomials": pol, "relaxation_parameters": s_h,
             "equilibrium": [H, QX, QY, 0.0]},
            {"velocities": vel, "conserved_moments": QX, "polynomials": pol, "relaxation_parameters": s_q,
             "equilibrium": [QX, QX ** 2 / H + G * H ** 2 / 2, QX * QY / H, 0.0]},
            {"velocities": vel, "conserved_moments": QY, "polynomials": pol, "relaxation_parameters": s_q,
             "equilibrium": [QY, QX * QY / H, QY ** 2 / H + G * H ** 2 / 2, 0.0]},
        ],
        "init": {H: (h_init, ()), QX: 0.0, QY: 0.0},
        "relative_velocity": [QX / H, QY / H],
        "generator": generator,
    }, [H, QX, QY]


def orszag_tang(dx, generator="hip"):
    """Orszag-Tang vortex of ideal MHD: six D2Q4 schemes for (rho, qx, qy, E, Bx, By)
    on the 2 pi torus (demo/2D/Orszag_Tang_vortex.py:24-195)."""
    GA, X, Y, LA = sp.symbols("GA, X, Y, lambda")
    rho, qx, qy, E, Bx, By = sp.symbols("rho, qx, qy, E, Bx, By")
    gamma, la, s = 5.0 / 3.0, 10.0, 1.95
    p = (GA - 1) * (E - (qx ** 2 + qy ** 2) / (2 * rho) - (Bx ** 2 + By ** 2) / 2)
    ps = p + (Bx ** 2 + By ** 2) / 2
    vB = (qx * Bx + qy * By) / rho

    def d2q4(c, eq):
        return {"velocities": list(range(1, 5)), "conserved_moments": c,
                "polynomials": [1, LA * X, LA * Y, X ** 2 - Y ** 2], "relaxation_parameters": [0.0, s, s, s],
                "equilibrium": [c] + eq + [0.0]}

    def rho0(x, y, g):
        return g ** 2 * np.ones(x.shape)

    def qx0(x, y, g):
        return -(g ** 2) * np.sin(y)

    def qy0(x, y, g):
        return g ** 2 * np.sin(x)

    def bx0(x, y):
        return -np.sin(y)

    def by0(x, y):
        return np.sin(2 * x)

    def e0(x, y, g):
        return 0.5 * (qx0(x, y, g) ** 2 + qy0(x, y, g) ** 2) / rho0(x, y, g) + 0.5 * (bx0(x, y) ** 2 + by0(x, y) ** 2) \
            + g / (g - 1)

    return {
        "box": {"x": [0.0, 2 * math.pi], "y": [0.0, 2 * math.pi], "label": -1},
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [
            d2q4(rho, [qx, qy]),
            d2q4(qx, [qx ** 2 / rho + ps - Bx ** 2, qx * qy / rho - Bx * By]),
            d2q4(qy, [qx * qy / rho - Bx * By, qy ** 2 / rho + ps - By ** 2]),
            d2q4(E, [(E + ps) * qx / rho - vB * Bx, (E + ps) * qy / rho - vB * By]),
            d2q4(Bx, [0, (qy * Bx - qx * By) / rho]),
            d2q4(By, [(qx * By - qy * Bx) / rho, 0]),
        ],
        "init": {rho: (rho0, (gamma,)), qx: (qx0, (gamma,)), qy: (qy0, (gamma,)), E: (e0, (gamma,)),
                 Bx: bx0, By: by0},
        "parameters": {LA: la, GA: gamma},
        "generator": generator,
    }, [rho, qx, qy, E, Bx, By]


def poiseuille_vec(dx, generator="hip"):
    """Vectorial Poiseuille channel: three D2Q4 schemes for (p, ux, uy) with Bouzidi
    bounce-back / anti-bounce-back walls per scheme, a valued inflow and a valued
    NeumannX / Bouzidi outflow (demo/2D/Poiseuille_vec.py:24-139)."""
    X, Y, LA = sp.symbols("X, Y, lambda")
    p, ux, uy = sp.symbols("p, ux, uy")
    width, height = 1.0, 0.5
    la, max_velocity, mu, zeta, cte = 1.0, 0.1, 0.00185, 1.0e-5, 3.0
    grad_pressure = -mu * max_velocity * 8.0 / height ** 2
    dummy = 3.0 / (la * dx)
    s1 = 1.0 / (0.5 + zeta * dummy)
    s2 = 1.0 / (0.5 + mu * dummy)
    vel, pol = list(range(1, 5)), [1, LA * X, LA * Y, X ** 2 - Y ** 2]

    def bc_in(f, m, x, y, width, height, max_velocity, grad_pressure, cte):
        m[p] = (x - 0.5 * width) * grad_pressure * cte
        m[ux] = max_velocity * (1.0 - 4.0 * y ** 2 / height ** 2)

    def bc_out(f, m, x, y, width, grad_pressure, cte):
        m[p] = (x - 0.5 * width) * grad_pressure * cte

    return {
        "box": {"x": [0.0, width], "y": [-0.5 * height, 0.5 * height], "label": [2, 1, 0, 0]},
        "space_step": dx,
        "scheme_velocity": la,
        "schemes": [
            {"velocities": vel, "polynomials": pol, "relaxation_parameters": [0.0, s1, s1, 1.0],
             "equilibrium": [p, ux, uy, 0.0], "conserved_moments": p},
            {"velocities": vel, "polynomials": pol, "relaxation_parameters": [0.0, s2, s2, 1.0],
             "equilibrium": [ux, ux ** 2 + p / cte, ux * uy, 0.0], "conserved_moments": ux},
            {"velocities": vel, "polynomials": pol, "relaxation_parameters": [0.0, s2, s2, 1.0],
             "equilibrium": [uy, ux * uy, uy ** 2 + p / cte, 0.0], "conserved_moments": uy},
        ],
        "parameters": {LA: la},
        "init": {p: 0.0, ux: 0.0, uy: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack, 1: bc.BouzidiAntiBounceBack, 2: bc.BouzidiAntiBounceBack}},
            1: {"method": {0: bc.BouzidiAntiBounceBack, 1: bc.NeumannX, 2: bc.NeumannX},
                "value": (bc_out, (width, grad_pressure, cte))},
            2: {"method": {0: bc.BouzidiAntiBounceBack, 1: bc.BouzidiAntiBounceBack, 2: bc.BouzidiAntiBounceBack},
                "value": (bc_in, (width, height, max_velocity, grad_pressure, cte))},
        },
        "generator": generator,
    }, [p, ux, uy]


def rayleigh_benard(dx, generator="hip"):
    """Rayleigh-Benard convection: D2Q9 MRT flow with the buoyancy source
    S(qy) = alpha g T coupled to a D2Q5 temperature, periodic in x, Bouzidi
    bounce-back / anti-bounce-back plates, the lower plate's temperature a
    time-dependent value (demo/2D/Rayleigh-Benard.py:54-231, time_bc=True)."""
    X, Y = sp.symbols("X, Y")
    rho, qx, qy, T = sp.symbols("rho, qx, qy, T")
    Td, Tu = 0.5, -0.5
    xmin, xmax, ymin, ymax = 0.0, 2.0, 0.0, 1.0
    Ra, Pr, alpha, g = 2000, 0.71, 0.005, 9.81
    nu = np.sqrt(Pr * alpha * 9.81 * (Td - Tu) * (ymax - ymin) / Ra)
    kappa = nu / Pr
    snu = 1.0 / (0.5 + 3 * nu)
    seta = 1.0 / (0.5 + 3 * nu)
    sq = 8 * (2 - snu) / (8 - snu)
    sf = [0.0, 0.0, 0.0, seta, seta, sq, sq, snu, snu]
    a = 0.5
    skappa = 1.0 / (0.5 + 10 * kappa / (4 + a))
    se = 1.0 / (0.5 + np.sqrt(3) / 3)
    sT = [0.0, skappa, skappa, se, se]

    def init_T(x, y):
        xmid, ymid = (xmax + xmin) / 2, (ymax + ymin) / 2
        return Td + (Tu - Td) / (ymax - ymin) * (y - ymin) + Td * (x < 1.1 * xmid) * (x > 0.9 * xmid) * (y < ymid)

    def bc_up(f, m, x, y, Tu):
        m[qx] = 0.0
        m[qy] = 0.0
        m[T] = Tu

    def bc_down_with_time(f, m, t, x, y, Td, Tu):
        m[qx] = 0.0
        m[qy] = 0.0
        m[T] = Td if 0 <= t % 10.0 <= 5 else Tu

    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "label": [-1, -1, 0, 1]},
        "space_step": dx,
        "scheme_velocity": 1.0,
        "schemes": [
            {"velocities": list(range(9)), "conserved_moments": [rho, qx, qy],
             "polynomials": [1, X, Y, 3 * (X ** 2 + Y ** 2) - 4,
                             0.5 * (9 * (X ** 2 + Y ** 2) ** 2 - 21 * (X ** 2 + Y ** 2) + 8),
                             3 * X * (X ** 2 + Y ** 2) - 5 * X, 3 * Y * (X ** 2 + Y ** 2) - 5 * Y,
                             X ** 2 - Y ** 2, X * Y],
             "relaxation_parameters": sf,
             "equilibrium": [rho, qx, qy, -2 * rho + 3 * (qx ** 2 + qy ** 2), rho - 3 * (qx ** 2 + qy ** 2), -qx, -qy,
                             qx ** 2 - qy ** 2, qx * qy],
             "source_terms": {qy: alpha * g * T}},
            {"velocities": list(range(5)), "conserved_moments": T,
             "polynomials": [1, X, Y, 5 * (X ** 2 + Y ** 2) - 4, (X ** 2 - Y ** 2)],
             "equilibrium": [T, T * qx, T * qy, a * T, 0.0], "relaxation_parameters": sT},
        ],
        "init": {rho: 1.0, qx: 0.0, qy: 0.0, T: (init_T, ())},
        "boundary_conditions": {
            0: {"method": {0: bc.BouzidiBounceBack, 1: bc.BouzidiAntiBounceBack},
                "value": (bc_down_with_time, (Td, Tu)), "time_bc": True},
            1: {"method": {0: bc.BouzidiBounceBack, 1: bc.BouzidiAntiBounceBack}, "value": (bc_up, (Tu,))},
        },
        "generator": generator,
    }, [rho, qx, qy, T]


def air_conditioning(dx, generator="hip"):
    """Air conditioning: the Rayleigh-Benard flow / temperature pair in the unit
    box with two wall blocks (Parallelogram obstacles, Bouzidi walls at T0), a
    Bouzidi inflow of cold air on the left and NeumannX on the right
    (demo/2D/air_conditioning.py:10-175)."""
    X, Y = sp.symbols("X, Y")
    rho, qx, qy, T = sp.symbols("rho, qx, qy, T")
    T0, Tin = 0.5, -0.5
    xmin, xmax, ymin, ymax = 0.0, 1.0, 0.0, 1.0
    Ra, Pr, alpha, rhoo, g, uo = 2000, 0.71, 0.005, 1.0, 9.81, 0.025
    nu = np.sqrt(Pr * alpha * 9.81 * (T0 - Tin) * (ymax - ymin) / Ra)
    kappa = nu / Pr
    snu = 1.0 / (0.5 + 3 * nu)
    seta = 1.0 / (0.5 + 3 * nu)
    sq = 8 * (2 - snu) / (8 - snu)
    sf = [0.0, 0.0, 0.0, seta, seta, sq, sq, snu, snu]
    a = 0.5
    skappa = 1.0 / (0.5 + 10 * kappa / (4 + a))
    se = 1.0 / (0.5 + np.sqrt(3) / 3)
    sT = [0.0, skappa, skappa, se, se]

    def init_T(x, y, T0):
        return T0

    def bc_wall(f, m, x, y, T0):
        m[qx] = 0.0
        m[qy] = 0.0
        m[T] = T0

    def bc_in(f, m, x, y, T0, Tin, ymax, rhoo, uo):
        m[qx] = rhoo * uo
        m[qy] = 0.0
        m[T] = T0 + (Tin - T0) * (ymax - y) * (y - 0.8) * 100

    walls = {0: bc.BouzidiBounceBack, 1: bc.BouzidiAntiBounceBack}
    return {
        "box": {"x": [xmin, xmax], "y": [ymin, ymax], "label": [1, 2, 0, 0]},
        "elements": [Parallelogram([xmin, ymin], [0.1, 0], [0, 0.8], label=0),
                     Parallelogram([xmax, ymin], [-0.1, 0], [0, 0.8], label=0)],
        "space_step": dx,
        "scheme_velocity": 1.0,
        "schemes": [
            {"velocities": list(range(9)), "conserved_moments": [rho, qx, qy],
             "polynomials": [1, X, Y, 3 * (X ** 2 + Y ** 2) - 4,
                             sp.Rational(1, 2) * (9 * (X ** 2 + Y ** 2) ** 2 - 21 * (X ** 2 + Y ** 2) + 8),
                             3 * X * (X ** 2 + Y ** 2) - 5 * X, 3 * Y * (X ** 2 + Y ** 2) - 5 * Y,
                             X ** 2 - Y ** 2, X * Y],
             "relaxation_parameters": sf,
             "equilibrium": [rho, qx, qy, -2 * rho + 3 * (qx ** 2 + qy ** 2), rho - 3 * (qx ** 2 + qy ** 2), -qx, -qy,
                             qx ** 2 - qy ** 2, qx * qy],
             "source_terms": {qy: alpha * g * T}},
            {"velocities": list(range(5)), "conserved_moments": T,
             "polynomials": [1, X, Y, 5 * (X ** 2 + Y ** 2) - 4, (X ** 2 - Y ** 2)],
             "equilibrium": [T, T * qx, T * qy, a * T, 0.0], "relaxation_parameters": sT},
        ],
        "init": {rho: 1.0, qx: 0.0, qy: 0.0, T: (init_T, (T0,))},
        "boundary_conditions": {
            0: {"method": dict(walls), "value": (bc_wall, (T0,))},
            1: {"method": dict(walls), "value": (bc_in, (T0, Tin, ymax, rhoo, uo))},
            2: {"method": {0: bc.NeumannX, 1: bc.NeumannX}},
        },
        "generator": generator,
    }, [rho, qx, qy, T]


# name of the fixture -> (dictionary builder, dx, Tf)
DEMOS = {
    "test1D_advection": (advection_1d, 1.0 / 64, 0.5),
    "test1D_advection_reaction": (advection_reaction, 1.0 / 64, 0.5),
    "test1D_burgers": (burgers_1d, 1.0 / 64, 0.5),
    "test1D_euler": (euler_1d, 1.0 / 64, 0.5),
    "test1D_riemann_advection": (riemann_advection, 1.0 / 64, 0.5),
    "test1D_riemann_burgers": (riemann_burgers, 1.0 / 64, 0.5),
    "test1D_riemann_euler": (riemann_euler, 1.0 / 64, 0.5),
    "test1D_riemann_euler_isothermal": (riemann_euler_isothermal, 1.0 / 64, 0.5),
    "test1D_riemann_p_system": (riemann_p_system, 1.0 / 64, 0.5),
    "test1D_riemann_shallow_water": (riemann_shallow_water, 1.0 / 64, 0.5),
    "test3D_advection": (advection_3d, 1.0 / 64, 0.5),
    "test3D_poseuille": (poiseuille_3d, 1.0 / 64, 0.5),
    "test2D_advection": (advection_2d, 1.0 / 64, 0.5),
    "test2D_advection_init_f": (advection_2d_init_f, 1.0 / 64, 0.5),
    "test2D_kelvin_Helmoltz": (kelvin_helmholtz, 1.0 / 64, 0.5),
    "test2D_orszag_Tang_vortex": (orszag_tang, 2.0 * math.pi / 64, 0.5),
    "test2D_poiseuille": (poiseuille, 1.0 / 64, 0.5),
    "test2D_poiseuille_vec": (poiseuille_vec, 1.0 / 64, 0.5),
    "test2D_rayleigh_benard": (rayleigh_benard, 1.0 / 64, 0.5),
    "test2D_air_conditioning": (air_conditioning, 1.0 / 64, 0.5),
    "test2D_coude": (bend, 1.0 / 64, 0.5),
    "test2D_shallow_water": (shallow_water, 1.0 / 64, 0.5),
    "test2D_lid_driven_cavity": (lid_driven_cavity, 1.0 / 64, 0.5),
    "test2D_karman_vortex_street": (karman_vortex_street, 1.0 / 64, 0.5),
}
