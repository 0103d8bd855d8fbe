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

import sympy as sp

from . import bc

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


# name of the fixture -> (dictionary builder, dx, Tf)
DEMOS = {
    "test1D_advection_reaction": (advection_reaction, 1.0 / 64, 0.5),
    "test2D_poiseuille": (poiseuille, 1.0 / 64, 0.5),
    "test2D_lid_driven_cavity": (lid_driven_cavity, 1.0 / 64, 0.5),
    "test2D_karman_vortex_street": (karman_vortex_street, 1.0 / 64, 0.5),
}
