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


# name of the fixture -> (dictionary builder, dx, Tf)
DEMOS = {
    "test1D_advection_reaction": (advection_reaction, 1.0 / 64, 0.5),
    "test2D_poiseuille": (poiseuille, 1.0 / 64, 0.5),
}
