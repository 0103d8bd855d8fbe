"""Scheme dictionaries of the workloads this time-stepper is measured on.

Each function returns a pylbm-format dict (the same keys pylbm's
``Simulation`` takes: box, space_step, scheme_velocity, schemes, init,
boundary_conditions, generator), in lattice units (dx = 1, la = 1, so dt = 1).

* :func:`lid_cavity_d2q9`  -- D2Q9 MRT (d'Humieres basis of
  demo/2D/lid_driven_cavity.py:95-127), lid u = 0.05, bounce-back walls.
* :func:`karman_d2q9`      -- D2Q9 MRT channel, cylinder, anti-bounce-back
  inflow, NeumannX outflow, bounce-back walls and cylinder.
* :func:`taylor_green_d3q19` -- D3Q19 MRT periodic Taylor-Green vortex.
* :func:`sphere_packing_d3q19` -- D3Q19 MRT with body force through random
  spheres, periodic in x, bounce-back walls in y and z.
* :func:`decaying_d2q9`    -- D2Q9 MRT periodic decaying turbulence from a
  random Fourier velocity field.

pylbm ships no D3Q19 MRT dict (only the stencil and a D3Q15 MRT in
demo/3D/lid_cavity.py:61-126); :func:`d3q19_mrt` is the standard orthogonal
basis of d'Humieres et al. (2002) written in pylbm's format.
"""
import math

import numpy as np
import sympy as sp

from . import bc
from .elements import Circle, Parallelogram, Sphere

X, Y, Z = sp.symbols("X, Y, Z")
RHO, QX, QY, QZ = sp.symbols("rho, qx, qy, qz")
LA = sp.symbols("lambda", constants=True)

__all__ = ["d2q9_mrt", "d3q19_mrt", "lid_cavity_d2q9", "karman_d2q9", "taylor_green_d3q19",
           "sphere_packing_d3q19", "decaying_d2q9", "poiseuille_d2q9", "with_bouzidi", "RHO", "QX", "QY", "QZ"]


def d2q9_mrt(s_e, s_eps, s_q, s_nu, rho0=1.0):
    """D2Q9 d'Humieres MRT scheme entry (demo/2D/lid_driven_cavity.py:95-127)."""
    inv = 1.0 / rho0
    q2 = inv * QX ** 2 + inv * QY ** 2
    return {
        "velocities": list(range(9)),
        "conserved_moments": [RHO, QX, QY],
        "polynomials": [
            1, X, Y,
            3 * (X ** 2 + Y ** 2) - 4 * LA ** 2,
            sp.Rational(1, 2) * (9 * (X ** 2 + Y ** 2) ** 2 - 21 * (X ** 2 + Y ** 2) * LA ** 2 + 8 * LA ** 4),
            3 * X * (X ** 2 + Y ** 2) - 5 * X * LA ** 2,
            3 * Y * (X ** 2 + Y ** 2) - 5 * Y * LA ** 2,
            X ** 2 - Y ** 2,
            X * Y,
        ],
        "equilibrium": [
            RHO, QX, QY,
            -2 * RHO * LA ** 2 + 3 * q2,
            RHO * LA ** 2 - 3 * q2,
            -QX * LA ** 2,
            -QY * LA ** 2,
            inv * QX ** 2 - inv * QY ** 2,
            inv * QX * QY,
        ],
        "relaxation_parameters": [0, 0, 0, s_e, s_eps, s_q, s_q, s_nu, s_nu],
    }


def d3q19_mrt(s_nu, s_e=1.19, s_eps=1.4, s_q=1.2, s_pi=1.4, s_m=1.98, source_qx=None):
    """D3Q19 orthogonal MRT (d'Humieres, Ginzburg, Krafczyk, Lallemand, Luo 2002), rho0 = 1."""
    r2 = X ** 2 + Y ** 2 + Z ** 2
    j2 = QX ** 2 + QY ** 2 + QZ ** 2
    half = sp.Rational(1, 2)
    entry = {
        "velocities": list(range(19)),
        "conserved_moments": [RHO, QX, QY, QZ],
        "polynomials": [
            1,
            19 * r2 - 30,
            half * (21 * r2 ** 2 - 53 * r2 + 24),
            X, (5 * r2 - 9) * X,
            Y, (5 * r2 - 9) * Y,
            Z, (5 * r2 - 9) * Z,
            3 * X ** 2 - r2, (3 * r2 - 5) * (3 * X ** 2 - r2),
            Y ** 2 - Z ** 2, (3 * r2 - 5) * (Y ** 2 - Z ** 2),
            X * Y, Y * Z, X * Z,
            (Y ** 2 - Z ** 2) * X, (Z ** 2 - X ** 2) * Y, (X ** 2 - Y ** 2) * Z,
        ],
        "equilibrium": [
            RHO,
            -11 * RHO + 19 * j2,
            3 * RHO - sp.Rational(11, 2) * j2,
            QX, -sp.Rational(2, 3) * QX,
            QY, -sp.Rational(2, 3) * QY,
            QZ, -sp.Rational(2, 3) * QZ,
            3 * QX ** 2 - j2, -half * (3 * QX ** 2 - j2),
            QY ** 2 - QZ ** 2, -half * (QY ** 2 - QZ ** 2),
            QX * QY, QY * QZ, QX * QZ,
            0, 0, 0,
        ],
        "relaxation_parameters": [0, s_e, s_eps, 0, s_q, 0, s_q, 0, s_q,
                                  s_nu, s_pi, s_nu, s_pi, s_nu, s_nu, s_nu, s_m, s_m, s_m],
    }
    if source_qx is not None:
        entry["source_terms"] = {QX: source_qx}
    return entry


def _s_nu(nu):
    return 1.0 / (0.5 + 3.0 * nu)


def lid_cavity_d2q9(n=128, u_lid=0.05, reynolds=100.0, generator="hip"):
    """Lid-driven cavity, n x n cells, nu = u_lid * n / Re, s = [0,0,0,1.5,1.5,1.5,1.5,s_nu,s_nu]."""
    nu = u_lid * n / reynolds
    s_nu = _s_nu(nu)

    def lid(f, m, x, y):
        m[RHO] = 1.0
        m[QX] = u_lid
        m[QY] = 0.0

    return {
        "parameters": {LA: 1},
        "box": {"x": [0.0, float(n)], "y": [0.0, float(n)], "label": [0, 0, 0, 1]},
        "space_step": 1.0,
        "scheme_velocity": LA,
        "schemes": [d2q9_mrt(1.5, 1.5, 1.5, s_nu)],
        "init": {RHO: 1.0, QX: 0.0, QY: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BounceBack}},
            1: {"method": {0: bc.BounceBack}, "value": lid},
        },
        "generator": generator,
    }


def karman_d2q9(nx=1024, ny=256, radius=20.0, center=(256.0, 128.0), u_in=0.05, reynolds=100.0,
                amplitude=1e-3, seed=1234, generator="hip"):
    """Channel with a cylinder: ABB inflow (left, label 1), NeumannX outflow (right, label 2),
    bounce-back on the top/bottom walls and the cylinder (label 0).
    nu = u_in * 2 r / Re.  Initial velocity (u_in, 0) + U(-amplitude, amplitude) per component."""
    nu = u_in * 2.0 * radius / reynolds
    s_nu = _s_nu(nu)

    def inflow(f, m, x, y):
        m[RHO] = 1.0
        m[QX] = u_in
        m[QY] = 0.0

    def qx0(x, y):
        rng = np.random.default_rng(seed)
        return u_in + rng.uniform(-amplitude, amplitude, size=(x.size, y.size))

    def qy0(x, y):
        rng = np.random.default_rng(seed + 1)
        return rng.uniform(-amplitude, amplitude, size=(x.size, y.size))

    return {
        "parameters": {LA: 1},
        "box": {"x": [0.0, float(nx)], "y": [0.0, float(ny)], "label": [1, 2, 0, 0]},
        "elements": [Circle(list(center), radius, label=0)],
        "space_step": 1.0,
        "scheme_velocity": LA,
        "schemes": [d2q9_mrt(1.5, 1.5, 1.5, s_nu)],
        "init": {RHO: 1.0, QX: qx0, QY: qy0},
        "boundary_conditions": {
            0: {"method": {0: bc.BounceBack}},
            1: {"method": {0: bc.AntiBounceBack}, "value": inflow},
            2: {"method": {0: bc.NeumannX}},
        },
        "generator": generator,
    }


def poiseuille_d2q9(nx=4, ny=12, y_low=1.1, y_high=10.6, g=1e-5, s_nu=1.0, method=bc.BounceBack,
                    generator="hip"):
    """Force-driven channel, periodic in x, between two solid slabs whose faces
    y = y_low and y = y_high cut the links at fractional distances (label 0, method
    `method`); body force g on qx.  Steady state: a parabola vanishing at the walls."""
    sch = d2q9_mrt(1.2, 1.2, 1.2, s_nu)
    sch["source_terms"] = {QX: g}
    H = float(ny)
    return {
        "parameters": {LA: 1},
        "box": {"x": [0.0, float(nx)], "y": [0.0, H], "label": [-1, -1, 0, 0]},
        "elements": [Parallelogram((-2.0, -1.0), (nx + 4.0, 0.0), (0.0, y_low + 1.0), label=0),
                     Parallelogram((-2.0, y_high), (nx + 4.0, 0.0), (0.0, H - y_high + 1.0), label=0)],
        "space_step": 1.0,
        "scheme_velocity": LA,
        "schemes": [sch],
        "init": {RHO: 1.0, QX: 0.0, QY: 0.0},
        "boundary_conditions": {0: {"method": {0: method}}},
        "generator": generator,
    }


def with_bouzidi(d):
    """The same case with its bounce-back / anti-bounce-back methods replaced by the
    interpolated (Bouzidi) ones."""
    swap = {bc.BounceBack: bc.BouzidiBounceBack, bc.AntiBounceBack: bc.BouzidiAntiBounceBack}
    out = dict(d)
    out["boundary_conditions"] = {
        label: dict(spec, method={k: swap.get(cls, cls) for k, cls in spec["method"].items()})
        for label, spec in d["boundary_conditions"].items()}
    return out


def taylor_green_d3q19(n=256, u0=0.04, nu=0.005, generator="hip"):
    """Periodic Taylor-Green vortex on n^3 cells, wave number 2 pi / n, rho0 = 1."""
    k = 2.0 * math.pi / n

    def rho(x, y, z):
        return 1.0 + 3.0 * u0 ** 2 / 16.0 * (np.cos(2 * k * x) + np.cos(2 * k * y)) * (np.cos(2 * k * z) + 2.0)

    def qx(x, y, z):
        return u0 * np.sin(k * x) * np.cos(k * y) * np.cos(k * z)

    def qy(x, y, z):
        return -u0 * np.cos(k * x) * np.sin(k * y) * np.cos(k * z)

    return {
        "box": {"x": [0.0, float(n)], "y": [0.0, float(n)], "z": [0.0, float(n)], "label": -1},
        "space_step": 1.0,
        "scheme_velocity": 1,
        "schemes": [d3q19_mrt(_s_nu(nu))],
        "init": {RHO: rho, QX: qx, QY: qy, QZ: 0.0},
        "generator": generator,
    }


def forced_periodic_d2q9(nx=12, ny=7, dx=0.5, g0=1e-3, seed=11, generator="hip"):
    """Periodic D2Q9 MRT box driven by a body force that depends on space and time,
    S(qx) = rho g0 (1 + Y) (1 + t), S(qy) = rho g0 X / 4 (X, Y cell centres, t the
    time of the step), from a random density field."""
    rng = np.random.default_rng(seed)
    drho = rng.uniform(-0.05, 0.05, size=(nx + 2, ny + 2))

    def rho(x, y):
        return 1.0 + drho

    t = sp.Symbol("t")
    entry = d2q9_mrt(1.5, 1.4, 1.3, 1.7)
    entry["source_terms"] = {QX: RHO * g0 * (1 + Y) * (1 + t), QY: RHO * g0 * X / 4}
    return {
        "parameters": {LA: 1},
        "box": {"x": [0.0, nx * dx], "y": [0.0, ny * dx], "label": -1},
        "space_step": dx,
        "scheme_velocity": LA,
        "schemes": [entry],
        "init": {RHO: rho, QX: 0.0, QY: 0.0},
        "generator": generator,
    }


def sphere_centres(shape, rmin=4.0, rmax=12.0, solid_fraction=0.10, seed=2024):
    """Random sphere packing: centres uniform in the box, radii uniform in
    [rmin, rmax], added (overlaps allowed) until the solid cells -- cells of the
    box whose centre lies inside a sphere, the domain's own rule -- make up
    solid_fraction of the box (dx = 1, cell i centred at i + 1/2)."""
    rng = np.random.default_rng(seed)
    solid = np.zeros(tuple(int(n) for n in shape), dtype=bool)
    target = solid_fraction * solid.size
    count = 0
    out = []
    while count < target:
        r = float(rng.uniform(rmin, rmax))
        c = [float(rng.uniform(0.0, s)) for s in shape]
        out.append((c, r))
        lo = [max(0, int(math.floor(c[a] - r - 0.5))) for a in range(len(shape))]
        hi = [min(int(shape[a]), int(math.ceil(c[a] + r + 0.5)) + 1) for a in range(len(shape))]
        axes = np.meshgrid(*[np.arange(lo[a], hi[a]) + 0.5 - c[a] for a in range(len(shape))],
                           sparse=True, indexing="ij")
        inside = sum(x * x for x in axes) <= r * r
        box = solid[tuple(slice(lo[a], hi[a]) for a in range(len(shape)))]
        count += int(np.count_nonzero(inside & ~box))
        box |= inside
    return out


def sphere_packing_d3q19(nx=512, ny=256, nz=256, g=1e-6, nu=1.0 / 6.0, seed=2024,
                         solid_fraction=0.10, generator="hip"):
    """Body-force-driven flow through random spheres: periodic in x (labels -1),
    bounce-back walls in y and z (label 0) and on the spheres (label 1);
    source term S(qx) = rho g applied as two dt/2 half steps."""
    spheres = sphere_centres((nx, ny, nz), seed=seed, solid_fraction=solid_fraction)
    return {
        "box": {"x": [0.0, float(nx)], "y": [0.0, float(ny)], "z": [0.0, float(nz)],
                "label": [-1, -1, 0, 0, 0, 0]},
        "elements": [Sphere(c, r, label=1) for c, r in spheres],
        "space_step": 1.0,
        "scheme_velocity": 1,
        "schemes": [d3q19_mrt(_s_nu(nu), source_qx=RHO * g)],
        "init": {RHO: 1.0, QX: 0.0, QY: 0.0, QZ: 0.0},
        "boundary_conditions": {
            0: {"method": {0: bc.BounceBack}},
            1: {"method": {0: bc.BounceBack}},
        },
        "generator": generator,
    }


def fourier_velocity(n, kmin=4.0, kmax=16.0, rms=0.05, seed=0):
    """Divergence-free random periodic velocity with E(k) ~ k^(-5/3) on kmin <= |k| <= kmax."""
    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, d=1.0 / n)
    kx, ky = np.meshgrid(k1, k1, indexing="ij")
    kk = np.sqrt(kx ** 2 + ky ** 2)
    band = (kk >= kmin) & (kk <= kmax)
    amp = np.zeros_like(kk)
    # |u_hat|^2 ~ E(k) / (2 pi k): psi_hat = u_hat / k
    amp[band] = np.sqrt(kk[band] ** (-5.0 / 3.0) / (2.0 * np.pi * kk[band])) / kk[band]
    phase = np.exp(2j * np.pi * rng.random(kk.shape))
    psi = np.real(np.fft.ifft2(amp * phase))
    ux = (np.roll(psi, -1, axis=1) - np.roll(psi, 1, axis=1)) / 2.0
    uy = -(np.roll(psi, -1, axis=0) - np.roll(psi, 1, axis=0)) / 2.0
    scale = rms / np.sqrt(np.mean(ux ** 2 + uy ** 2))
    return ux * scale, uy * scale


def decaying_d2q9(n=512, nu=1e-3, rms=0.05, seed=0, generator="hip"):
    """Periodic decaying flow on n x n cells from :func:`fourier_velocity`."""
    ux, uy = fourier_velocity(n, rms=rms, seed=seed)

    def halo(a):
        return np.pad(a, 1, mode="wrap")

    def qx(x, y):
        return halo(ux)

    def qy(x, y):
        return halo(uy)

    return {
        "parameters": {LA: 1},
        "box": {"x": [0.0, float(n)], "y": [0.0, float(n)], "label": -1},
        "space_step": 1.0,
        "scheme_velocity": LA,
        "schemes": [d2q9_mrt(1.5, 1.5, 1.5, _s_nu(nu))],
        "init": {RHO: 1.0, QX: qx, QY: qy},
        "generator": generator,
    }


# ------------------------------------------------------------ small test schemes
U, T = sp.symbols("u, T")


def d1q3_advection(n=64, c=0.2, walls=False, generator="hip"):
    """D1Q3 advection-diffusion of u on [0, n]: periodic, or bounce-back left
    (label 0, value u = 1) and Neumann right (label 1) when ``walls``."""

    def left(f, m, x):
        m[U] = 1.0

    def u0(x):
        return np.exp(-((x - 0.5 * n) / (0.1 * n)) ** 2)

    d = {
        "box": {"x": [0.0, float(n)], "label": [0, 1] if walls else -1},
        "space_step": 1.0,
        "scheme_velocity": 1,
        "schemes": [{
            "velocities": [0, 1, 2],
            "conserved_moments": U,
            "polynomials": [1, X, X ** 2 / 2],
            "equilibrium": [U, c * U, sp.Rational(1, 3) * U],
            "relaxation_parameters": [0, 1.6, 1.3],
        }],
        "init": {U: u0},
        "generator": generator,
    }
    if walls:
        d["boundary_conditions"] = {0: {"method": {0: bc.BounceBack}, "value": left},
                                    1: {"method": {0: bc.Neumann}}}
    return d


def d2q5q9_vectorial(n=32, generator="hip", walls=False):
    """Vectorial scheme: D2Q9 MRT flow (rho, qx, qy) and a D2Q5 passive scalar T
    advected by it (block-diagonal M over the two schemes, pylbm/scheme.py:509-516)."""
    k = 2 * math.pi / n

    def qx(x, y):
        return 0.02 * np.sin(k * y) + 0.0 * x

    def t0(x, y):
        return 1.0 + 0.1 * np.cos(k * x) * np.cos(k * y)

    flow = d2q9_mrt(1.2, 1.3, 1.4, 1.7)
    scal = {
        "velocities": list(range(5)),
        "conserved_moments": T,
        "polynomials": [1, X, Y, X ** 2 + Y ** 2, X ** 2 - Y ** 2],
        "equilibrium": [T, QX * T, QY * T, sp.Rational(1, 2) * T, 0],
        "relaxation_parameters": [0, 1.5, 1.5, 1.2, 1.2],
    }
    d = {
        "parameters": {LA: 1},
        "box": {"x": [0.0, float(n)], "y": [0.0, float(n)], "label": [0, 1, 0, 0] if walls else -1},
        "space_step": 1.0,
        "scheme_velocity": LA,
        "schemes": [flow, scal],
        "init": {RHO: 1.0, QX: qx, QY: 0.0, T: t0},
        "generator": generator,
    }
    if walls:
        d["boundary_conditions"] = {0: {"method": {0: bc.BounceBack, 1: bc.BounceBack}},
                                    1: {"method": {0: bc.NeumannX, 1: bc.Neumann}}}
    return d
