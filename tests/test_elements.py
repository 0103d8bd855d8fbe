"""Geometric elements beyond the circle / sphere / polygons: Ellipse, Ellipsoid, the
cylinders (circle, ellipse, triangle bases) and the parallelepiped
(pylbm/elements/ellipse.py, ellipsoid.py, cylinder.py, base.py).

Pins: the bounds of pylbm's own tests/test_elements.py; the ellipse / ellipsoid
spanned by r e_i give the circle's / sphere's link distances bit for bit (the same
quadric formula); for the 3D extrusions every link distance equals the first entry
of the segment into the element found by bisection on point_inside, and the label is
that of the face crossed (side, bottom, top)."""
import numpy as np
import pytest

import pylbm_amd as pa
from pylbm_amd import Simulation  # noqa: F401  (the package imports without a GPU)
from pylbm_amd.domain import Domain


@pytest.mark.parametrize("dim, elem", [
    (2, lambda: pa.Circle([0, 0], 1)),
    (2, lambda: pa.Ellipse([0, 0], [1, 0], [0, 1])),
    (2, lambda: pa.Triangle([-1, -1], [0, 2], [2, 0])),
    (2, lambda: pa.Parallelogram([-1, -1], [0, 2], [2, 0])),
    (3, lambda: pa.Ellipsoid([0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1])),
    (3, lambda: pa.Sphere([0, 0, 0], 1)),
], ids=["Circle", "Ellipse", "Triangle", "Parallelogram", "Ellipsoid", "Sphere"])
def test_bounds_as_reference(dim, elem):
    """pylbm tests/test_elements.py: every element of its list spans [-1, 1]^dim."""
    lo, hi = elem().get_bounds()
    assert np.allclose(lo, [-1] * dim) and np.allclose(hi, [1] * dim)


def _grid2(n=41, off=0.013):
    x = np.linspace(-2, 2, n) + off
    return np.meshgrid(x, x + 0.7 * off, sparse=True, indexing="ij")


def _grid3(n=17, off=0.011):
    x = np.linspace(-2, 2, n) + off
    return np.meshgrid(x, x + 0.7 * off, x - 0.3 * off, sparse=True, indexing="ij")


def test_ellipse_of_equal_axes_is_the_circle():
    r, c = 0.83, [0.1, -0.2]
    circ, ell = pa.Circle(c, r, label=4), pa.Ellipse(c, [r, 0], [0, r], label=4)
    grid = _grid2()
    for v in ([0.1, 0.0], [0.1, 0.1], [-0.1, 0.05], [0.0, -0.1]):
        a1, b1 = circ.distance(grid, np.asarray(v), 1.0)
        a2, b2 = ell.distance(grid, np.asarray(v), 1.0)
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert np.array_equal(circ.point_inside(grid), ell.point_inside(grid))


def test_ellipsoid_of_equal_axes_is_the_sphere():
    r, c = 0.77, [0.1, -0.2, 0.05]
    sph, ell = pa.Sphere(c, r, label=2), pa.Ellipsoid(c, [r, 0, 0], [0, r, 0], [0, 0, r], label=2)
    grid = _grid3()
    for v in ([0.2, 0.0, 0.0], [0.2, 0.2, 0.0], [-0.2, 0.1, 0.2], [0.0, 0.0, -0.2]):
        a1, b1 = sph.distance(grid, np.asarray(v), 1.0)
        a2, b2 = ell.distance(grid, np.asarray(v), 1.0)
        assert np.array_equal(a1, a2) and np.array_equal(b1, b2)
    assert np.array_equal(sph.point_inside(grid), ell.point_inside(grid))


def test_ellipse_rotated_axes():
    """A rotated ellipse: inside iff (p.u1)^2/|v1|^4 + (p.u2)^2/|v2|^4 <= 1 (u_i = v_i)."""
    v1, v2 = np.array([0.6, 0.3]), np.array([-0.15, 0.3])
    e = pa.Ellipse([0, 0], v1, v2)
    x, y = _grid2(61)
    X, Y = np.broadcast_arrays(x, y)
    q = (X * v1[0] + Y * v1[1]) ** 2 / np.dot(v1, v1) ** 2 + (X * v2[0] + Y * v2[1]) ** 2 / np.dot(v2, v2) ** 2
    inside = e.point_inside((x, y))
    assert np.array_equal(inside[np.abs(q - 1) > 1e-9], (q <= 1)[np.abs(q - 1) > 1e-9])
    with pytest.raises(ValueError):
        pa.Ellipse([0, 0], [1, 0], [1, 1])


def _first_entry(elem, p, v):
    """First t in (0, 1] where p + t v is inside elem, by sampling then bisection;
    None when the segment stays outside (or enters within 1e-3 of its ends)."""
    ts = np.linspace(0.0, 1.0, 2001)
    pts = [p[d] + ts * v[d] for d in range(3)]
    ins = elem.point_inside(tuple(pts))
    if ins[0] or not ins.any():
        return None
    k = int(np.argmax(ins))
    lo, hi = ts[k - 1], ts[k]
    if lo < 1e-3 or hi > 1 - 1e-3:
        return None
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        if elem.point_inside(tuple(np.array([p[d] + mid * v[d]]) for d in range(3)))[0]:
            hi = mid
        else:
            lo = mid
    return 0.5 * (lo + hi)


EXTRUSIONS = {
    "cyl_circle": lambda: pa.CylinderCircle([0.05, -0.1, 0.02], [0.7, 0, 0], [0, 0.7, 0], [0, 0, 0.6],
                                            label=[1, 2, 3]),
    "cyl_ellipse_tilted": lambda: pa.CylinderEllipse([0.03, 0.02, -0.04], [0, 0.5, 0.5], [0, -0.6, 0.6],
                                                     [0.5, -0.2, 0.2], label=[1, 2, 3]),
    "cyl_triangle": lambda: pa.CylinderTriangle([-0.5, -0.45, 0.01], [1.1, 0, 0], [0, 1.0, 0], [0, 0, 0.7],
                                                label=[1, 2, 3, 4, 5]),
    "parallelepiped": lambda: pa.Parallelepiped([-0.6, -0.55, -0.5], [1.2, 0, 0], [0, 1.1, 0.1], [0, 0, 1.0],
                                                label=[1, 2, 3, 4, 5, 6]),
}


@pytest.mark.parametrize("name", sorted(EXTRUSIONS))
def test_extrusion_distances_by_bisection(name):
    elem = EXTRUSIONS[name]()
    rng = np.random.default_rng(7)
    checked = 0
    for v in [np.array(c) * 0.45 for c in ((1, 0, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (1, 1, 0), (-1, 0, 1),
                                              (0, 1, -1), (1, -1, 1))]:
        for _ in range(200):
            p = rng.uniform(-1.2, 1.2, size=3)
            grid = tuple(np.array([[[p[d]]]]) for d in range(3))
            if elem.point_inside(grid).ravel()[0]:
                continue
            alpha, border = elem.distance(grid, v, 1.0)
            a = float(alpha.ravel()[0])
            t = _first_entry(elem, p, v)
            if t is None:
                continue
            assert a > 0 and abs(a - t) < 1e-9, (name, p, v, a, t)
            assert int(border.ravel()[0]) in elem.label
            checked += 1
    assert checked > 50


def test_cylinder_face_labels():
    """z-aligned circular cylinder |z| <= 0.6: links crossing the top plane take the
    last label, the bottom plane the second last, the side the first."""
    cyl = EXTRUSIONS["cyl_circle"]()
    c = cyl.center

    def hit(p, v):
        grid = tuple(np.array([[[p[d]]]]) for d in range(3))
        a, b = cyl.distance(grid, np.asarray(v, dtype=float), 1.0)
        return float(a.ravel()[0]), int(b.ravel()[0])

    a, b = hit([c[0] + 0.1, c[1], c[2] + 0.9], [0, 0, -0.5])
    assert b == 3 and abs(a - 0.6) < 1e-12
    a, b = hit([c[0], c[1] + 0.2, c[2] - 0.8], [0, 0, 0.4])
    assert b == 2 and abs(a - 0.5) < 1e-12
    a, b = hit([c[0] + 0.9, c[1], c[2]], [-0.4, 0, 0])
    assert b == 1 and abs(a - 0.5) < 1e-12


def test_domain_with_reference_cylinder_and_ellipsoid():
    """The domains of pylbm's tests/test_domain.py with a CylinderEllipse and an
    Ellipsoid build, carry the elements' labels, and make solid exactly the cell
    centres inside the element (within its bounding box)."""
    d1 = {"box": {"x": [-3, 3], "y": [-3, 3], "z": [-3, 3], "label": 0},
          "elements": [pa.CylinderEllipse([0.5, 0, 0], [0, 1, 1], [0, -1.5, 1.5], [1, -1, 0], label=[1, 2, 3])],
          "space_step": 0.5, "schemes": [{"velocities": list(range(19))}]}
    d2 = {"box": {"x": [0, 3], "y": [0, 3], "z": [0, 3], "label": list(range(1, 7))},
          "elements": [pa.Ellipsoid((1.5, 1.5, 1.5), [0.5, 0, 0], [0, 0.5, 0], [0, 0, 1], label=0)],
          "space_step": 0.5, "schemes": [{"velocities": list(range(19))}]}
    for d, labels in ((d1, {0, 1, 2, 3}), (d2, {0, 1, 2, 3, 4, 5, 6})):
        dom = Domain(d)
        assert labels <= set(dom.list_of_labels())
        elem = d["elements"][0]
        inner = tuple(slice(h, -h) for h in dom.stencil.vmax)
        grid = np.meshgrid(*dom.coords, sparse=True, indexing="ij")
        inside = np.broadcast_to(elem.point_inside(grid), dom.fluid[inner].shape)
        lo, hi = elem.get_bounds()
        X, Y, Z = np.broadcast_arrays(*grid)
        box = np.ones(X.shape, dtype=bool)
        for a, g in enumerate((X, Y, Z)):
            box &= (g >= lo[a] - 2 * dom.dx) & (g <= hi[a] + 2 * dom.dx)
        assert np.array_equal(~dom.fluid[inner][box], inside[box])
        assert inside.any()
