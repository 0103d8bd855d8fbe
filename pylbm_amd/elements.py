"""Geometric elements that carve solids (or fluid) out of the box.

``Circle`` (pylbm/elements/circle.py:24-171), ``Sphere``
(pylbm/elements/sphere.py:24-176), ``Parallelogram``
(pylbm/elements/parallelogram.py) and ``Triangle``
(pylbm/elements/triangle.py).  ``point_inside`` decides the solid mask and
``distance`` returns, for every grid point, the fraction of the link dx*v
after which the element's border is crossed (-1 when it is not crossed
within dmax), and the label of that border.

The floating-point operations of ``distance`` are performed in the same
order as pylbm/elements/utils.py:21-277 does them, because the link lists
must be bit-identical to the reference's: a link whose distance differs in
the last bit can change which label (and so which boundary method) owns it.
"""
import numpy as np

__all__ = ["Circle", "Sphere", "Parallelogram", "Triangle"]

_FAR = 1.0e16


def _labels(label, n):
    return [label] * n if isinstance(label, (int, np.integer)) else list(label)


def _nearest_positive_root(a, b, c, shape):
    """Smallest positive root of a d^2 + b d + c = 0, -1 when none (utils.py:93-167)."""
    delta = b ** 2 - 4 * a * c
    real = delta >= 0
    delta[real] = np.sqrt(delta[real])
    r1 = np.full(shape, _FAR)
    r2 = np.full(shape, _FAR)
    if a != 0:
        r1[real] = (-b[real] - delta[real]) / (2 * a)
        r2[real] = (-b[real] + delta[real]) / (2 * a)
    r1[r1 < 0] = _FAR
    r2[r2 < 0] = _FAR
    d = np.full(shape, -1.0)
    d[real] = np.minimum(r1[real], r2[real])
    d[d == _FAR] = -1
    return d


def _clip(d, dmax, label, shape):
    alpha = np.full(shape, -1.0)
    border = np.full(shape, -1.0)
    hit = (d > 0) if dmax is None else np.logical_and(d > 0, d <= dmax)
    alpha[hit] = d[hit]
    border[hit] = label
    return alpha, border


class Circle:
    number_of_bounds = 1
    dim = 2

    def __init__(self, center, radius, label=0, isfluid=False):
        if radius < 0:
            raise ValueError("the radius of the circle should be positive")
        self.center = np.asarray(center)
        self.radius = radius
        self.label = _labels(label, 1)
        self.isfluid = isfluid

    def get_bounds(self):
        return self.center - self.radius, self.center + self.radius

    def point_inside(self, grid):
        x, y = grid
        return ((x - self.center[0]) ** 2 + (y - self.center[1]) ** 2) <= self.radius ** 2

    def distance(self, grid, v, dmax=None, normal=False):
        x, y = grid
        e1 = self.radius * np.array([1, 0])
        e2 = self.radius * np.array([0, 1])
        px = x - self.center[0]
        py = y - self.center[1]
        cx2 = e1[0] ** 2 + e2[0] ** 2
        cy2 = e1[1] ** 2 + e2[1] ** 2
        cxy = 2 * (e1[0] * e1[1] + e2[0] * e2[1])
        rhs = (e1[0] * e2[1] - e1[1] * e2[0]) ** 2
        a = v[0] ** 2 * cy2 + v[1] ** 2 * cx2 - v[0] * v[1] * cxy
        b = 2 * px * v[0] * cy2 + 2 * py * v[1] * cx2 - (px * v[1] + py * v[0]) * cxy
        c = px ** 2 * cy2 + py ** 2 * cx2 - px * py * cxy - rhs
        shape = np.broadcast(b, c).shape
        d = _nearest_positive_root(a, np.broadcast_to(b, shape).copy(), np.broadcast_to(c, shape), shape)
        return _clip(d, dmax, self.label[0], shape)


class Sphere:
    number_of_bounds = 1
    dim = 3

    def __init__(self, center, radius, label=0, isfluid=False):
        if radius < 0:
            raise ValueError("the radius of the sphere should be positive")
        self.center = np.asarray(center)
        self.radius = radius
        self.label = _labels(label, 1)
        self.isfluid = isfluid

    def get_bounds(self):
        return self.center - self.radius, self.center + self.radius

    def point_inside(self, grid):
        x, y, z = grid
        return ((x - self.center[0]) ** 2 + (y - self.center[1]) ** 2
                + (z - self.center[2]) ** 2) <= self.radius ** 2

    def quadric(self):
        """Coefficients of the sphere as the ellipsoid of utils.py:171-234."""
        e1 = self.radius * np.array([1, 0, 0])
        e2 = self.radius * np.array([0, 1, 0])
        e3 = self.radius * np.array([0, 0, 1])
        c12, c23, c31 = np.cross(e1, e2), np.cross(e2, e3), np.cross(e3, e1)
        rhs = np.inner(e1, c23) ** 2
        cxx = c12[0] ** 2 + c23[0] ** 2 + c31[0] ** 2
        cyy = c12[1] ** 2 + c23[1] ** 2 + c31[1] ** 2
        czz = c12[2] ** 2 + c23[2] ** 2 + c31[2] ** 2
        cxy = 2 * (c12[0] * c12[1] + c23[0] * c23[1] + c31[0] * c31[1])
        cyz = 2 * (c12[1] * c12[2] + c23[1] * c23[2] + c31[1] * c31[2])
        czx = 2 * (c12[2] * c12[0] + c23[2] * c23[0] + c31[2] * c31[0])
        return rhs, cxx, cyy, czz, cxy, cyz, czx

    def distance(self, grid, v, dmax=None, normal=False):
        x, y, z = grid
        shape = (x.size, y.size, z.size)
        px, py, pz = x - self.center[0], y - self.center[1], z - self.center[2]
        rhs, cxx, cyy, czz, cxy, cyz, czx = self.quadric()
        a = (cxx * v[0] ** 2 + cyy * v[1] ** 2 + czz * v[2] ** 2
             + cxy * v[0] * v[1] + cyz * v[1] * v[2] + czx * v[2] * v[0])
        b = ((2 * cxx * v[0] + cxy * v[1] + czx * v[2]) * px
             + (2 * cyy * v[1] + cyz * v[2] + cxy * v[0]) * py
             + (2 * czz * v[2] + czx * v[0] + cyz * v[1]) * pz)
        c = (cxx * px ** 2 + cyy * py ** 2 + czz * pz ** 2
             + cxy * px * py + cyz * py * pz + czx * pz * px - rhs)
        delta = b ** 2 - 4 * a * c
        real = delta >= 0
        delta[real] = np.sqrt(delta[real])
        r1 = np.full(shape, _FAR)
        r2 = np.full(shape, _FAR)
        r1[real] = (-b[real] - delta[real]) / (2 * a)
        r2[real] = (-b[real] + delta[real]) / (2 * a)
        r1[r1 < 0] = _FAR
        r2[r2 < 0] = _FAR
        d = np.full(shape, -1.0)
        d[real] = np.minimum(r1[real], r2[real])
        d[d == _FAR] = -1
        return _clip(d, dmax, self.label[0], shape)


def _line_hits(px, py, v, starts, edges, dmax, labels):
    """Nearest crossing of the segment (p, p + v) with polygon edges (utils.py:21-90)."""
    if px.shape[1] == 1 and py.shape[0] == 1:
        shape = (px.size, py.size)
    else:
        shape = px.shape
    alpha = np.full(shape, _FAR)
    border = np.full(shape, -1.0)
    for q, e, lab in zip(starts, edges, labels):
        det = v[1] * e[0] - v[0] * e[1]
        if det == 0:
            continue
        inv = 1.0 / det
        c1 = q[0] - px
        c2 = q[1] - py
        t = (-e[1] * c1 + e[0] * c2) * inv
        u = (-v[1] * c1 + v[0] * c2) * inv
        if dmax is None:
            ok = np.logical_and(t > 0, np.logical_and(u >= 0, u <= 1))
        else:
            ok = np.logical_and(np.logical_and(t > 0, t <= dmax), np.logical_and(u >= 0, u <= 1))
        ok = np.where(np.logical_and(alpha > t, ok))
        alpha[ok] = t[ok]
        border[ok] = lab
    alpha[alpha == _FAR] = -1.0
    return alpha, border


class Parallelogram:
    number_of_bounds = 4
    dim = 2

    def __init__(self, point, vecta, vectb, label=0, isfluid=False):
        self.point = np.asarray(point)
        self.v1 = np.asarray(vecta)
        self.v2 = np.asarray(vectb)
        self.label = _labels(label, self.number_of_bounds)
        self.isfluid = isfluid

    def _corners(self):
        return [self.point, self.point + self.v1, self.point + self.v1 + self.v2, self.point + self.v2]

    def get_bounds(self):
        box = np.asarray(self._corners())
        return np.min(box, axis=0), np.max(box, axis=0)

    def _barycentric(self, grid):
        x, y = grid
        wx, wy = x - self.point[0], y - self.point[1]
        inv = 1.0 / (self.v1[0] * self.v2[1] - self.v1[1] * self.v2[0])
        return (wx * self.v2[1] - wy * self.v2[0]) * inv, (wy * self.v1[0] - wx * self.v1[1]) * inv

    def point_inside(self, grid):
        u, v = self._barycentric(grid)
        return np.logical_and(np.logical_and(u >= 0, v >= 0), np.logical_and(u <= 1, v <= 1))

    def distance(self, grid, v, dmax=None, normal=False):
        x, y = grid
        return _line_hits(x - self.point[0], y - self.point[1], v,
                          [[0, 0], [0, 0], self.v2, self.v1], [self.v1, self.v2, self.v1, self.v2],
                          dmax, self.label)


class Triangle(Parallelogram):
    number_of_bounds = 3

    def _corners(self):
        return [self.point, self.point + self.v1, self.point + self.v2]

    def point_inside(self, grid):
        u, v = self._barycentric(grid)
        return np.logical_and(np.logical_and(u >= 0, v >= 0), u + v <= 1)

    def distance(self, grid, v, dmax=None, normal=False):
        x, y = grid
        return _line_hits(x - self.point[0], y - self.point[1], v,
                          [[0, 0], [0, 0], self.v1], [self.v1, self.v2, self.v2 - self.v1],
                          dmax, self.label)
