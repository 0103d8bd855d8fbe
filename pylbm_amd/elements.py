"""Geometric elements that carve solids (or fluid) out of the box.

``Circle`` (pylbm/elements/circle.py:24-171), ``Ellipse``
(pylbm/elements/ellipse.py:22-196), ``Sphere`` (pylbm/elements/sphere.py:24-176),
``Ellipsoid`` (pylbm/elements/ellipsoid.py:21-211), ``Parallelogram``
(pylbm/elements/parallelogram.py), ``Triangle`` (pylbm/elements/triangle.py), and
the 3D extrusions ``CylinderCircle``, ``CylinderEllipse``, ``CylinderTriangle``
and ``Parallelepiped`` (pylbm/elements/cylinder.py:27-566 with their unit bases,
pylbm/elements/base.py:136-608).  ``point_inside`` decides the solid mask and
``distance`` returns, for every grid point, the fraction of the link dx*v
after which the element's border is crossed (-1 when it is not crossed
within dmax), and the label of that border.

The floating-point operations of ``distance`` are performed in the same
order as pylbm/elements/utils.py:21-277 does them, because the link lists
must be bit-identical to the reference's: a link whose distance differs in
the last bit can change which label (and so which boundary method) owns it.
"""
import numpy as np

__all__ = ["Circle", "Ellipse", "Sphere", "Ellipsoid", "Parallelogram", "Triangle",
           "CylinderCircle", "CylinderEllipse", "CylinderTriangle", "Parallelepiped"]

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
        return _ellipse_distance(x - self.center[0], y - self.center[1], v, self.radius * np.array([1, 0]),
                                 self.radius * np.array([0, 1]), dmax, self.label[0])


def _ellipse_quadric(e1, e2):
    """X^2 cy2 + Y^2 cx2 - X Y cxy = rhs: the ellipse spanned by e1, e2 about its centre."""
    cx2 = e1[0] ** 2 + e2[0] ** 2
    cy2 = e1[1] ** 2 + e2[1] ** 2
    cxy = 2 * (e1[0] * e1[1] + e2[0] * e2[1])
    return cx2, cy2, cxy, (e1[0] * e2[1] - e1[1] * e2[0]) ** 2


def _ellipse_distance(px, py, v, e1, e2, dmax, label):
    """Fraction of v after which the segment from (px, py) (relative to the centre)
    enters or leaves the ellipse spanned by e1, e2 (utils.py:93-167)."""
    cx2, cy2, cxy, rhs = _ellipse_quadric(e1, e2)
    a = v[0] ** 2 * cy2 + v[1] ** 2 * cx2 - v[0] * v[1] * cxy
    b = 2 * px * v[0] * cy2 + 2 * py * v[1] * cx2 - (px * v[1] + py * v[0]) * cxy
    c = px ** 2 * cy2 + py ** 2 * cx2 - px * py * cxy - rhs
    shape = np.broadcast(b, c).shape
    d = _nearest_positive_root(a, np.broadcast_to(b, shape).copy(), np.broadcast_to(c, shape), shape)
    return _clip(d, dmax, label, shape)


class Ellipse:
    """The solid (or fluid) ellipse of centre c spanned by orthogonal v1, v2."""
    number_of_bounds = 1
    dim = 2

    def __init__(self, center, v1, v2, label=0, isfluid=False):
        if abs(v1[0] * v2[0] + v1[1] * v2[1]) > 1.0e-14:
            raise ValueError("the vectors of the ellipse are not orthogonal")
        self.center = np.asarray(center)
        self.v1 = np.asarray(v1)
        self.v2 = np.asarray(v2)
        self.label = _labels(label, 1)
        self.isfluid = isfluid

    def get_bounds(self):
        r = max(np.linalg.norm(self.v1), np.linalg.norm(self.v2))
        return self.center - r, self.center + r

    def point_inside(self, grid):
        x, y = grid
        px, py = x - self.center[0], y - self.center[1]
        cx2, cy2, cxy, _ = _ellipse_quadric(self.v1, self.v2)
        det = self.v1[0] * self.v2[1] - self.v1[1] * self.v2[0]
        return px ** 2 * cy2 + py ** 2 * cx2 - px * py * cxy <= det ** 2

    def distance(self, grid, v, dmax=None, normal=False):
        x, y = grid
        return _ellipse_distance(x - self.center[0], y - self.center[1], v, self.v1, self.v2, dmax,
                                 self.label[0])


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
        return _ellipsoid_quadric(self.radius * np.array([1, 0, 0]), self.radius * np.array([0, 1, 0]),
                                  self.radius * np.array([0, 0, 1]))

    def distance(self, grid, v, dmax=None, normal=False):
        return _ellipsoid_distance(grid, v, self.center, self.quadric(), dmax, self.label[0])


def _ellipsoid_quadric(e1, e2, e3):
    """rhs and the coefficients of cxx X^2 + cyy Y^2 + czz Z^2 + cxy XY + cyz YZ + czx ZX =
    rhs, the ellipsoid spanned by e1, e2, e3 about its centre (from the cross products)."""
    c12, c23, c31 = np.cross(e1, e2), np.cross(e2, e3), np.cross(e3, e1)
    rhs = np.inner(e1, c23) ** 2
    cxx = c12[0] ** 2 + c23[0] ** 2 + c31[0] ** 2
    cyy = c12[1] ** 2 + c23[1] ** 2 + c31[1] ** 2
    czz = c12[2] ** 2 + c23[2] ** 2 + c31[2] ** 2
    cxy = 2 * (c12[0] * c12[1] + c23[0] * c23[1] + c31[0] * c31[1])
    cyz = 2 * (c12[1] * c12[2] + c23[1] * c23[2] + c31[1] * c31[2])
    czx = 2 * (c12[2] * c12[0] + c23[2] * c23[0] + c31[2] * c31[0])
    return rhs, cxx, cyy, czz, cxy, cyz, czx


def _ellipsoid_distance(grid, v, center, quadric, dmax, label):
    """Fraction of v after which the segment from each grid point crosses the
    ellipsoid (utils.py:171-234)."""
    x, y, z = grid
    shape = (x.size, y.size, z.size)
    px, py, pz = x - center[0], y - center[1], z - center[2]
    rhs, cxx, cyy, czz, cxy, cyz, czx = quadric
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
    return _clip(d, dmax, label, shape)


class Ellipsoid:
    """The solid (or fluid) ellipsoid of centre c spanned by orthogonal v1, v2, v3."""
    number_of_bounds = 1
    dim = 3

    def __init__(self, center, v1, v2, v3, label=0, isfluid=False):
        p12 = abs(v1[0] * v2[0] + v1[1] * v2[1] + v1[2] * v2[2])
        p23 = abs(v2[0] * v3[0] + v2[1] * v3[1] + v2[2] * v3[2])
        p31 = abs(v3[0] * v1[0] + v3[1] * v1[1] + v3[2] * v1[2])
        if max(p12, p23, p31) > 1.0e-14:
            raise ValueError("the vectors of the ellipsoid are not orthogonal")
        self.center = np.asarray(center)
        self.v1, self.v2, self.v3 = np.asarray(v1), np.asarray(v2), np.asarray(v3)
        self.label = _labels(label, 1)
        self.isfluid = isfluid

    def get_bounds(self):
        r = max(np.linalg.norm(self.v1), np.linalg.norm(self.v2), np.linalg.norm(self.v3))
        return self.center - r, self.center + r

    def quadric(self):
        return _ellipsoid_quadric(self.v1, self.v2, self.v3)

    def point_inside(self, grid):
        x, y, z = grid
        px, py, pz = x - self.center[0], y - self.center[1], z - self.center[2]
        rhs, cxx, cyy, czz, cxy, cyz, czx = self.quadric()
        return (cxx * px ** 2 + cyy * py ** 2 + czz * pz ** 2 + cxy * px * py + cyz * py * pz
                + czx * pz * px <= rhs)

    def distance(self, grid, v, dmax=None, normal=False):
        return _ellipsoid_distance(grid, v, self.center, self.quadric(), dmax, self.label[0])


def _line_hits(px, py, v, starts, edges, dmax, labels):
    """Nearest crossing of the segment (p, p + v) with polygon edges (utils.py:21-90)."""
    shape = np.broadcast(px, py).shape
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


# ------------------------------------------------------------ 3D extrusions
class _UnitDisc:
    """The base of CylinderCircle / CylinderEllipse in the cylinder's own frame: the
    unit disc (pylbm/elements/base.py:136-362)."""

    @staticmethod
    def point_inside(grid):
        x, y = grid
        return (x ** 2 + y ** 2) <= 1.0

    @staticmethod
    def distance(grid, v, dmax, labels):
        x, y = grid
        return _ellipse_distance(x - 0.0, y - 0.0, v, np.asarray([1, 0]), np.asarray([0, 1]), dmax, labels[0])


class _UnitTriangle:
    """The base of CylinderTriangle: the triangle (0, 0), (1, 0), (0, 1)
    (pylbm/elements/base.py:365-480)."""

    @staticmethod
    def point_inside(grid):
        x, y = grid
        return np.logical_and(np.logical_and(x >= 0, y >= 0), x + y <= 1)

    @staticmethod
    def distance(grid, v, dmax, labels):
        x, y = grid
        return _line_hits(x, y, v, [[0, 0], [0, 0], [1, 0]], [[1, 0], [0, 1], [-1, 1]], dmax, labels)


class _UnitSquare:
    """The base of Parallelepiped: the unit square (pylbm/elements/base.py:483-608)."""

    @staticmethod
    def point_inside(grid):
        x, y = grid
        return np.logical_and(np.logical_and(x >= 0, y >= 0), np.logical_and(x <= 1, y <= 1))

    @staticmethod
    def distance(grid, v, dmax, labels):
        x, y = grid
        return _line_hits(x, y, v, [[0, 0], [0, 0], [1, 0], [0, 1]], [[1, 0], [0, 1], [0, 1], [1, 0]], dmax,
                          labels)


class _Cylinder:
    """A 2D base extruded along w (pylbm/elements/cylinder.py:27-216).

    Points are written in the frame (v1, v2, w) about the centre, where the element
    is base x [-1, 1]; the lateral faces take the base's labels and the bottom
    (z = -1) and top (z = 1) faces the last two.  A link's crossing is the nearest of
    the lateral crossing (kept when it lies between the end planes) and the end-plane
    crossings that land inside the base."""
    dim = 3
    _base = None

    def _frame(self, center, v1, v2, w, label, isfluid):
        self.center = np.asarray(center)
        self.v1, self.v2, self.w = np.asarray(v1), np.asarray(v2), np.asarray(w)
        A = np.empty((3, 3))
        A[:, 0], A[:, 1], A[:, 2] = self.v1, self.v2, self.w
        self.A = A
        self.iA = np.linalg.inv(A)
        self.label = _labels(label, self.number_of_bounds)
        if len(self.label) != self.number_of_bounds:
            raise ValueError("%s needs %d labels" % (type(self).__name__, self.number_of_bounds))
        self.isfluid = isfluid

    def _local(self, grid):
        x, y, z = grid
        xx, yy, zz = x - self.center[0], y - self.center[1], z - self.center[2]
        iA = self.iA
        return (iA[0, 0] * xx + iA[0, 1] * yy + iA[0, 2] * zz,
                iA[1, 0] * xx + iA[1, 1] * yy + iA[1, 2] * zz,
                iA[2, 0] * xx + iA[2, 1] * yy + iA[2, 2] * zz)

    def get_bounds(self):
        """pylbm's bounds (cylinder.py:51-64): the base's box widened by |w_k| / 2 along
        each axis.  The element itself spans the centre +- w, so a domain carves only
        its part within these bounds, as pylbm's does."""
        half = [0.5 * abs(self.w[k]) for k in range(len(self.w))]
        lo, hi = self._base_bounds()
        return lo - half, hi + half

    def point_inside(self, grid):
        xc, yc, zc = self._local(grid)
        return np.logical_and(self._base.point_inside((xc, yc)), np.abs(zc) <= 1.0)

    def distance(self, grid, v, dmax=None, normal=False):
        vc = self.iA.dot(np.asarray(v))
        xc, yc, zc = self._local(grid)
        alpha, border = self._base.distance((xc, yc), vc[:-1], dmax, self.label[:-2])
        alpha[alpha < 0] = _FAR
        beyond = np.logical_and(alpha > 0, np.abs(zc + alpha * vc[2]) > 1.0)
        alpha[beyond] = _FAR
        border[beyond] = -1.0
        tiny = 1.0e-16 if vc[2] == 0 else 0.0
        ends = []
        for top in (True, False):
            t = (1.0 - zc) / (vc[2] + tiny) if top else -(1.0 + zc) / (vc[2] + tiny)
            miss = np.logical_or(np.logical_or(t < 0, t > dmax),
                                 np.logical_not(self._base.point_inside((xc + t * vc[0], yc + t * vc[1]))))
            t[miss] = _FAR
            ends.append(t)
        alpha = np.amin([alpha, ends[0], ends[1]], axis=0)
        border[alpha == ends[0]] = self.label[-1]
        border[alpha == ends[1]] = self.label[-2]
        alpha[alpha == _FAR] = -1.0
        return alpha, border


class CylinderCircle(_Cylinder):
    """Circular cylinder: base the disc of centre c spanned by v1, v2 (equal norms),
    extruded over c + [-1, 1] w; labels: side, bottom, top."""
    number_of_bounds = 3
    _base = _UnitDisc

    def __init__(self, center, v1, v2, w, label=0, isfluid=False):
        if np.linalg.norm(v1) != np.linalg.norm(v2):
            raise ValueError("the vectors of the circular base must have the same norm")
        self._frame(center, v1, v2, w, label, isfluid)

    def _base_bounds(self):
        r = np.linalg.norm(self.v1)
        return self.center - r, self.center + r


class CylinderEllipse(_Cylinder):
    """Elliptic cylinder: base the ellipse of centre c spanned by orthogonal v1, v2,
    extruded over c + [-1, 1] w; labels: side, bottom, top."""
    number_of_bounds = 3
    _base = _UnitDisc

    def __init__(self, center, v1, v2, w, label=0, isfluid=False):
        if abs(np.inner(np.asarray(v1), np.asarray(v2))) > 1.0e-10:
            raise ValueError("the vectors of the elliptic base must be orthogonal")
        self._frame(center, v1, v2, w, label, isfluid)

    def _base_bounds(self):
        r = max(np.linalg.norm(self.v1), np.linalg.norm(self.v2))
        return self.center - r, self.center + r


class CylinderTriangle(_Cylinder):
    """Triangular prism: base the triangle c, c + v1, c + v2, extruded over
    c + [-1, 1] w; labels: the three sides (edges v1, v2, v2 - v1), bottom, top."""
    number_of_bounds = 5
    _base = _UnitTriangle

    def __init__(self, center, v1, v2, w, label=0, isfluid=False):
        self._frame(center, v1, v2, w, label, isfluid)

    def _base_bounds(self):
        box = np.asarray([self.center, self.center + self.v1, self.center + self.v1 + self.v2,
                          self.center + self.v2])
        return np.min(box, axis=0), np.max(box, axis=0)


class Parallelepiped(_Cylinder):
    """Parallelepiped point + [0, 1] v0 + [0, 1] v1 + [0, 1] v2 (the unit square
    extruded along v2 from its mid-plane); labels: the four faces along v0 and v1,
    the face at point, the opposite face."""
    number_of_bounds = 6
    _base = _UnitSquare

    def __init__(self, point, v0, v1, v2, label=0, isfluid=False):
        self.point = np.asarray(point)
        w = 0.5 * np.asarray(v2)
        self._frame(self.point + w, v0, v1, w, label, isfluid)

    def _base_bounds(self):
        box = np.asarray([self.center, self.center + self.v1, self.center + self.v1 + self.v2,
                          self.center + self.v2])
        return np.min(box, axis=0), np.max(box, axis=0)
