"""Dense domain (solid mask, link distances, link flags) -- oracle restatement.

Follows pylbm/domain.py:
  * coordinates with halo (np.linspace)            domain.py:385-442
  * box walls: half-link distances (i+0.5)/|v|,
    first axis wins at corners                     domain.py:463-520
  * elements: int()-truncated boxes +- vmax, solid
    mask, per-velocity link distance, nearest wins domain.py:523-620
  * clean (drop links of solid cells)              domain.py:622-635
  * labels: sorted union                           domain.py:637-642
and the element geometry of pylbm/elements/circle.py:86-151,
sphere.py:88-150, parallelogram.py:92-175, triangle.py:103-149,
utils.py:21-277 (same floating-point operation order, so results are
bit-identical to numpy running the reference).

Pinned by: the seven fixtures tests/domain/data/*.npz (copied as data into
tests/golden/domain/) and the inline known answers of
tests/domain/test_domain1D.py:15-89, plus brute-force per-link checks on
tiny 3D boxes with a sphere (tests/test_oracle_domain.py).
"""
import numpy as np

from .stencil import Stencil

VALIN = 999
VALOUT = -1
TGV = 1.0e16


# ----------------------------------------------------------------- elements
def _intersection_two_lines(p1, v1, p2, v2):
    det = v1[1] * v2[0] - v1[0] * v2[1]
    if det == 0:
        return None, None
    invdet = 1.0 / det
    c1 = p2[0] - p1[0]
    c2 = p2[1] - p1[1]
    alpha = (-v2[1] * c1 + v2[0] * c2) * invdet
    beta = (-v1[1] * c1 + v1[0] * c2) * invdet
    return alpha, beta


def _distance_lines(x, y, v, p, vt, dmax, label):
    if x.shape[1] == 1 and y.shape[0] == 1:
        shape = (x.size, y.size)
    else:
        shape = x.shape
    alpha = TGV * np.ones(shape)
    border = -np.ones(shape)
    for i, vti in enumerate(vt):
        t1, t2 = _intersection_two_lines([x, y], v, p[i], vti)
        if t1 is None:
            continue
        if dmax is None:
            ok = np.logical_and(t1 > 0, np.logical_and(t2 >= 0, t2 <= 1))
        else:
            ok = np.logical_and(np.logical_and(t1 > 0, t1 <= dmax),
                                np.logical_and(t2 >= 0, t2 <= 1))
        ok = np.where(np.logical_and(alpha > t1, ok))
        alpha[ok] = t1[ok]
        border[ok] = label[i]
    alpha[alpha == TGV] = -1.0
    return alpha, border


def _quadratic_nearest(a, b, c, shape):
    delta = b ** 2 - 4 * a * c
    ind = delta >= 0
    delta[ind] = np.sqrt(delta[ind])
    d1 = TGV * np.ones(shape)
    d2 = TGV * np.ones(shape)
    if a != 0:
        d1[ind] = (-b[ind] - delta[ind]) / (2 * a)
        d2[ind] = (-b[ind] + delta[ind]) / (2 * a)
    d1[d1 < 0] = TGV
    d2[d2 < 0] = TGV
    d = -np.ones(shape)
    d[ind] = np.minimum(d1[ind], d2[ind])
    d[d == TGV] = -1
    return d


class Circle:
    number_of_bounds = 1
    dim = 2

    def __init__(self, center, radius, label=0, isfluid=False):
        self.center = np.asarray(center)
        self.radius = radius
        self.label = [label] if isinstance(label, int) else list(label)
        self.isfluid = isfluid

    def get_bounds(self):
        return self.center - self.radius, self.center + self.radius

    def point_inside(self, grid):
        x, y = grid
        return ((x - self.center[0]) ** 2 + (y - self.center[1]) ** 2) <= self.radius ** 2

    def distance(self, grid, v, dmax):
        x, y = grid
        v1 = self.radius * np.array([1, 0])
        v2 = self.radius * np.array([0, 1])
        Xc = x - self.center[0]
        Yc = y - self.center[1]
        vx2 = v1[0] ** 2 + v2[0] ** 2
        vy2 = v1[1] ** 2 + v2[1] ** 2
        vxy = 2 * (v1[0] * v1[1] + v2[0] * v2[1])
        dd = (v1[0] * v2[1] - v1[1] * v2[0]) ** 2
        a = v[0] ** 2 * vy2 + v[1] ** 2 * vx2 - v[0] * v[1] * vxy
        b = 2 * Xc * v[0] * vy2 + 2 * Yc * v[1] * vx2 - (Xc * v[1] + Yc * v[0]) * vxy
        c = Xc ** 2 * vy2 + Yc ** 2 * vx2 - Xc * Yc * vxy - dd
        shape = (b ** 2 - 4 * a * c).shape
        d = _quadratic_nearest(a, b, c, shape)
        alpha = -np.ones(shape)
        border = -np.ones(shape)
        ind = np.logical_and(d > 0, d <= dmax)
        alpha[ind] = d[ind]
        border[ind] = self.label[0]
        return alpha, border


class Sphere:
    number_of_bounds = 1
    dim = 3

    def __init__(self, center, radius, label=0, isfluid=False):
        self.center = np.asarray(center)
        self.radius = radius
        self.label = [label] if isinstance(label, int) else list(label)
        self.isfluid = isfluid

    def get_bounds(self):
        return self.center - self.radius, self.center + self.radius

    def point_inside(self, grid):
        x, y, z = grid
        return ((x - self.center[0]) ** 2 + (y - self.center[1]) ** 2
                + (z - self.center[2]) ** 2) <= self.radius ** 2

    def distance(self, grid, v, dmax):
        x, y, z = grid
        v1 = self.radius * np.array([1, 0, 0])
        v2 = self.radius * np.array([0, 1, 0])
        v3 = self.radius * np.array([0, 0, 1])
        shape = (x.size, y.size, z.size)
        Xc = x - self.center[0]
        Yc = y - self.center[1]
        Zc = z - self.center[2]
        v12 = np.cross(v1, v2)
        v23 = np.cross(v2, v3)
        v31 = np.cross(v3, v1)
        dd = np.inner(v1, v23) ** 2
        cxx = v12[0] ** 2 + v23[0] ** 2 + v31[0] ** 2
        cyy = v12[1] ** 2 + v23[1] ** 2 + v31[1] ** 2
        czz = v12[2] ** 2 + v23[2] ** 2 + v31[2] ** 2
        cxy = 2 * (v12[0] * v12[1] + v23[0] * v23[1] + v31[0] * v31[1])
        cyz = 2 * (v12[1] * v12[2] + v23[1] * v23[2] + v31[1] * v31[2])
        czx = 2 * (v12[2] * v12[0] + v23[2] * v23[0] + v31[2] * v31[0])
        a = (cxx * v[0] ** 2 + cyy * v[1] ** 2 + czz * v[2] ** 2
             + cxy * v[0] * v[1] + cyz * v[1] * v[2] + czx * v[2] * v[0])
        b = ((2 * cxx * v[0] + cxy * v[1] + czx * v[2]) * Xc
             + (2 * cyy * v[1] + cyz * v[2] + cxy * v[0]) * Yc
             + (2 * czz * v[2] + czx * v[0] + cyz * v[1]) * Zc)
        c = (cxx * Xc ** 2 + cyy * Yc ** 2 + czz * Zc ** 2
             + cxy * Xc * Yc + cyz * Yc * Zc + czx * Zc * Xc - dd)
        delta = b ** 2 - 4 * a * c
        ind = delta >= 0
        delta[ind] = np.sqrt(delta[ind])
        d1 = 1e16 * np.ones(shape)
        d2 = 1e16 * np.ones(shape)
        d1[ind] = (-b[ind] - delta[ind]) / (2 * a)
        d2[ind] = (-b[ind] + delta[ind]) / (2 * a)
        d1[d1 < 0] = 1e16
        d2[d2 < 0] = 1e16
        d = -np.ones(shape)
        d[ind] = np.minimum(d1[ind], d2[ind])
        d[d == 1e16] = -1
        alpha = -np.ones(shape)
        border = -np.ones(shape)
        ind = np.logical_and(d > 0, d <= dmax)
        alpha[ind] = d[ind]
        border[ind] = self.label[0]
        return alpha, border


class Parallelogram:
    number_of_bounds = 4
    dim = 2

    def __init__(self, point, vecta, vectb, label=0, isfluid=False):
        self.point = np.asarray(point)
        self.v1 = np.asarray(vecta)
        self.v2 = np.asarray(vectb)
        self.label = [label] * 4 if isinstance(label, int) else list(label)
        self.isfluid = isfluid

    def get_bounds(self):
        box = np.asarray([self.point, self.point + self.v1,
                          self.point + self.v1 + self.v2, self.point + self.v2])
        return np.min(box, axis=0), np.max(box, axis=0)

    def _uv(self, grid):
        x, y = grid
        w = [x - self.point[0], y - self.point[1]]
        invdelta = 1.0 / (self.v1[0] * self.v2[1] - self.v1[1] * self.v2[0])
        u = (w[0] * self.v2[1] - w[1] * self.v2[0]) * invdelta
        v = (w[1] * self.v1[0] - w[0] * self.v1[1]) * invdelta
        return u, v

    def point_inside(self, grid):
        u, v = self._uv(grid)
        return np.logical_and(np.logical_and(u >= 0, v >= 0), np.logical_and(u <= 1, v <= 1))

    def distance(self, grid, v, dmax):
        x, y = grid
        p = [[0, 0], [0, 0], self.v2, self.v1]
        vt = [self.v1, self.v2, self.v1, self.v2]
        return _distance_lines(x - self.point[0], y - self.point[1], v, p, vt, dmax, self.label)


class Triangle(Parallelogram):
    number_of_bounds = 3

    def __init__(self, point, vecta, vectb, label=0, isfluid=False):
        super().__init__(point, vecta, vectb, label, isfluid)
        self.label = [label] * 3 if isinstance(label, int) else list(label)

    def get_bounds(self):
        box = np.asarray([self.point, self.point + self.v1, self.point + self.v2])
        return np.min(box, axis=0), np.max(box, axis=0)

    def point_inside(self, grid):
        u, v = self._uv(grid)
        return np.logical_and(np.logical_and(u >= 0, v >= 0), u + v <= 1)

    def distance(self, grid, v, dmax):
        x, y = grid
        p = [[0, 0], [0, 0], self.v1]
        vt = [self.v1, self.v2, self.v2 - self.v1]
        return _distance_lines(x - self.point[0], y - self.point[1], v, p, vt, dmax, self.label)


def _as_oracle_element(e):
    """Rebuild a user element (any package) as the oracle's own element class."""
    name = type(e).__name__
    if name == "Circle":
        return Circle(e.center, e.radius, e.label, e.isfluid)
    if name == "Sphere":
        return Sphere(e.center, e.radius, e.label, e.isfluid)
    if name == "Parallelogram":
        return Parallelogram(e.point, e.v1, e.v2, e.label, e.isfluid)
    if name == "Triangle":
        return Triangle(e.point, e.v1, e.v2, e.label, e.isfluid)
    raise NotImplementedError("oracle: element %s" % name)


# ------------------------------------------------------------------- domain
class Domain:
    def __init__(self, dico):
        box = dico["box"]
        bounds = [box["x"]] + ([box["y"]] if "y" in box else []) + \
                 ([box["z"]] if "z" in box and "y" in box else [])
        self.dim = len(bounds)
        self.bounds = np.asarray(bounds, dtype="f8")
        lab = box.get("label", -1)
        self.box_label = [lab] * 2 * self.dim if isinstance(lab, int) else list(lab)
        self.elements = [_as_oracle_element(e) for e in (dico.get("elements", []) or [])]
        self.stencil = Stencil(self.dim, [s["velocities"] for s in dico["schemes"]])
        self.dx = dico["space_step"]
        self._coords()
        shape = list(self.shape_halo)
        st = self.stencil
        self.in_or_out = VALIN * np.ones(shape)
        self.distance = VALIN * np.ones([st.unvtot] + shape)
        self.flag = VALIN * np.ones([st.unvtot] + shape, dtype="int")
        self._add_box()
        for e in self.elements:
            self._add_elem(e)
        self._clean()

    def _coords(self):
        n = []
        for k in range(self.dim):
            g = (self.bounds[k][1] - self.bounds[k][0]) / self.dx
            if not g.is_integer():
                r = round(g)
                if abs(r - g) < self.dx:
                    g = r
                else:
                    raise ValueError("box length is not a multiple of the space step")
            n.append(int(g))
        self.global_size = np.asarray(n, dtype=int)
        h = np.asarray(self.stencil.vmax)
        hb = self.dx * (h - 0.5)
        self.coords_halo = [np.linspace(self.bounds[k][0] - hb[k],
                                        self.bounds[k][0] + self.dx * n[k] + hb[k],
                                        n[k] + 2 * h[k]) for k in range(self.dim)]
        self.coords = [self.coords_halo[k][h[k]:(-h[k] if h[k] > 0 else None)]
                       for k in range(self.dim)]
        self.shape_halo = [c.size for c in self.coords_halo]
        self.shape_in = [c.size for c in self.coords]

    def _add_box(self):
        h = np.asarray(self.stencil.vmax)
        label = self.box_label
        inner = tuple(slice(hh, -hh) if hh > 0 else slice(None) for hh in h)
        self.in_or_out[:] = VALOUT
        self.in_or_out[inner] = VALIN
        uv = self.stencil.uv
        for d in range(self.dim):
            n = self.shape_in[d]
            for k in range(self.stencil.unvtot):
                vk = int(uv[k, d])
                if vk < 0 and label[2 * d] != -2:
                    layers = [(i, -(i + 0.5) / vk, label[2 * d]) for i in range(-vk)]
                elif vk > 0 and label[2 * d + 1] != -2:
                    layers = [(n - 1 - i, (i + 0.5) / vk, label[2 * d + 1]) for i in range(vk)]
                else:
                    layers = []
                for li, dvik, lab in layers:
                    sl = list(inner)
                    sl[d] = slice(h[d] + li, h[d] + li + 1)
                    dist = self.distance[k][tuple(sl)]
                    flg = self.flag[k][tuple(sl)]
                    sel = dist > dvik
                    dist[sel] = dvik
                    flg[sel] = lab

    def _add_elem(self, elem):
        st = self.stencil
        vmax = np.asarray(st.vmax)
        ebl, eur = elem.get_bounds()
        pbl = np.asarray([c[0] for c in self.coords_halo])
        nmin = np.maximum(vmax, np.array((ebl - pbl) / self.dx, int) - vmax)
        nmax = np.minimum(vmax + np.asarray(self.shape_in),
                          np.array((eur - pbl) / self.dx, int) + vmax + 1)
        sl = tuple(slice(a, b) for a, b in zip(nmin, nmax))
        ioo = self.in_or_out[sl]
        grid = np.meshgrid(*(self.coords_halo[d][s] for d, s in enumerate(sl)),
                           sparse=True, indexing="ij")
        inside = elem.point_inside(grid)
        if not elem.isfluid:
            ind_solid = inside
            ind_fluid = np.logical_not(inside)
            ioo[ind_solid] = VALOUT
        else:
            ind_fluid = inside
            ind_solid = np.logical_not(inside)
            ioo[ind_fluid] = VALIN
        for k in range(st.unvtot):
            vk = st.uv[k]
            if not np.any(vk != 0):
                continue
            shifted = tuple(slice(a + vk[d], b + vk[d]) for d, (a, b) in enumerate(zip(nmin, nmax)))
            out_cells = self.in_or_out[shifted] == VALOUT
            alpha, border = elem.distance(grid, self.dx * vk, 1.0)
            indx = np.logical_and(alpha > 0, ind_fluid)
            if out_cells.size != 0:
                indx = np.logical_and(indx, out_cells)
            dk = self.distance[k][sl]
            fk = self.flag[k][sl]
            if elem.isfluid:
                b2i = np.logical_and(np.logical_not(out_cells), ioo == VALIN)
                dk[b2i] = VALIN
                fk[b2i] = VALIN
            else:
                dk[ind_solid] = VALIN
                fk[ind_solid] = VALIN
            ind4 = np.where(indx)
            if not elem.isfluid:
                ind3 = np.where(alpha[ind4] < dk[ind4])[0]
            else:
                ind3 = np.where(np.logical_or(alpha[ind4] > dk[ind4], dk[ind4] == VALIN))[0]
            ind = tuple(i[ind3] for i in ind4)
            dk[ind] = alpha[ind]
            fk[ind] = border[ind]

    def _clean(self):
        for k in range(self.stencil.unvtot):
            sel = np.logical_and(self.distance[k] > 0, self.in_or_out == VALOUT)
            self.distance[k][sel] = VALIN
            self.flag[k][sel] = VALIN

    def list_of_labels(self):
        labels = np.unique(self.box_label)
        el = np.empty(0)
        for e in self.elements:
            el = np.union1d(el, e.label)
        return np.union1d(labels, el)
