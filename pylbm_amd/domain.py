"""Computational domain: coordinates, solid mask and boundary links.

Produces the same ``in_or_out``, ``distance`` and ``flag`` values as
pylbm's ``Domain`` (pylbm/domain.py:252-635) -- box walls with half-link
distances (i + 0.5)/|v| where the first axis wins at corners, elements
added in order with the nearest crossing winning, links of solid cells
cleaned -- but organised for large grids: the per-velocity arrays are built
one velocity at a time (every velocity is independent in the reference
algorithm), and only the boundary links (positions, distances, labels) are
kept.  The reference holds three dense (unique velocities x cells) arrays,
about 26 GB for a 512 x 256 x 256 D3Q19 domain; this keeps one velocity's
arrays at a time.  ``distance`` / ``flag`` are still available as dense
properties for small domains.
"""
import numpy as np

from .elements import (Circle, CylinderCircle, CylinderEllipse, CylinderTriangle, Ellipse,  # noqa: F401
                       Ellipsoid, Parallelepiped, Parallelogram, Sphere, Triangle)  # re-export
from .stencil import Stencil, num_from_velocity

__all__ = ["Domain"]

VALIN = 999
VALOUT = -1


def _box(dico):
    box = dico["box"]
    bounds = [box["x"]]
    if box.get("y") is not None:
        bounds.append(box["y"])
        if box.get("z") is not None:
            bounds.append(box["z"])
    lab = box.get("label", -1)
    dim = len(bounds)
    if isinstance(lab, (int, np.integer)):
        labels = [int(lab)] * 2 * dim
    else:
        labels = list(lab)
        if len(labels) != 2 * dim:
            raise ValueError("the list label of the box has the wrong size (must be 2*dim)")
    return np.asarray(bounds, dtype="f8"), labels


class _ElementFootprint:
    """Box of an element on the halo grid, its sparse grid and inside mask."""

    def __init__(self, dom, elem):
        vmax = np.asarray(dom.stencil.vmax)
        lo, hi = elem.get_bounds()
        origin = np.asarray([c[0] for c in dom.coords_halo])
        self.nmin = np.maximum(vmax, np.array((lo - origin) / dom.dx, int) - vmax)
        self.nmax = np.minimum(vmax + np.asarray(dom.shape_in),
                               np.array((hi - origin) / dom.dx, int) + vmax + 1)
        self.sl = tuple(slice(a, b) for a, b in zip(self.nmin, self.nmax))
        self.grid = np.meshgrid(*(dom.coords_halo[d][s] for d, s in enumerate(self.sl)),
                                sparse=True, indexing="ij")
        self.inside = np.asarray(elem.point_inside(self.grid))
        shape = tuple(max(0, b - a) for a, b in zip(self.nmin, self.nmax))
        self.inside = np.broadcast_to(self.inside, shape)
        self.elem = elem

    def shifted(self, v):
        return tuple(slice(a + v[d], b + v[d]) for d, (a, b) in enumerate(zip(self.nmin, self.nmax)))


class Domain:
    """pylbm.Domain restricted to what the time-stepper needs (no normals, no viewer)."""

    def __init__(self, dico, need_validation=True):
        self.valin = VALIN
        self.valout = VALOUT
        self.bounds, self.box_label = _box(dico)
        self.dim = self.bounds.shape[0]
        self.elements = list(dico.get("elements", None) or [])
        for e in self.elements:
            if e.dim != self.dim:
                raise ValueError("element must have the same dimension as the box")
        self.stencil = Stencil(dico, need_validation=False)
        self.dx = dico["space_step"]
        self._create_coords()
        self._build_mask()
        self._links = None
        # pylbm's MPI topology (pylbm/domain.py:377-383), passed to H5File by pylbm's
        # scripts; one GPU and one process here, so there is none
        self.mpi_topo = None

    # ------------------------------------------------------------- coords
    def _create_coords(self):
        n = []
        for k in range(self.dim):
            g = (self.bounds[k][1] - self.bounds[k][0]) / self.dx
            if not float(g).is_integer():
                r = round(g)
                if abs(r - g) >= self.dx:
                    raise ValueError("the length of the box in direction %d must be a multiple of "
                                     "the space step (%.15f points)" % (k, g))
                g = r
            n.append(int(g))
        self.global_size = np.asarray(n, dtype=int)
        h = np.asarray(self.stencil.vmax)
        beg = self.dx * (h - 0.5)
        self.coords_halo = [np.linspace(self.bounds[k][0] - beg[k],
                                        self.bounds[k][0] + self.dx * n[k] + beg[k], n[k] + 2 * h[k])
                            for k in range(self.dim)]
        self.coords = [self.coords_halo[k][h[k]:self.coords_halo[k].size - h[k]] for k in range(self.dim)]

    @property
    def shape_halo(self):
        return [c.size for c in self.coords_halo]

    @property
    def shape_in(self):
        return [c.size for c in self.coords]

    @property
    def x(self):
        return self.coords[0]

    @property
    def y(self):
        return self.coords[1]

    @property
    def z(self):
        return self.coords[2]

    @property
    def x_halo(self):
        return self.coords_halo[0]

    @property
    def y_halo(self):
        return self.coords_halo[1]

    @property
    def z_halo(self):
        return self.coords_halo[2]

    def _interior(self):
        h = self.stencil.vmax
        return tuple(slice(hh, -hh) if hh > 0 else slice(None) for hh in h)

    # --------------------------------------------------------------- mask
    def _build_mask(self):
        self._box_fluid = np.zeros(self.shape_halo, dtype=bool)
        self._box_fluid[self._interior()] = True
        fluid = self._box_fluid.copy()
        self._foot = []
        for e in self.elements:
            fp = _ElementFootprint(self, e)
            self._apply(fluid, fp)
            self._foot.append(fp)
        self.fluid = fluid

    @staticmethod
    def _apply(fluid, fp):
        view = fluid[fp.sl]
        if fp.elem.isfluid:
            view[fp.inside] = True
        else:
            view[fp.inside] = False

    @property
    def in_or_out(self):
        return np.where(self.fluid, float(VALIN), float(VALOUT))

    # ------------------------------------------------ per-velocity arrays
    def velocity_arrays(self, u):
        """(distance, flag) of unique velocity ``u`` on the halo grid (reference semantics)."""
        shape = self.shape_halo
        dist = np.full(shape, float(VALIN))
        flag = np.full(shape, VALIN, dtype=np.int64 if np.prod(shape) < (1 << 22) else np.int32)
        v = np.asarray(self.stencil.uv[u])
        inner = list(self._interior())
        labels = self.box_label
        h = self.stencil.vmax
        for d in range(self.dim):
            vk = int(v[d])
            n = self.shape_in[d]
            if vk < 0 and labels[2 * d] != -2:
                layers = [(h[d] + i, -(i + 0.5) / vk, labels[2 * d]) for i in range(-vk)]
            elif vk > 0 and labels[2 * d + 1] != -2:
                layers = [(h[d] + n - 1 - i, (i + 0.5) / vk, labels[2 * d + 1]) for i in range(vk)]
            else:
                layers = []
            for idx, dv, lab in layers:
                sl = list(inner)
                sl[d] = slice(idx, idx + 1)
                dview = dist[tuple(sl)]
                fview = flag[tuple(sl)]
                sel = dview > dv
                dview[sel] = dv
                fview[sel] = lab
        if self._foot and np.any(v != 0):
            fluid_t = self._box_fluid.copy()
            for fp in self._foot:
                self._apply(fluid_t, fp)
                e = fp.elem
                out_cells = np.logical_not(fluid_t[fp.shifted(v)])
                alpha, border = e.distance(fp.grid, self.dx * v, 1.0)
                inside = fp.inside
                not_inside = np.logical_not(inside)
                indx = np.logical_and(alpha > 0, inside if e.isfluid else not_inside)
                if out_cells.size != 0:
                    indx = np.logical_and(indx, out_cells)
                dv = dist[fp.sl]
                fv = flag[fp.sl]
                if e.isfluid:
                    b2i = np.logical_and(np.logical_not(out_cells), fluid_t[fp.sl])
                    dv[b2i] = VALIN
                    fv[b2i] = VALIN
                else:
                    dv[inside] = VALIN
                    fv[inside] = VALIN
                ind4 = np.where(indx)
                if not e.isfluid:
                    pick = np.where(alpha[ind4] < dv[ind4])[0]
                else:
                    pick = np.where(np.logical_or(alpha[ind4] > dv[ind4], dv[ind4] == VALIN))[0]
                ind = tuple(i[pick] for i in ind4)
                dv[ind] = alpha[ind]
                fv[ind] = border[ind]
        solid_link = np.logical_and(dist > 0, np.logical_not(self.fluid))
        dist[solid_link] = VALIN
        flag[solid_link] = VALIN
        return dist, flag

    @property
    def distance(self):
        return np.stack([self.velocity_arrays(u)[0] for u in range(self.stencil.unvtot)])

    @property
    def flag(self):
        return np.stack([self.velocity_arrays(u)[1] for u in range(self.stencil.unvtot)])

    # -------------------------------------------------------------- labels
    def list_of_labels(self):
        labels = np.unique(self.box_label)
        el = np.empty(0)
        for e in self.elements:
            el = np.union1d(el, e.label)
        return np.union1d(labels, el)

    def links(self):
        """{(label, u): (flat C-order positions, distances)} of every labelled link.

        ``u`` is the unique velocity along which the fluid cell sees the label
        (the reference's ``flag[u] == label``)."""
        if self._links is None:
            wanted = [lab for lab in self.list_of_labels() if lab not in (-1, -2)]
            out = {}
            for u in (range(self.stencil.unvtot) if wanted else []):
                if not np.any(self.stencil.uv[u] != 0):
                    continue
                dist, flag = self.velocity_arrays(u)
                ff = flag.reshape(-1)
                cand = np.flatnonzero(ff != VALIN)
                for lab in wanted:
                    pos = cand[ff[cand] == lab]
                    if pos.size:
                        out[(lab, u)] = (pos, dist.reshape(-1)[pos])
            self._links = out
        return self._links

    def sym_index(self, u):
        """Unique-velocity index of -v_u."""
        v = [-int(c) for c in self.stencil.uv[u]]
        return int(self.stencil.unum2index[num_from_velocity(v)])
