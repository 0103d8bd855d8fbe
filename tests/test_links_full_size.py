"""Link lists of the full-size sphere packing (512 x 256 x 256 D3Q19, about 7.8 M
bounce-back links): the invariants of pylbm's link construction
(pylbm/boundary.py:23-47, 411-419) checked over every link, vectorised.

* every store slot (k, x_out) is unique;
* every load (ksym, x_out + v_k) is a fluid cell and ksym is the opposite velocity;
* the set of store slots equals the brute-force set of (k, x - v_k), x a fluid
  interior cell and y = x - v_k not fluid, where either y is not across a
  periodic face, or y is across one and the link x -> y meets a sphere surface
  nearer than half a link.  The second case is pylbm's element rule
  (pylbm/domain.py:565-620): an element's distance replaces the box's (1/2 on a
  periodic face, label -1, no link) only when it is smaller, so a sphere near
  the periodic face gets a link across it when its surface lies within half a
  link of the fluid cell.  That distance is computed here from the sphere list
  by the quadratic formula, independently of the domain's distance arrays.

The brute-force rule is the one tests/test_links.py applies cell by cell to the
oracle's lists at small sizes; here it is first checked against that loop on a
small packing, then applied to the full one.
"""
import numpy as np
import pytest

from pylbm_amd import cases
from pylbm_amd.boundary import Boundary
from pylbm_amd.domain import Domain


def _crossings(x, y, spheres):
    """For each segment x[i] -> y[i], the smallest fraction t in (0, 1] at which it meets
    the surface of one of the spheres (inf when it meets none): pylbm's element
    distance (pylbm/elements/utils.py:93-167, the smaller positive root of the
    quadratic, dmax = 1), minimised over the elements."""
    best = np.full(x.shape[0], np.inf)
    d = y - x
    mid = 0.5 * (x + y)
    for c, r in spheres:
        sel = np.sum((mid - c) ** 2, axis=1) <= (r + 1.0) ** 2
        if not sel.any():
            continue
        xs, ds = x[sel] - c, d[sel]
        a = np.sum(ds * ds, axis=1)
        b = 2.0 * np.sum(ds * xs, axis=1)
        cc = np.sum(xs * xs, axis=1) - r * r
        disc = b * b - 4.0 * a * cc
        ok = disc >= 0
        sq = np.sqrt(np.where(ok, disc, 0.0))
        t1 = (-b - sq) / (2.0 * a)
        t2 = (-b + sq) / (2.0 * a)
        t1 = np.where(t1 > 0, t1, np.inf)
        t2 = np.where(t2 > 0, t2, np.inf)
        t = np.where(ok, np.minimum(t1, t2), np.inf)
        t = np.where(t <= 1.0, t, np.inf)
        best[sel] = np.minimum(best[sel], t)
    return best


def _brute_force_slots(dom, v, vmax, box_label, spheres):
    """Sorted flat (k, x_out) slots of every link by the brute-force rule."""
    fluid = dom.in_or_out == 999
    shape = fluid.shape
    inner = np.zeros(shape, dtype=bool)
    inner[tuple(slice(h, n - h) for h, n in zip(vmax, shape))] = True
    # halo cells across a periodic face carry no link (the pull reads the wrap)
    periodic_halo = np.zeros(shape, dtype=bool)
    for a in range(len(shape)):
        if box_label[2 * a] == -1:
            lo = [slice(None)] * len(shape)
            hi = [slice(None)] * len(shape)
            lo[a] = slice(0, vmax[a])
            hi[a] = slice(shape[a] - vmax[a], shape[a])
            periodic_halo[tuple(lo)] = True
            periodic_halo[tuple(hi)] = True
    x_ok = fluid & inner
    # spheres that can reach a periodic face; the others cannot meet a link across it
    spheres = [(c, r) for c, r in spheres
               if any(box_label[2 * a] == -1 and (c[a] - r < vmax[a] + 1 or c[a] + r > shape[a] - 2 * vmax[a] - 1)
                      for a in range(len(shape)))]
    slots = []
    for k, vk in enumerate(v):
        # y = x - v_k; every interior x has y inside the halo-inclusive array
        y_bad = ~fluid & ~periodic_halo
        shift = tuple(int(c) for c in vk)
        axes = tuple(range(len(shape)))
        hit = x_ok & np.roll(y_bad, shift=shift, axis=axes)
        xs = np.nonzero(hit)
        y = tuple(xs[a] - int(vk[a]) for a in range(len(shape)))
        slots.append(np.ravel_multi_index((np.full(y[0].size, k),) + y, (len(v),) + shape))
        # solid cells across a periodic face: a link when a sphere is met nearer than 1/2
        cand = x_ok & np.roll(~fluid & periodic_halo, shift=shift, axis=axes)
        xs = np.stack(np.nonzero(cand), axis=1)
        if xs.size:
            ys = xs - vk
            # cell centres: halo-inclusive index i sits at i - 1/2 (dx = 1, box from 0)
            t = _crossings(xs - 0.5, ys - 0.5, spheres)
            assert not np.any(np.abs(t - 0.5) < 1e-9)        # no tie with the box's 1/2
            ys = ys[t < 0.5]
            slots.append(np.ravel_multi_index((np.full(ys.shape[0], k),) + tuple(ys.T), (len(v),) + shape))
    return np.sort(np.concatenate(slots))


def _product_links(d):
    dom = Domain(d)
    b = Boundary(dom, d)
    for m in b.methods:
        m.set_iload()
    return dom, b.methods


def _check(d, spheres):
    from pylbm_amd.scheme import Scheme
    dom, methods = _product_links(d)
    st = Scheme(d).stencil
    v = np.asarray(st.get_all_velocities())
    ksym = np.asarray(st.get_symmetric())
    fluid = dom.in_or_out == 999
    shape = fluid.shape
    slots, total = [], 0
    for m in methods:
        k = m.istore[0]
        x = m.istore[1:] + v[k].T
        assert fluid[tuple(x)].all()                                  # x_out + v_k is fluid
        ld = m.iload[0].T if m._fixed else m.iload[0]
        assert np.array_equal(ld[0], ksym[k])                         # load population ksym
        assert np.array_equal(ld[1:], x)                              # at x_out + v_k
        flat = np.ravel_multi_index(tuple(m.istore), (len(v),) + shape)
        assert np.unique(flat).size == flat.size                      # stores unique
        slots.append(flat)
        total += k.size
    got = np.sort(np.concatenate(slots))
    want = _brute_force_slots(dom, v, st.vmax, d["box"]["label"],
                              [(np.asarray(c, dtype=np.float64), float(r)) for c, r in spheres])
    assert got.size == total
    assert np.array_equal(got, want)
    return total


def test_brute_force_rule_small_packing():
    """A small packing (the size tests/test_links.py checks cell by cell against the oracle)."""
    args = dict(seed=3, solid_fraction=0.15)
    sph = cases.sphere_centres((20, 14, 12), **args)
    assert _check(cases.sphere_packing_d3q19(20, 14, 12, **args), sph) > 0


@pytest.mark.slow
def test_full_size_sphere_packing_links():
    sph = cases.sphere_centres((512, 256, 256), seed=2024, solid_fraction=0.10)
    n = _check(cases.sphere_packing_d3q19(), sph)
    assert n > 7_000_000, n
