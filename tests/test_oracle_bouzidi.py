"""Pins of the oracle's interpolated (Bouzidi) bounce-back (oracle/boundary.py).

* Every wall at half a link (q = 1/2, the box walls) makes s = 1 and the
  second load's weight 0, so Bouzidi bounce-back / anti-bounce-back equal plain
  bounce-back / anti-bounce-back bit for bit (pins load0, the signs and rhs).
* Loads by geometry: for each link, x_out + v_k is a fluid cell; for q < 1/2 the
  second load is population ksym one link further into the fluid, otherwise the
  incoming population k at the fluid cell; s = 2q or 1/(2q).
* A force-driven channel between walls at fractional positions reaches an exact
  parabola (fit residual ~1e-10) whose roots are the effective walls: plain
  bounce-back puts them at the half-link positions whatever the true walls,
  Bouzidi moves them to within 0.08 of the true walls (the remaining offset is
  the known relaxation-dependent slip of linear interpolation).  Taking the
  wrong branch, the wrong second load or the wrong s moves the roots away.
  (Walls with q < 1/2 on both sides: with q > 1/2 the diagonal links of the
  edge columns that cross the periodic x face keep the box's periodic flag,
  pylbm/domain.py:603-610 replacing a distance only by a smaller one, so those
  links get no boundary value and the profile is not a parabola for plain
  bounce-back either.)
"""
import numpy as np
import pytest

from oracle.lbm import Simulation as Oracle
from pylbm_amd import bc, cases


def _run(d, n):
    o = Oracle(d)
    for _ in range(n):
        o.one_time_step()
    return o


@pytest.mark.parametrize("make", [lambda: cases.lid_cavity_d2q9(12),
                                  lambda: {k: v for k, v in cases.karman_d2q9(24, 12).items() if k != "elements"}])
def test_bouzidi_at_half_link_equals_bounce_back(make):
    d = make()
    a = _run(d, 15)
    b = _run(cases.with_bouzidi(d), 15)
    assert any(m.kind.startswith("Bouzidi") for m in b.methods)
    for m in b.methods:
        if m.kind.startswith("Bouzidi"):
            assert np.all(m.distance == 0.5) and np.all(m.s == 1.0)
    assert np.array_equal(a.F, b.F)


def test_bouzidi_loads_by_geometry():
    d = cases.poiseuille_d2q9(y_low=0.8, y_high=10.6, method=bc.BouzidiBounceBack)
    o = Oracle(d)
    (m,) = o.methods
    st = o.scheme.stencil
    v, ksym = st.all_velocities(), st.symmetric()
    from oracle.domain import VALIN
    fluid = o.domain.in_or_out == VALIN
    k = m.istore[0]
    xf = m.istore[1:] + v[k].T
    assert np.all(fluid[tuple(xf)])
    y = o.domain.coords_halo[1][xf[1]]
    # walls: horizontal faces y = 0.8 (below, q = 0.7) and y = 10.6 (above, q = 0.1)
    q = np.where(v[k][:, 1] > 0, y - 0.8, 10.6 - y) / np.abs(v[k][:, 1])
    assert np.allclose(m.distance, q, atol=1e-12)
    near = q < 0.5
    assert np.allclose(m.s, np.where(near, 2 * q, 0.5 / q), rtol=1e-14)
    assert np.array_equal(m.iload[0], ksym[k])
    assert np.array_equal(m.iload1[0], np.where(near, ksym[k], k))
    assert np.array_equal(m.iload1[1:], np.where(near, xf + v[k].T, xf))
    assert near.any() and (~near).any()


@pytest.mark.parametrize("walls", [(1.1, 10.6), (1.45, 10.8)])
def test_bouzidi_channel_wall_positions(walls):
    yb, yt = walls
    y = np.arange(12) + 0.5
    inside = (y > yb) & (y < yt)
    roots = {}
    for name, meth in (("bb", bc.BounceBack), ("bz", bc.BouzidiBounceBack)):
        o = _run(cases.poiseuille_d2q9(y_low=yb, y_high=yt, method=meth), 3000)
        ux = (o.m(cases.QX) / o.m(cases.RHO))[0]
        c = np.polyfit(y[inside], ux[inside], 2)
        assert np.max(np.abs(np.polyval(c, y[inside]) - ux[inside])) < 1e-6 * ux.max()
        roots[name] = np.sort(np.roots(c).real)
    assert np.allclose(roots["bb"], [1.0, 11.0], atol=0.01)
    assert np.allclose(roots["bz"], [yb, yt], atol=0.08)
