"""End-to-end runs against the reference's own fixtures.

pylbm's demo tests run ``run(dx, Tf)`` of a demo and compare every conserved
moment, solid cells set to 0, with tests/reference/<name>.h5 at
``rel=1e-14, abs=1e-7`` (pylbm tests/conftest.py:233-262, 306-307; the
comparison is ``pytest.approx``, so a value passes when it is within
max(1e-14 |ref|, 1e-7) of the fixture).  The fixtures are converted to
tests/golden/h5/*.npz by tests/golden/make_h5_fixtures.py, and the demos'
dictionaries are in pylbm_amd/demos.py.

* CPU: the oracle (numpy FP64) against every fixture of DEMOS -- this pins the
  oracle's whole step (sources, Bouzidi, Neumann, valued boundaries) to the
  reference's generated code.
* GPU: the HIP FP64 path against the same fixtures, at the same tolerance.
"""
import os

import numpy as np
import pytest

from pylbm_amd.demos import DEMOS

HERE = os.path.dirname(os.path.abspath(__file__))
RTOL, ATOL = 1e-14, 1e-7
VALIN = 999.0


def _fixture(name):
    return np.load(os.path.join(HERE, "golden", "h5", name + ".npz"))


def _assert_close(name, sym, got, ref):
    ref = np.asarray(ref)
    assert got.shape == ref.shape, (name, sym, got.shape, ref.shape)
    bound = np.maximum(RTOL * np.abs(ref), ATOL)
    bad = np.abs(got - ref) > bound
    assert not bad.any(), "%s %s: %d values off, max |diff| %.3e" % (
        name, sym, int(bad.sum()), float(np.max(np.abs(got - ref))))


def _interior_mask(in_or_out, halo):
    sl = tuple(slice(int(h), in_or_out.shape[i] - int(h)) for i, h in enumerate(halo))
    return in_or_out[sl] == VALIN


@pytest.mark.parametrize("name", sorted(DEMOS))
def test_oracle_matches_reference_fixture(name):
    from oracle.lbm import Simulation

    build, dx, Tf = DEMOS[name]
    dico, syms = build(dx)
    sim = Simulation(dico)
    while sim.t < Tf:
        sim.one_time_step()
    fluid = _interior_mask(sim.domain.in_or_out, sim.h)
    ref = _fixture(name)
    for s in syms:
        got = sim.m(s).copy()
        got[~fluid] = 0.0
        _assert_close(name, s, got, ref[s.name])


@pytest.mark.gpu
@pytest.mark.parametrize("name", sorted(DEMOS))
def test_hip_matches_reference_fixture(hip, name):
    from pylbm_amd import Simulation

    build, dx, Tf = DEMOS[name]
    dico, syms = build(dx)
    sim = Simulation(dico)
    while sim.t < Tf:
        sim.one_time_step()
    fluid = _interior_mask(sim.domain.in_or_out, sim.domain.stencil.vmax)
    ref = _fixture(name)
    for s in syms:
        got = np.array(sim.m[s], dtype=np.float64)
        got[~fluid] = 0.0
        _assert_close(name, s, got, ref[s.name])
