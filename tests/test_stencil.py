"""Velocity numbering: oracle and product against the reference's tables
(tests/conftest.py:19-165 of the reference, as tests/golden/stencil_tables.json)
and the symmetry involution of the reference's tests/test_velocity.py:14-32."""
import json
import os

import numpy as np
import pytest

import oracle.stencil as ost
from pylbm_amd.stencil import Stencil, Velocity, num_from_velocity, velocity_from_num

TABLES = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "stencil_tables.json")))


def _entries():
    for name, t in TABLES.items():
        dim = int(name[1])
        for sub in (t if isinstance(t, list) else [t]):
            yield name, dim, sub


@pytest.mark.parametrize("name,dim,tab", list(_entries()), ids=lambda x: x if isinstance(x, str) else "")
def test_tables(name, dim, tab):
    comps = [tab["vx"], tab.get("vy"), tab.get("vz")][:dim]
    for i, num in enumerate(tab["num"]):
        v = tuple(c[i] for c in comps)
        assert ost.num_to_v(dim, num) == v
        assert ost.v_to_num(v) == num
        assert velocity_from_num(dim, num) == v
        assert num_from_velocity(v) == num


@pytest.mark.parametrize("dim", [1, 2, 3])
def test_symmetric_involution(dim):
    rng = np.random.default_rng(dim)
    for num in rng.integers(0, 100 if dim < 3 else 300, size=100):
        v = Velocity(dim=dim, num=int(num))
        for axis in [None] + list(range(dim)):
            assert v.get_symmetric(axis).get_symmetric(axis).num == v.num
        s = v.get_symmetric()
        assert s.v == [-c for c in v.v]


def test_stencil_fields_match_oracle():
    for name, t in TABLES.items():
        dim = int(name[1])
        subs = t if isinstance(t, list) else [t]
        d = {"dim": dim, "schemes": [{"velocities": s["num"]} for s in subs]}
        st = Stencil(d)
        o = ost.Stencil(dim, [s["num"] for s in subs])
        assert np.array_equal(st.get_all_velocities(), o.all_velocities())
        assert np.array_equal(st.nv_ptr, o.nv_ptr)
        assert np.array_equal(st.unum, o.unum)
        assert np.array_equal(st.vmax, o.vmax)
        if all(num_from_velocity([-c for c in v]) in s["num"]
               for s in subs for v in [velocity_from_num(dim, n) for n in s["num"]]):
            assert np.array_equal(st.get_symmetric(), o.symmetric())
