"""Boundary link lists: product (pylbm_amd.boundary) == oracle (oracle.boundary),
bit for bit in content and order, plus the invariants that pin the oracle
(pylbm/boundary.py:23-160, 411-419, 777-890)."""
import numpy as np
import pytest

from oracle.boundary import build_links
from oracle.domain import Domain as ODomain
from oracle.lbm import method_kind
from pylbm_amd import bc, cases
from pylbm_amd.boundary import Boundary, _resolve_sequential
from pylbm_amd.domain import Domain

CASES = {
    "lid": lambda: cases.lid_cavity_d2q9(16),
    "karman": lambda: cases.karman_d2q9(64, 32, 5.0, (16.0, 16.0)),
    "spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.15),
    "d1q3": lambda: cases.d1q3_advection(12, walls=True),
    "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
}


def _product(d):
    dom = Domain(d)
    b = Boundary(dom, d)
    for m in b.methods:
        m.set_iload()
    return dom, b.methods


def _oracle(d):
    dom = ODomain(d)
    from oracle.scheme import Scheme
    st = Scheme(d).stencil
    return dom, build_links(dom, st, d["boundary_conditions"], method_kind)


@pytest.mark.parametrize("name", sorted(CASES))
def test_links_identical(name):
    d = CASES[name]()
    _, pm = _product(d)
    _, om = _oracle(d)
    assert [m.name for m in pm] == [m.kind for m in om]
    for a, b in zip(pm, om):
        assert np.array_equal(a.istore, b.istore)
        assert np.array_equal(a.iload[0], b.iload)
        assert np.array_equal(a.distance, b.distance)
        assert np.array_equal(a.ilabel, b.ilabel)


@pytest.mark.parametrize("name", sorted(CASES))
def test_link_invariants(name):
    d = CASES[name]()
    dom, om = _oracle(d)
    from oracle.scheme import Scheme
    st = Scheme(d).stencil
    v = st.all_velocities()
    fluid = dom.in_or_out == 999
    total = 0
    for m in om:
        k = m.istore[0]
        x = m.istore[1:] + v[k].T
        assert fluid[tuple(x)].all()                           # x_out + v_k is fluid
        total += k.size
        # store slots are unique
        flat = np.ravel_multi_index(tuple(m.istore), (st.nv_ptr[-1],) + tuple(dom.shape_halo))
        assert np.unique(flat).size == flat.size
    # brute-force count: (fluid x, k) with x - v_k outside the fluid and not periodic
    per = [lab == -1 for lab in dom.box_label]
    count = 0
    shape = dom.shape_halo
    for k, vk in enumerate(v):
        for x in zip(*np.nonzero(fluid)):
            y = [x[a] - vk[a] for a in range(dom.dim)]
            wrapped = False
            for a in range(dom.dim):
                h = st.vmax[a]
                if y[a] < h or y[a] >= shape[a] - h:
                    if per[2 * a]:
                        wrapped = True
            if not wrapped and not fluid[tuple(y)]:
                count += 1
    assert total == count


def test_resolve_sequential_matches_loop():
    rng = np.random.default_rng(0)
    for trial in range(50):
        n, size = 30, 40
        st = rng.integers(0, size, n)
        ld = rng.integers(0, size, n)
        sign = rng.choice([-1.0, 1.0], n)
        rhs = rng.random(n)
        f = rng.random(size)
        seq = f.copy()
        for i in range(n):
            seq[st[i]] = sign[i] * seq[ld[i]] + rhs[i]
        s2, l2, a2, r2 = _resolve_sequential(st, ld, sign, rhs)
        par = f.copy()
        gathered = a2 * f[l2] + r2
        par[s2] = gathered
        assert np.allclose(par, seq, rtol=0, atol=1e-12)
        assert np.unique(s2).size == s2.size


def test_unflat_inverts_flat():
    """_unflat recovers (k, canonical coordinates) of _flat indices in both arrangements."""
    from pylbm_amd.boundary import _flat, _unflat
    from pylbm_amd.runtime import LBM_F32, LBM_F64
    from pylbm_amd.storage import make_layout
    rng = np.random.default_rng(3)
    for shape, ns in (((7, 6, 9), 19), ((11, 13), 9), ((17,), 3)):
        for inter in (False, True):
            for st in (LBM_F32, LBM_F64):
                L = make_layout(ns, shape, (1,) * len(shape), st, interleaved=inter)
                rows = np.stack([rng.integers(0, ns, 200)] + [rng.integers(0, n, 200) for n in shape])
                k, o, m, i = _unflat(_flat(rows, L), L)
                canon = [np.zeros(200, np.int64)] * (3 - len(shape)) + [rows[1 + d] for d in range(len(shape))]
                assert np.array_equal(k, rows[0])
                assert np.array_equal(o, canon[0]) and np.array_equal(m, canon[1]) and np.array_equal(i, canon[2])


def test_flat_indices_both_arrangements():
    """Flat element indices of link rows and their population, for planes and
    interleaved rows, against a direct numpy index computation."""
    from pylbm_amd.boundary import _flat, _population
    from pylbm_amd.runtime import LBM_F32
    from pylbm_amd.storage import make_layout, layout_elements
    rng = np.random.default_rng(3)
    shape = (7, 6, 9)
    for inter in (False, True):
        L = make_layout(19, shape, (1, 1, 1), LBM_F32, interleaved=inter)
        rows = np.stack([rng.integers(0, 19, 200)] + [rng.integers(0, s, 200) for s in shape])
        idx = _flat(rows, L)
        if inter:
            want = ((rows[1] * L.n[1] + rows[2]) * 19 + rows[0]) * L.pitch + L.lead + rows[3]
        else:
            want = ((rows[0] * L.n[0] + rows[1]) * L.n[1] + rows[2]) * L.pitch + L.lead + rows[3]
        assert np.array_equal(idx, want)
        assert np.array_equal(_population(idx, L), rows[0])
        assert idx.max() < layout_elements(L)


class _HostLayout:
    """Layout whose flat indices address an oracle array (ns, *shape_halo) directly."""

    storage = 0  # LBM_F64: no deviation shift

    def __init__(self, ns, shape):
        n = [1] * (3 - len(shape)) + list(shape)
        self.ns, self.n = ns, n
        self.pitch = self.row_stride = n[2]
        self.lead = 0
        self.pop_stride = n[0] * n[1] * n[2]


def _emulated_step(o, prod_methods, dom):
    """The device's sequence on the host: ghosts poisoned, periodic-face slabs
    filled (first ghost axis periodic, populations entering through every ghost
    axis), the unwritten-ghost copy list, then the link lists."""
    from pylbm_amd.boundary import unwritten_ghost_reads
    from oracle.boundary import apply_links
    F = o.F.copy()
    ghost = np.ones(F.shape[1:], dtype=bool)
    ghost[o._interior()[1:]] = False
    F[:, ghost] = np.nan
    shape = F.shape[1:]
    h = o.h
    labels = dom.box_label
    for idx in zip(*np.nonzero(ghost)):
        first = next(b for b in range(o.dim) if idx[b] < h[b] or idx[b] >= shape[b] - h[b])
        side = 0 if idx[first] < h[first] else 1
        if labels[2 * first + side] != -1:
            continue
        for k, vk in enumerate(o.v):
            x = [idx[b] + vk[b] for b in range(o.dim)]
            if all(h[b] <= x[b] < shape[b] - h[b] for b in range(o.dim)):
                w = [int(idx[b]) + int(shape[b] - 2 * h[b]) * (int(idx[b] < h[b]) - int(idx[b] >= shape[b] - h[b]))
                     for b in range(o.dim)]
                F[(k,) + tuple(idx)] = F[(k,) + tuple(w)]
    lay = _HostLayout(o.ns, shape)
    st, ld = unwritten_ghost_reads(dom, dom.stencil, prod_methods, lay)
    flat = F.reshape(-1)
    flat[st] = flat[ld]
    for meth in o.methods:
        apply_links(F, meth)
    return F, st.size


def _neumann_periodic():
    """D2Q9 with NeumannX on both x faces and periodic y: the corner links load ghost
    slots of the periodic y slabs that the entering-population fill does not write."""
    d = cases.karman_d2q9(24, 12, radius=2.0, center=(8.0, 6.0))
    d["box"] = {"x": [0.0, 24.0], "y": [0.0, 12.0], "label": [2, 2, -1, -1]}
    d["elements"] = []
    d["boundary_conditions"] = {2: {"method": {0: bc.NeumannX}}}
    return d


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_restricted_ghost_fill_equals_full_wrap(name):
    from oracle.lbm import Simulation as O
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3": lambda: cases.d1q3_advection(10, walls=False),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    for _ in range(2):
        o.one_time_step()
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    F_dev, ncopies = _emulated_step(o, pm, dom)
    F_ref = o.F.copy()
    o.ghost_fill(F_ref)
    from oracle.boundary import apply_links
    for meth in o.methods:
        apply_links(F_ref, meth)
    a, b = np.empty_like(F_ref), np.empty_like(F_ref)
    o.pull_step(F_dev, a)
    o.pull_step(F_ref, b)
    inner = o._interior()
    assert not np.isnan(a[inner]).any()
    assert np.array_equal(a[inner], b[inner])
    if name == "spheres":
        assert ncopies > 0


def test_merge_flags_synthetic():
    """Dependencies that must split a group, the group size cap, two-phase lists."""
    from pylbm_amd.boundary import MAX_LINK_GROUP, merge_flags
    a = (np.array([0, 1]), np.array([10, 11]), False)
    assert merge_flags([a, (np.array([2]), np.array([12]), False)]) == [0, 1]          # independent
    assert merge_flags([a, (np.array([3]), np.array([0]), False)]) == [0, 0]           # loads a's store
    assert merge_flags([a, (np.array([11]), np.array([13]), False)]) == [0, 0]         # stores a's load
    assert merge_flags([a, (np.array([1]), np.array([14]), False)]) == [0, 0]          # stores a's store
    two = (np.array([20]), np.array([21]), True)
    assert merge_flags([a, two]) == [0, 0] and merge_flags([two, a]) == [0, 0]
    # a dependency on an earlier member of the group, not only the previous list
    b = (np.array([2]), np.array([12]), False)
    assert merge_flags([a, b, (np.array([4]), np.array([1]), False)]) == [0, 1, 0]
    many = [(np.array([100 + i]), np.array([200 + i]), False) for i in range(MAX_LINK_GROUP + 2)]
    assert merge_flags(many) == [0] + [1] * (MAX_LINK_GROUP - 1) + [0, 1]


def _apply(F, lst):
    st, ld, sign, rhs = lst[:4]
    flat = F.reshape(-1)
    flat[st] = sign * flat[ld] + rhs  # gather before scatter: one-phase and folded lists alike


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_merged_link_groups_equal_sequential(name):
    """The step's lists (ghost copy list first) applied one after another equal the
    groups merge_flags forms, each applied as one gather-then-scatter, bit for bit."""
    from oracle.lbm import Simulation as O
    from pylbm_amd.boundary import host_lists, merge_flags, unwritten_ghost_reads
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    lists = []
    st, ld = unwritten_ghost_reads(dom, dom.stencil, pm, lay)
    if st.size:
        lists.append(host_lists("ghost_copy", lay, None, flat=(st, ld)))
    rng = np.random.default_rng(1)
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        lists.append(host_lists(m, lay, None))
    flags = merge_flags([(x[0], x[1], x[4]) for x in lists])
    F0 = rng.normal(size=o.F.shape)
    seq = F0.copy()
    for lst in lists:
        _apply(seq, lst)
    grp = F0.copy()
    j = 0
    while j < len(lists):
        k = j + 1
        while k < len(lists) and flags[k]:
            k += 1
        snap = grp.reshape(-1).copy()
        vals = [lst[2] * snap[lst[1]] + lst[3] for lst in lists[j:k]]
        for lst, v in zip(lists[j:k], vals):
            grp.reshape(-1)[lst[0]] = v
        j = k
    assert np.array_equal(seq, grp)
    if name in ("karman", "vectorial"):
        assert sum(flags) > 0, flags


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_resolve_step_equals_sequential_lists(name):
    """The step's lists (ghost copy list first) applied one after another to a random
    FP64 array leave, in every slot they store, exactly sign * f0[origin] + rhs
    from resolve_step, for every chain that applied at most one rhs."""
    from oracle.lbm import Simulation as O
    from pylbm_amd.boundary import host_lists, resolve_step, unwritten_ghost_reads
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    lists = []
    st, ld = unwritten_ghost_reads(dom, dom.stencil, pm, lay)
    if st.size:
        lists.append(host_lists("ghost_copy", lay, None, flat=(st, ld)))
    rng = np.random.default_rng(2)
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        lists.append(host_lists(m, lay, None))
    F0 = rng.normal(size=o.F.shape)
    seq = F0.copy()
    for lst in lists:
        _apply(seq, lst)
    slot, org, sg, rh, has, bad, _, _ = resolve_step([(x[0], x[1], x[2], x[3] if np.any(x[3] != 0) else None)
                                                for x in lists])
    flat0, flat1 = F0.reshape(-1), seq.reshape(-1)
    val = sg * flat0[org] + np.where(has, rh, 0.0)
    assert slot.size > 0
    assert np.array_equal(val[~bad], flat1[slot][~bad])
    # slots no list stores are untouched
    untouched = np.ones(flat0.size, bool)
    untouched[slot] = False
    assert np.array_equal(flat0[untouched], flat1[untouched])


_BZ = {
    "gap-bb": lambda: cases.poiseuille_d2q9(ny=5, y_low=1.3, y_high=2.4, method=bc.BouzidiBounceBack),
    # loads of slots stored later in the loop: the old values, as two-phase gives
    "gap-abb": lambda: cases.poiseuille_d2q9(ny=5, y_low=1.3, y_high=2.4, method=bc.BouzidiAntiBounceBack),
    "karman-bz": lambda: cases.with_bouzidi(cases.karman_d2q9(48, 24, radius=4.0, center=(12.0, 12.0))),
    "spheres-bz": lambda: cases.with_bouzidi(cases.sphere_packing_d3q19(16, 12, 12, seed=3)),
}


@pytest.mark.parametrize("name", sorted(_BZ))
def test_bouzidi_lists_equal_sequential_loop(name):
    """The device form of an interpolated list (boundary.bouzidi_lists: last store
    wins, gather before scatter when loads hit stores, sorted otherwise), applied
    to a flat array of the device layout, equals the oracle's loop in link order
    (the copy of f for bounce-back, f itself for anti-bounce-back), bit for bit."""
    from oracle.boundary import _apply_bouzidi
    from pylbm_amd.boundary import _flat, bouzidi_lists
    from pylbm_amd.runtime import LBM_F64
    from pylbm_amd.storage import layout_elements, make_layout
    d = _BZ[name]()
    dom, pm = _product(d)
    _, om = _oracle(d)
    ns = int(dom.stencil.nv_ptr[-1])
    shape = tuple(dom.shape_halo)
    L = make_layout(ns, shape, list(dom.stencil.vmax), LBM_F64)
    rng = np.random.default_rng(11)
    F = rng.standard_normal((ns,) + shape)
    every = np.array(np.unravel_index(np.arange(F.size), F.shape))
    pos = _flat(every, L)
    flat = np.full(layout_elements(L), np.nan)
    flat[pos] = F.ravel()
    checked = 0
    for p, o in zip(pm, om):
        if not p.kind.bouzidi:
            continue
        assert o.kind == p.name
        rhs = rng.standard_normal(p.ncond)
        p.rhs, o.rhs = rhs.copy(), rhs.copy()
        st, l0, l1, ca, cb, r, two, index, shift = bouzidi_lists(p, L, np.zeros(ns))
        vals = ca * flat[l0] + cb * flat[l1] + r      # every load before any store
        flat[st] = vals
        _apply_bouzidi(F, o)
        assert np.array_equal(flat[pos], F.ravel())
        if name.startswith("gap"):
            assert two
        checked += 1
    assert checked


@pytest.mark.parametrize("gap", [(0.9, 1.8), (0.7, 1.6), (1.3, 2.4)])
def test_bouzidi_anti_bounce_back_levels_equal_sequential_loop(gap):
    """Bouzidi anti-bounce-back reads f in loop order: in a one-row gap whose upper
    wall is the nearer one, the upper links (populations pointing down, later in the
    velocity order) load slots of the lower solid row that the lower links stored
    earlier.  One list cannot carry those values (bouzidi_lists raises); the
    dependency levels of bouzidi_levels, each applied one-phase in order, equal the
    oracle's in-place loop bit for bit, and every level is free of internal
    conflicts (no slot stored twice, no load of a slot the level stores)."""
    from oracle.boundary import _apply_bouzidi
    from pylbm_amd.boundary import _flat, bouzidi_levels, bouzidi_lists
    from pylbm_amd.runtime import LBM_F64
    from pylbm_amd.storage import layout_elements, make_layout
    d = cases.poiseuille_d2q9(ny=5, y_low=gap[0], y_high=gap[1], method=bc.BouzidiAntiBounceBack)
    dom, pm = _product(d)
    _, om = _oracle(d)
    ns = int(dom.stencil.nv_ptr[-1])
    shape = tuple(dom.shape_halo)
    L = make_layout(ns, shape, list(dom.stencil.vmax), LBM_F64)
    rng = np.random.default_rng(5)
    F = rng.standard_normal((ns,) + shape)
    every = np.array(np.unravel_index(np.arange(F.size), F.shape))
    pos = _flat(every, L)
    flat = np.full(layout_elements(L), np.nan)
    flat[pos] = F.ravel()
    p, o = pm[0], om[0]
    rhs = rng.standard_normal(p.ncond)
    p.rhs, o.rhs = rhs.copy(), rhs.copy()
    levels = bouzidi_levels(p, L, np.zeros(ns))
    if gap == (1.3, 2.4):
        assert len(levels) == 1       # the nearer wall is the lower one: one list
    else:
        assert len(levels) > 1
        with pytest.raises(NotImplementedError):
            bouzidi_lists(p, L, np.zeros(ns))
    covered = np.zeros(p.ncond, dtype=int)
    for st, l0, l1, ca, cb, r, two, index, shift in levels:
        if len(levels) > 1:
            assert not two and np.unique(st).size == st.size
            assert not np.isin(np.concatenate([l0, l1]), st).any()
        covered[index] += 1
        vals = ca * flat[l0] + cb * flat[l1] + r
        flat[st] = vals
    if len(levels) > 1:
        assert np.all(covered == 1)   # every link runs exactly once
    _apply_bouzidi(F, o)
    assert np.array_equal(flat[pos], F.ravel())


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_merged_link_groups_equal_sequential(name):
    """The step's lists (ghost copy list first) applied one after another equal the
    groups merge_flags forms, each applied as one gather-then-scatter, bit for bit."""
    from oracle.lbm import Simulation as O
    from pylbm_amd.boundary import host_lists, merge_flags, unwritten_ghost_reads
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    lists = []
    st, ld = unwritten_ghost_reads(dom, dom.stencil, pm, lay)
    if st.size:
        lists.append(host_lists("ghost_copy", lay, None, flat=(st, ld)))
    rng = np.random.default_rng(1)
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        lists.append(host_lists(m, lay, None))
    flags = merge_flags([(x[0], x[1], x[4]) for x in lists])
    F0 = rng.normal(size=o.F.shape)
    seq = F0.copy()
    for lst in lists:
        _apply(seq, lst)
    grp = F0.copy()
    j = 0
    while j < len(lists):
        k = j + 1
        while k < len(lists) and flags[k]:
            k += 1
        snap = grp.reshape(-1).copy()
        vals = [lst[2] * snap[lst[1]] + lst[3] for lst in lists[j:k]]
        for lst, v in zip(lists[j:k], vals):
            grp.reshape(-1)[lst[0]] = v
        j = k
    assert np.array_equal(seq, grp)
    if name in ("karman", "vectorial"):
        assert sum(flags) > 0, flags


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_resolve_step_equals_sequential_lists(name):
    """The step's lists (ghost copy list first) applied one after another to a random
    FP64 array leave, in every slot they store, exactly sign * f0[origin] + rhs
    from resolve_step, for every chain that applied at most one rhs."""
    from oracle.lbm import Simulation as O
    from pylbm_amd.boundary import host_lists, resolve_step, unwritten_ghost_reads
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    lists = []
    st, ld = unwritten_ghost_reads(dom, dom.stencil, pm, lay)
    if st.size:
        lists.append(host_lists("ghost_copy", lay, None, flat=(st, ld)))
    rng = np.random.default_rng(2)
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        lists.append(host_lists(m, lay, None))
    F0 = rng.normal(size=o.F.shape)
    seq = F0.copy()
    for lst in lists:
        _apply(seq, lst)
    slot, org, sg, rh, has, bad, _, _ = resolve_step([(x[0], x[1], x[2], x[3] if np.any(x[3] != 0) else None)
                                                for x in lists])
    flat0, flat1 = F0.reshape(-1), seq.reshape(-1)
    val = sg * flat0[org] + np.where(has, rh, 0.0)
    assert slot.size > 0
    assert np.array_equal(val[~bad], flat1[slot][~bad])
    # slots no list stores are untouched
    untouched = np.ones(flat0.size, bool)
    untouched[slot] = False
    assert np.array_equal(flat0[untouched], flat1[untouched])


_BZ = {
    "gap-bb": lambda: cases.poiseuille_d2q9(ny=5, y_low=1.3, y_high=2.4, method=bc.BouzidiBounceBack),
    # loads of slots stored later in the loop: the old values, as two-phase gives
    "gap-abb": lambda: cases.poiseuille_d2q9(ny=5, y_low=1.3, y_high=2.4, method=bc.BouzidiAntiBounceBack),
    "karman-bz": lambda: cases.with_bouzidi(cases.karman_d2q9(48, 24, radius=4.0, center=(12.0, 12.0))),
    "spheres-bz": lambda: cases.with_bouzidi(cases.sphere_packing_d3q19(16, 12, 12, seed=3)),
}


@pytest.mark.parametrize("name", sorted(_BZ))
def test_bouzidi_lists_equal_sequential_loop(name):
    """The device form of an interpolated list (boundary.bouzidi_lists: last store
    wins, gather before scatter when loads hit stores, sorted otherwise), applied
    to a flat array of the device layout, equals the oracle's loop in link order
    (the copy of f for bounce-back, f itself for anti-bounce-back), bit for bit."""
    from oracle.boundary import _apply_bouzidi
    from pylbm_amd.boundary import _flat, bouzidi_lists
    from pylbm_amd.runtime import LBM_F64
    from pylbm_amd.storage import layout_elements, make_layout
    d = _BZ[name]()
    dom, pm = _product(d)
    _, om = _oracle(d)
    ns = int(dom.stencil.nv_ptr[-1])
    shape = tuple(dom.shape_halo)
    L = make_layout(ns, shape, list(dom.stencil.vmax), LBM_F64)
    rng = np.random.default_rng(11)
    F = rng.standard_normal((ns,) + shape)
    every = np.array(np.unravel_index(np.arange(F.size), F.shape))
    pos = _flat(every, L)
    flat = np.full(layout_elements(L), np.nan)
    flat[pos] = F.ravel()
    checked = 0
    for p, o in zip(pm, om):
        if not p.kind.bouzidi:
            continue
        assert o.kind == p.name
        rhs = rng.standard_normal(p.ncond)
        p.rhs, o.rhs = rhs.copy(), rhs.copy()
        st, l0, l1, ca, cb, r, two, index, shift = bouzidi_lists(p, L, np.zeros(ns))
        vals = ca * flat[l0] + cb * flat[l1] + r      # every load before any store
        flat[st] = vals
        _apply_bouzidi(F, o)
        assert np.array_equal(flat[pos], F.ravel())
        if name.startswith("gap"):
            assert two
        checked += 1
    assert checked


def test_bouzidi_anti_bounce_back_refuses_earlier_store_loads():
    """Bouzidi anti-bounce-back reads f in loop order: in a one-row gap whose upper
    wall is the nearer one, the upper links (populations pointing down, later in the
    velocity order) load slots of the lower solid row that the lower links stored
    earlier, which the two-phase form cannot give."""
    from pylbm_amd.boundary import bouzidi_lists
    from pylbm_amd.runtime import LBM_F64
    from pylbm_amd.storage import make_layout
    d = cases.poiseuille_d2q9(ny=5, y_low=0.9, y_high=1.8, method=bc.BouzidiAntiBounceBack)
    dom, pm = _product(d)
    L = make_layout(int(dom.stencil.nv_ptr[-1]), tuple(dom.shape_halo), list(dom.stencil.vmax), LBM_F64)
    with pytest.raises(NotImplementedError):
        bouzidi_lists(pm[0], L, np.zeros(9))


@pytest.mark.parametrize("name", ["spheres", "lid", "d1q3_walls", "karman", "vectorial", "neumann_periodic"])
def test_wrap_mode_pulls_equal_halo_pulls(name):
    """The step's pulled values with every list folded and the pulls read through the
    periodic wrap (the wrap mode of Simulation: table sources in ghost slots moved to
    the wrapped interior slot) equal, for every interior cell and population, the
    pulls of pylbm's sequence: wrap of the whole halo (the oracle's ghost fill,
    pylbm/storage.py:306-367), every list in order, then f[k, x - v_k]."""
    from oracle.lbm import Simulation as O
    from pylbm_amd.boundary import _unflat, host_lists, resolve_step, unwritten_ghost_reads
    d = {"spheres": lambda: cases.sphere_packing_d3q19(20, 14, 12, seed=3, solid_fraction=0.25),
         "lid": lambda: cases.lid_cavity_d2q9(12),
         "d1q3_walls": lambda: cases.d1q3_advection(10, walls=True),
         "karman": lambda: cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0)),
         "vectorial": lambda: cases.d2q5q9_vectorial(12, walls=True),
         "neumann_periodic": _neumann_periodic}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    lists = []
    st, ld = unwritten_ghost_reads(dom, dom.stencil, pm, lay)
    if st.size:
        lists.append(host_lists("ghost_copy", lay, None, flat=(st, ld)))
    rng = np.random.default_rng(4)
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        lists.append(host_lists(m, lay, None))
    F0 = rng.normal(size=o.F.shape)
    # pylbm: wrap every halo slot, the lists in order, then the pulls
    seq = F0.copy()
    o.ghost_fill(seq)
    for lst in lists:
        _apply(seq, lst)
    # wrap mode: folded values where a list stores the pulled slot, else the wrapped read
    slot, org, sg, rh, has, bad, _, _ = resolve_step([(x[0], x[1], x[2], x[3] if np.any(x[3] != 0) else None)
                                                for x in lists])
    n = np.array(lay.n)
    h = np.array([0] * (3 - o.dim) + [int(v) for v in o.h])
    nin = n - 2 * h

    def wrap(c):
        return [np.where(c[a] < h[a], c[a] + nin[a], np.where(c[a] >= n[a] - h[a], c[a] - nin[a], c[a]))
                for a in range(3)]

    k, oo, mm, ii = _unflat(org, lay)
    w = wrap([np.asarray(oo), np.asarray(mm), np.asarray(ii)])
    org_w = k * lay.pop_stride + (w[0] * n[1] + w[1]) * lay.row_stride + w[2]
    folded = dict(zip(slot.tolist(), (sg * F0.reshape(-1)[org_w] + np.where(has, rh, 0.0)).tolist()))
    vel = np.asarray(o.v)
    v3 = np.zeros((o.ns, 3), dtype=np.int64)
    v3[:, 3 - o.dim:] = vel
    grid = np.stack(np.meshgrid(*[np.arange(h[a], n[a] - h[a]) for a in range(3)], indexing="ij"), axis=0)
    grid = grid.reshape(3, -1)
    flat_seq = seq.reshape(-1)
    for q in range(o.ns):
        src = grid - v3[q][:, None]
        idx = q * lay.pop_stride + (src[0] * n[1] + src[1]) * lay.row_stride + src[2]
        want = flat_seq[idx]
        ws = wrap(src)
        idx_w = q * lay.pop_stride + (ws[0] * n[1] + ws[1]) * lay.row_stride + ws[2]
        got = F0.reshape(-1)[idx_w].copy()
        for j, s in enumerate(idx.tolist()):
            if s in folded:
                got[j] = folded[s]
        assert np.array_equal(got, want), (name, q, int(np.sum(got != want)))


def test_resolve_step_interpolated_lists():
    """resolve_step with an interpolated (Bouzidi) list between one-value lists: every
    slot not flagged equals the sequential application of the lists in order
    (copies and sign flips resolved into the sources, ca f0 + cb f1 + rhs with
    contraction off, as the link kernels compute it), bit for bit; a later list that
    reads an interpolated value, and an interpolation of a value with an rhs, are
    flagged."""
    from pylbm_amd.boundary import resolve_step
    rng = np.random.default_rng(21)
    f = rng.standard_normal(400)
    slots = rng.permutation(400)
    sa, la = slots[:40], slots[200:240]               # copies, sign -1 on half of them
    sign_a = np.where(np.arange(40) % 2, -1.0, 1.0)
    sb = slots[40:80]                                   # interpolated stores
    l0 = np.concatenate([sa[:10], slots[240:270]])      # ten loads of copied values
    l1 = np.concatenate([slots[270:300], sa[10:20]])
    ca, cb = rng.uniform(-1, 1, 40), rng.uniform(0, 1, 40)
    rb = rng.standard_normal(40)
    sc, lc = slots[80:100], np.concatenate([sb[:5], slots[300:315]])   # five valued reads of interpolations
    rc = rng.standard_normal(20)
    sf, lf = slots[120:130], sb[5:15]                                   # signed copies of interpolations
    sign_f = np.where(np.arange(10) % 2, -1.0, 1.0)
    sd, ld = slots[100:110], slots[315:325]              # a valued list read by an interpolation
    rd = rng.standard_normal(10)
    se = slots[110:120]
    l0e, l1e = np.concatenate([sd[:3], slots[325:332]]), slots[332:342]
    cae, cbe = rng.uniform(-1, 1, 10), rng.uniform(0, 1, 10)
    lists = [(sa, la, sign_a, None), (sb, l0, ca, rb, l1, cb), (sc, lc, np.ones(20), rc),
             (sd, ld, np.ones(10), rd), (se, l0e, cae, None, l1e, cbe), (sf, lf, sign_f, None)]
    g = f.copy()
    g[sa] = sign_a * g[la]
    v = ca * g[l0] + cb * g[l1]
    g[sb] = v + rb
    g[sc] = g[lc] + rc
    g[sd] = g[ld] + rd
    g[se] = cae * g[l0e] + cbe * g[l1e]
    g[sf] = sign_f * g[lf]
    slot, org, sg, rh, has, bad, org2, c2 = resolve_step(lists)
    for i, s_ in enumerate(slot):
        if bad[i]:
            continue
        val = sg[i] * f[org[i]]
        if org2[i] >= 0:
            val = val + c2[i] * f[org2[i]]
        if has[i]:
            val = val + rh[i]
        assert val == g[s_], (s_, val, g[s_])
    flagged = set(slot[bad].tolist())
    assert set(sc[:5].tolist()) <= flagged and set(se[:3].tolist()) <= flagged
    assert not (set(sb.tolist()) & flagged) and not (set(sf.tolist()) & flagged)
    assert np.all(org2[np.isin(slot, sf)] >= 0)
    assert np.all(org2[np.isin(slot, sb)] >= 0)


@pytest.mark.parametrize("name", ["karman_bouzidi", "poiseuille", "lid_demo"])
def test_wrap_mode_pulls_equal_halo_pulls_interpolated(name):
    """As test_wrap_mode_pulls_equal_halo_pulls, with interpolated (Bouzidi) lists in
    the step: their device form (boundary.bouzidi_lists) applied in order between the
    other lists after the wrap of the whole halo, against the step's fold
    (resolve_step with two-source entries, ghost sources moved to the wrapped
    interior slot), bit for bit."""
    from oracle.lbm import Simulation as O
    from pylbm_amd import demos
    from pylbm_amd.boundary import _unflat, bouzidi_lists, host_lists, resolve_step
    d = {"karman_bouzidi": lambda: cases.with_bouzidi(cases.karman_d2q9(40, 20, radius=4.0, center=(10.0, 10.0))),
         "poiseuille": lambda: demos.poiseuille(1.0 / 16)[0],
         "lid_demo": lambda: demos.lid_driven_cavity(1.0 / 16)[0]}[name]()
    o = O(d)
    dom, pm = _product(d)
    for m in pm:
        m.fix_iload()
    lay = _HostLayout(o.ns, o.F.shape[1:])
    rng = np.random.default_rng(5)
    lists, seq_ops = [], []
    for m in pm:
        m.rhs = rng.normal(size=m.rhs.shape)
        if m.kind.bouzidi:
            st, l0, l1, ca, cb, rhs, two, index, shift = bouzidi_lists(m, lay, np.zeros(o.ns))
            assert not two
            lists.append((st, l0, ca, rhs, l1, cb))
            seq_ops.append(("i", st, l0, l1, ca, cb, rhs))
        else:
            x = host_lists(m, lay, None)
            lists.append((x[0], x[1], x[2], x[3] if np.any(x[3] != 0) else None))
            seq_ops.append(("c", x[0], x[1], x[2], x[3]))
    F0 = rng.normal(size=o.F.shape)
    seq = F0.copy()
    o.ghost_fill(seq)
    flat = seq.reshape(-1)
    for op in seq_ops:
        if op[0] == "i":
            _, st, l0, l1, ca, cb, rhs = op
            flat[st] = ca * flat[l0] + cb * flat[l1] + rhs
        else:
            _, st, ld, sign, rhs = op
            flat[st] = sign * flat[ld] + rhs
    slot, org, sg, rh, has, bad, org2, c2 = resolve_step(lists)
    n = np.array(lay.n)
    h = np.array([0] * (3 - o.dim) + [int(v) for v in o.h])
    nin = n - 2 * h

    def wrap(c):
        return [np.where(c[a] < h[a], c[a] + nin[a], np.where(c[a] >= n[a] - h[a], c[a] - nin[a], c[a]))
                for a in range(3)]

    def wrapped(idx):
        k, oo, mm, ii = _unflat(idx, lay)
        w = wrap([np.asarray(oo), np.asarray(mm), np.asarray(ii)])
        return k * lay.pop_stride + (w[0] * n[1] + w[1]) * lay.row_stride + w[2]

    f0 = F0.reshape(-1)
    val = sg * f0[wrapped(org)]
    two_src = org2 >= 0
    val[two_src] = val[two_src] + c2[two_src] * f0[wrapped(org2[two_src])]
    val = np.where(has, val + rh, val)
    folded = dict(zip(slot.tolist(), val.tolist()))
    vel = np.asarray(o.v)
    v3 = np.zeros((o.ns, 3), dtype=np.int64)
    v3[:, 3 - o.dim:] = vel
    grid = np.stack(np.meshgrid(*[np.arange(h[a], n[a] - h[a]) for a in range(3)], indexing="ij"), axis=0)
    grid = grid.reshape(3, -1)
    pulled_bad = 0
    for q in range(o.ns):
        src = grid - v3[q][:, None]
        idx = q * lay.pop_stride + (src[0] * n[1] + src[1]) * lay.row_stride + src[2]
        want = flat[idx]
        ws = wrap(src)
        idx_w = q * lay.pop_stride + (ws[0] * n[1] + ws[1]) * lay.row_stride + ws[2]
        got = f0[idx_w].copy()
        bad_slots = set(slot[bad].tolist())
        for j, s in enumerate(idx.tolist()):
            if s in folded:
                got[j] = folded[s]
                pulled_bad += s in bad_slots
        assert np.array_equal(got, want), (name, q, int(np.sum(got != want)))
    assert pulled_bad == 0
