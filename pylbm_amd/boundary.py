"""Boundary link lists and their device form.

Host side mirrors pylbm/boundary.py:
  * ``Boundary``: labels ascending (periodic -1 / interface -2 skipped), one
    method instance per method class in order of first appearance, links of
    each method concatenated label -> scheme -> velocity -> C-order of the
    fluid cells, store slot (k, x_out = x - v_k)            boundary.py:23-160
  * ``set_iload`` per method (bounce-back family loads (ksym, x_out + v_k),
    Neumann loads (k, x_out + v_k) or one component of it) boundary.py:411-419, 777-890
  * ``prepare_rhs``: wall points x_out + (1 - d) v_k dx, user value function
    on moments, equilibrium, m2f (here the generated GPU kernels)
                                                            boundary.py:238-305
  * ``set_rhs``: BB feq_k - feq_ksym, ABB feq_k + feq_ksym   boundary.py:421-427, 637-643
  * ``fix_iload``: int32 (n, dim + 1) rows                   boundary.py:226-235

Device side: each method becomes flat uint32 element indices (store, load)
into the population array plus a compute-type rhs (omitted when zero), run
by ``lbm_links_apply``.  The reference loop is sequential; when a method's
loads hit its own stores (possible with Neumann on adjacent faces) the
read-after-write chains are folded on the host and the list runs in its
two-phase (gather all, then scatter) form, which reproduces the sequential
result exactly.
"""
import collections
import ctypes

import numpy as np

from . import bc as bcmod
from .runtime import DEVIATION_FORMATS, LBM_F64, LbmLinkSet

__all__ = ["Boundary", "BoundaryMethod"]


class BoundaryMethod:
    """Link list of one method class (pylbm/boundary.py:164-397)."""

    def __init__(self, kind, istore, ilabel, distance, stencil, value_bc, time_bc):
        self.kind = kind
        self.name = kind.name
        self.istore = istore                 # (dim + 1, n) before fix_iload
        self.ilabel = ilabel
        self.distance = distance
        self.stencil = stencil
        self.feq = np.zeros((int(stencil.nv_ptr[-1]), istore.shape[1]))
        self.rhs = np.zeros(istore.shape[1])
        self.value_bc = {k: value_bc[k] for k in np.unique(ilabel)}
        self.time_bc = {k: time_bc[k] for k in np.unique(ilabel)}
        self.iload = []
        self._fixed = False
        self.device = None
        self._tbc = []

    @property
    def ncond(self):
        return self.istore.shape[1] if not self._fixed else self.istore.shape[0]

    def set_iload(self):
        k = self.istore[0]
        v = self.stencil.get_all_velocities()
        if self.kind.bouzidi:
            # pylbm/boundary.py:506-535: a wall nearer than half a link (q < 0.5)
            # interpolates between the fluid cell and the one behind it (s = 2q), a
            # farther one between the bounced and the incoming population (s = 1/(2q))
            ksym = self.stencil.get_symmetric()[k][np.newaxis, :]
            vk = v[k].T
            near = np.asarray(self.distance) < 0.5
            x1 = self.istore[1:] + vk
            self.iload.append(np.concatenate([ksym, x1]))
            self.iload.append(np.where(near[np.newaxis, :], np.concatenate([ksym, x1 + vk]),
                                       np.concatenate([k[np.newaxis, :], x1])))
            dist = np.asarray(self.distance, dtype=np.float64)
            self.s = np.where(near, 2.0 * dist, 0.5 / np.where(near, 1.0, dist))
            return
        if self.kind.symmetric_load:
            ksym = self.stencil.get_symmetric()[k][np.newaxis, :]
            self.iload.append(np.concatenate([ksym, self.istore[1:] + v[k].T]))
        else:
            shift = v[k].T.copy()
            if self.kind.shift_axis is not None:
                keep = shift[self.kind.shift_axis].copy()
                shift[:] = 0
                shift[self.kind.shift_axis] = keep
            self.iload.append(np.concatenate([k[np.newaxis, :], self.istore[1:] + shift]))

    def fix_iload(self):
        self.iload = [np.ascontiguousarray(a.T, dtype=np.int32) for a in self.iload]
        self.istore = np.ascontiguousarray(self.istore.T, dtype=np.int32)
        self._fixed = True

    def _store_rows(self):
        return self.istore.T if self._fixed else self.istore

    def set_rhs(self):
        if self.kind.rhs_sign is None:
            return
        k = self._store_rows()[0]
        ksym = self.stencil.get_symmetric()[k]
        cols = np.arange(k.size)
        self.rhs[:] = self.feq[k, cols] + self.kind.rhs_sign * self.feq[ksym, cols]


class Boundary:
    """All link-list methods of a simulation (pylbm/boundary.py:50-160)."""

    def __init__(self, domain, dico):
        self.domain = domain
        stencil = domain.stencil
        dico_bound = dico.get("boundary_conditions", {}) or {}
        links = domain.links()
        store = collections.OrderedDict()
        labs, dists = {}, {}
        value_bc, time_bc = {}, {}
        shape = domain.shape_halo
        for label in domain.list_of_labels():
            if label in (-1, -2):
                continue
            if label not in dico_bound:
                raise KeyError("no boundary condition for label %s" % label)
            value_bc[label] = dico_bound[label].get("value", None)
            time_bc[label] = dico_bound[label].get("time_bc", False)
            for k, cls in dico_bound[label]["method"].items():
                if getattr(cls, "unsupported", False) or not issubclass(cls, bcmod.BoundaryMethodKind):
                    raise NotImplementedError("boundary method %s is not provided by the hip generator"
                                              % getattr(cls, "__name__", cls))
                for inumk, numk in enumerate(stencil.num(k)):
                    u = int(stencil.unum2index[numk])
                    key = (label, domain.sym_index(u))
                    if key not in links:
                        continue
                    pos, dist = links[key]
                    x = np.array(np.unravel_index(pos, shape))
                    x_out = x - np.asarray(stencil.uv[u])[:, None]
                    rows = np.concatenate([np.full((1, pos.size), stencil.nv_ptr[k] + inumk), x_out])
                    store.setdefault(cls, []).append(rows)
                    labs.setdefault(cls, []).append(np.full(pos.size, label))
                    dists.setdefault(cls, []).append(dist)
        self.methods = []
        for cls in store:
            m = BoundaryMethod(cls, np.concatenate(store[cls], axis=1).astype(np.int64),
                               np.concatenate(labs[cls]), np.concatenate(dists[cls]),
                               stencil, value_bc, time_bc)
            self.methods.append(m)


# --------------------------------------------------------------- device form
def _flat(rows, layout):
    """Element indices of (k, x...) rows (halo-inclusive reference axes) in the device layout."""
    dim = rows.shape[0] - 1
    c = [np.zeros(rows.shape[1], dtype=np.int64)] * (3 - dim) + [rows[1 + d].astype(np.int64) for d in range(dim)]
    idx = (rows[0].astype(np.int64) * layout.pop_stride
           + (c[0] * layout.n[1] + c[1]) * layout.row_stride + layout.lead + c[2])
    if idx.size and (idx.min() < 0 or idx.max() >= (1 << 32)):
        raise ValueError("link index out of the 32-bit range")
    return idx


def _unflat(idx, layout):
    """(k, o, m, i) of flat element indices: population and canonical halo-inclusive
    coordinates (outer, mid, inner), the inverse of _flat."""
    idx = np.asarray(idx, dtype=np.int64)
    if layout.pop_stride >= layout.row_stride:      # population planes (one row: equal)
        k = idx // layout.pop_stride
        r = idx % layout.pop_stride - layout.lead
        row, i = r // layout.row_stride, r % layout.row_stride
    else:                                           # rows of all populations interleaved
        r = idx - layout.lead
        row, rem = r // layout.row_stride, r % layout.row_stride
        k, i = rem // layout.pop_stride, rem % layout.pop_stride
    return k, row // layout.n[1], row % layout.n[1], i


def _population(idx, layout):
    """Population index of flat element indices (either arrangement of the layout)."""
    if layout.pop_stride > layout.row_stride:
        return (idx // layout.pop_stride).astype(np.int64)
    return ((idx // layout.pop_stride) % layout.ns).astype(np.int64)


def _resolve_sequential(store, load, sign, rhs):
    """Fold read-after-write chains of a sequential loop f[s_i] = a_i f[l_i] + r_i.

    Returns (store, load, sign, rhs) such that gathering every load before any
    store gives the sequential result; earlier duplicate stores are dropped."""
    written = {}
    out_l = np.empty_like(load)
    out_a = np.empty(load.size, dtype=np.float64)
    out_r = np.empty(load.size, dtype=np.float64)
    for i in range(store.size):
        l = int(load[i])
        a, r = float(sign[i]), float(rhs[i])
        if l in written:
            aw, lw, rw = written[l]
            a, l, r = a * aw, lw, a * rw + r
        written[int(store[i])] = (a, l, r)
        out_l[i], out_a[i], out_r[i] = l, a, r
    last = {}
    for i in range(store.size):
        last[int(store[i])] = i
    keep = np.array(sorted(last.values()), dtype=np.int64)
    return store[keep], out_l[keep], out_a[keep], out_r[keep]


MAX_LINK_GROUP = 8  # lists per launch of lbm_links_apply_sets


def resolve_step(lists):
    """The step's link lists, in order, as one value per stored slot.

    lists: [(store, load, sign, rhs or None)] or, for an interpolated (Bouzidi)
    list, [(store, load0, ca, rhs or None, load1, cb)] (flat indices; rhs None:
    the list adds nothing).  Each list reads the values the lists before it left
    (its loads are looked up among their stores) and the last store of a slot
    wins.  Returns (slot, origin, sign, rhs, has_rhs, two_rhs, origin2, c2), sorted
    by slot: the final value of slot is sign * f[origin] (+ c2 * f[origin2] where
    origin2 >= 0) + rhs (rhs only if has_rhs) with origins slots no list stores;
    two_rhs marks values that this form does not reproduce bit for bit: chains that
    applied two rhs (their value needs two roundings), interpolations of an
    interpolated value or of a value that carries an rhs, and an rhs added to an
    interpolated value (a signed copy of one is exact)."""
    slot = np.zeros(0, np.int64)
    org = np.zeros(0, np.int64)
    sg = np.zeros(0)
    rh = np.zeros(0)
    has = np.zeros(0, bool)
    bad = np.zeros(0, bool)
    org2 = np.zeros(0, np.int64)
    cc2 = np.zeros(0)

    def lookup(ld):
        if not slot.size:
            return np.zeros(ld.size, bool), np.zeros(0, np.int64)
        pos = np.minimum(np.searchsorted(slot, ld), slot.size - 1)
        hit = slot[pos] == ld
        return hit, pos[hit]

    for entry in lists:
        st, ld, s1, r1 = entry[:4]
        st = np.asarray(st, np.int64)
        ld = np.asarray(ld, np.int64)
        s1 = np.asarray(s1, np.float64)
        h1 = np.full(st.size, r1 is not None)
        r1 = np.asarray(r1, np.float64) if r1 is not None else np.zeros(st.size)
        o2, s2, r2, h2, b2 = ld.copy(), s1.copy(), r1.copy(), h1.copy(), np.zeros(st.size, bool)
        q2, c2 = np.full(st.size, -1, np.int64), np.zeros(st.size)
        interp = len(entry) > 4 and entry[4] is not None
        hit, p = lookup(ld)
        if interp:
            # ca f[load0] + cb f[load1] + rhs: a load may take a signed copy (exact)
            q2 = np.asarray(entry[4], np.int64).copy()
            c2 = np.asarray(entry[5], np.float64).copy()
            o2[hit] = org[p]
            s2[hit] = s1[hit] * sg[p]
            b2[hit] = has[p] | bad[p] | (org2[p] >= 0)
            hit1, p1 = lookup(q2)
            q2[hit1] = org[p1]
            c2[hit1] = c2[hit1] * sg[p1]
            b2[hit1] |= has[p1] | bad[p1] | (org2[p1] >= 0)
        elif p.size:
            o2[hit] = org[p]
            s2[hit] = s1[hit] * sg[p]
            # s1 (sg f + rh) + r1, exact when at most one of rh, r1 applies
            r2[hit] = np.where(has[p], s1[hit] * rh[p], 0.0) + np.where(h1[hit], r1[hit], 0.0)
            h2[hit] = h1[hit] | has[p]
            # a signed copy of an interpolated value is the interpolation with every
            # coefficient and the rhs signed (negation is exact); one with an rhs of
            # its own would round twice
            ih = org2[p] >= 0
            q2[hit] = np.where(ih, org2[p], -1)
            c2[hit] = np.where(ih, s1[hit] * cc2[p], 0.0)
            b2[hit] = bad[p] | (h1[hit] & has[p]) | (ih & h1[hit])
        # the stores of this list replace earlier stores of the same slots
        cat = [np.concatenate(a) for a in ((slot, st), (org, o2), (sg, s2), (rh, r2), (has, h2), (bad, b2),
                                           (org2, q2), (cc2, c2))]
        uniq, first = np.unique(cat[0][::-1], return_index=True)
        keep = cat[0].size - 1 - first
        slot, org, sg, rh, has, bad, org2, cc2 = (uniq,) + tuple(c[keep] for c in cat[1:])
    return slot, org, sg, rh, has, bad, org2, cc2


def merge_flags(lists):
    """merge_prev flags of the link lists of one step, in application order.

    lists: [(store, load, two_phase)] with flat element indices.  List j joins
    the group of the lists before it (flag 1) when it is one-phase, the group
    is one-phase and has fewer than MAX_LINK_GROUP lists, no list of the group
    stores a slot that j loads or stores, and j stores no slot the group loads.
    Then every link of the group reads the value the sequential order
    (pylbm/simulation.py:385-390) gives it and every slot receives its one
    store, so the group may run as one launch.  The grouping is the one
    lbm_links_apply_sets applies."""
    flags = []
    gst = gld = None
    gsize = 0
    gtwo = True
    for st, ld, two in lists:
        st = np.asarray(st)
        ld = np.asarray(ld)
        ok = (not two and not gtwo and 0 < gsize < MAX_LINK_GROUP
              and not np.isin(st, gst).any() and not np.isin(ld, gst).any() and not np.isin(st, gld).any())
        if ok:
            gst = np.union1d(gst, st)
            gld = np.union1d(gld, ld)
            gsize += 1
        else:
            gst, gld, gsize, gtwo = np.unique(st), np.unique(ld), 1, bool(two)
        flags.append(1 if ok else 0)
    return flags


def host_lists(method, layout, wshift, flat=None):
    """The index lists one link method runs with, as DeviceLinks uploads them:
    (store, load, sign, rhs, conflict, order) with flat element indices of layout.
    flat=(store, load): a plain copy list f[store] = f[load] (method is only a name).
    conflict: the list loads a slot it stores or stores one slot twice; its
    sequential chains are then folded into load/sign/rhs (run two-phase).
    order: the permutation that sorted a non-conflicting list by store (or None)."""
    if flat is None:
        rows_s = method._store_rows()
        rows_l = method.iload[0].T if method._fixed else method.iload[0]
        st = _flat(rows_s, layout)
        ld = _flat(rows_l, layout)
        sign = np.full(st.size, method.kind.sign)
        rhs = method.rhs.copy() if method.kind.rhs_sign is not None else np.zeros(st.size)
    else:
        st, ld = (np.asarray(a, dtype=np.int64) for a in flat)
        sign = np.ones(st.size)
        rhs = np.zeros(st.size)
    conflict = bool((np.intersect1d(st, ld).size > 0) or (np.unique(st).size != st.size))
    order = None
    if conflict:
        st, ld, sign, rhs = _resolve_sequential(st, ld, sign, rhs)
    elif st.size > 1:
        # no slot is both stored and loaded and every store is unique, so the list
        # may run in any order: sort by store address, which puts the links of one
        # (o, m) row (all populations, interleaved layout) next to each other and
        # lets the lanes of a wave share cache lines
        order = np.argsort(st, kind="stable")
        st, ld, sign, rhs = st[order], ld[order], sign[order], rhs[order]
    if layout.storage in DEVIATION_FORMATS:
        # deviations: d_s = a d_l + (r + a w_l - w_s)
        rhs = rhs + sign * wshift[_population(ld, layout)] - wshift[_population(st, layout)]
    return st, ld, sign, rhs, conflict, order


def _bouzidi_flat(method, layout):
    """Store, first and second load (flat element indices of layout) and the
    per-link s of an interpolated method, in its link order."""
    rows_s = method._store_rows()
    l0r, l1r = [(a.T if method._fixed else a) for a in method.iload[:2]]
    dim = rows_s.shape[0] - 1
    for rows in (rows_s, l0r, l1r):
        for d in range(dim):
            c = np.asarray(rows[1 + d])
            if c.size and (c.min() < 0 or c.max() >= layout.n[3 - dim + d]):
                raise ValueError("%s: a link slot lies outside the halo-inclusive grid" % method.name)
    return (_flat(rows_s, layout), _flat(l0r, layout), _flat(l1r, layout),
            np.asarray(method.s, dtype=np.float64))


def _earlier_store_loads(st, l0, l1):
    """Whether a link loads a slot that an earlier link of the list stores."""
    n = st.size
    if not n:
        return False
    us, first = np.unique(st, return_index=True)
    links = np.concatenate([np.arange(n), np.arange(n)])
    loads = np.concatenate([l0, l1])
    pos = np.minimum(np.searchsorted(us, loads), us.size - 1)
    hit = us[pos] == loads
    return bool(np.any(first[pos[hit]] < links[hit]))


def _bouzidi_form(method, layout, wshift, index, st, l0, l1, s, two):
    """(store, load0, load1, ca, cb, rhs, two_phase, index, shift) of the links index."""
    ca = method.kind.sign * s
    cb = 1.0 - s
    shift = np.zeros(st.size)
    if layout.storage in DEVIATION_FORMATS:
        # deviations: d_s = ca d_0 + cb d_1 + (r + ca w_0 + cb w_1 - w_s)
        shift = (ca * wshift[_population(l0, layout)] + cb * wshift[_population(l1, layout)]
                 - wshift[_population(st, layout)])
    rhs = np.asarray(method.rhs, dtype=np.float64)[index] + shift
    return st, l0, l1, ca, cb, rhs, two, index, shift


def bouzidi_lists(method, layout, wshift):
    """The lists an interpolated (Bouzidi) method runs with:
    (store, load0, load1, ca, cb, rhs, two_phase, index, shift) with flat element
    indices of layout, f[store] = ca f[load0] + cb f[load1] + rhs.

    Bouzidi bounce-back reads a copy of f taken before its loop (pylbm/boundary.py:
    545-549), so every load sees the old value and the last store of a slot wins:
    earlier duplicate stores are dropped and, when a load hits a stored slot, the
    list runs two-phase (all values gathered before any store).  Bouzidi
    anti-bounce-back reads f itself in loop order: a load of a slot that an
    earlier link of the loop stored needs that link's new value, which one list
    does not carry (NotImplementedError here; bouzidi_levels splits such a list);
    loads of slots stored later see the old value, as two-phase gives them.
    index: link positions of the method kept, in run order (rhs[index] + shift is
    the device rhs); shift: the rest-population term of deviation formats."""
    st, l0, l1, s = _bouzidi_flat(method, layout)
    n = st.size
    if not method.kind.reads_copy and _earlier_store_loads(st, l0, l1):
        raise NotImplementedError("%s: a link loads a slot an earlier link of the list stores"
                                  % method.name)
    # the last store of each slot
    _, last_rev = np.unique(st[::-1], return_index=True)
    index = np.sort(n - 1 - last_rev)
    st, l0, l1, s = st[index], l0[index], l1[index], s[index]
    two = bool(np.isin(np.concatenate([l0, l1]), st).any())
    if not two and st.size > 1:
        order = np.argsort(st, kind="stable")
        index, st, l0, l1, s = index[order], st[order], l0[order], l1[order], s[order]
    return _bouzidi_form(method, layout, wshift, index, st, l0, l1, s, two)


def bouzidi_levels(method, layout, wshift):
    """The interpolated method as lists run one after the other, each in the form of
    bouzidi_lists.  One list, unless Bouzidi anti-bounce-back has a link that loads
    a slot an earlier link of its loop stored (pylbm's in-place loop, pylbm/
    boundary.py:743-760: a one-cell gap whose nearer wall is the upper one).  Then
    the links are split into dependency levels, in link order: a link's level is
    above that of the latest earlier link storing a slot it loads (it reads that
    value), above every earlier link loading the slot it stores (they read the old
    value) and above every earlier link storing that slot (the last store wins).
    The links of one level touch no slot another one of the level stores, so each
    level runs one-phase, and the levels in order give the loop's result."""
    if method.kind.reads_copy:
        return [bouzidi_lists(method, layout, wshift)]
    st, l0, l1, s = _bouzidi_flat(method, layout)
    if not _earlier_store_loads(st, l0, l1):
        return [bouzidi_lists(method, layout, wshift)]
    n = st.size
    level = np.zeros(n, dtype=np.int64)
    stored, read = {}, {}          # slot -> level of its latest store / highest read
    for j in range(n):
        a, b, x = int(l0[j]), int(l1[j]), int(st[j])
        lv = max(stored.get(a, -1), stored.get(b, -1), read.get(x, -1), stored.get(x, -1)) + 1
        level[j] = lv
        stored[x] = lv
        read[a] = max(read.get(a, -1), lv)
        read[b] = max(read.get(b, -1), lv)
    out = []
    for lv in range(int(level.max()) + 1):
        index = np.nonzero(level == lv)[0]
        order = np.argsort(st[index], kind="stable")
        index = index[order]
        out.append(_bouzidi_form(method, layout, wshift, index, st[index], l0[index], l1[index], s[index],
                                 False))
    return out


class DeviceLinks:
    """Device arrays of one method for a given population layout."""

    def __init__(self, rt, method, layout, wshift, flat=None, level=0):
        """flat=(store, load): a plain copy list f[store] = f[load] given as flat
        element indices (method is then only a name).  level: which list of
        bouzidi_levels an interpolated method's object runs."""
        self.rt = rt
        self.method = method
        self.layout = layout
        self.level = level
        self.bouzidi = flat is None and bool(getattr(method.kind, "bouzidi", False))
        self.d_load1 = self.d_ca = self.d_cb = None
        if self.bouzidi:
            self._init_bouzidi(method, layout, wshift)
            return
        st, ld, sign, rhs, self.conflict, self.order = host_lists(method, layout, wshift, flat)
        self.n = st.size
        # host copies of the final lists, for merge_flags and link folding (dropped after use)
        self.h_store, self.h_load = st, ld
        self.h_sign = sign
        self.h_rhs = rhs if np.any(rhs != 0) else None
        ctype = np.float64 if layout.storage == LBM_F64 else np.float32
        self._ptrs = []
        self.d_store = self._up(st.astype(np.uint32))
        self.d_load = self._up(ld.astype(np.uint32))
        self.d_rhs = self._up(rhs.astype(ctype)) if np.any(rhs != 0) else None
        uniform = np.all(sign == sign[0]) if self.n else True
        self.sign = float(sign[0]) if self.n else 1.0
        self.d_sign = None if uniform else self._up(sign.astype(np.int8))
        self.d_work = rt.malloc(max(1, self.n) * np.dtype(ctype).itemsize) if self.conflict else None
        if self.d_work:
            self._ptrs.append(self.d_work)

    def bouzidi_form(self, layout, wshift):
        """This object's interpolated list in the form of bouzidi_lists, for layout."""
        return bouzidi_levels(self.method, layout, wshift)[self.level]

    def _init_bouzidi(self, method, layout, wshift):
        st, l0, l1, ca, cb, rhs, self.conflict, self.order, self._shift = self.bouzidi_form(layout, wshift)
        self.n = st.size
        self.h_store, self.h_load = st, np.concatenate([l0, l1])
        self.h_sign, self.h_rhs = None, None
        ctype = np.float64 if layout.storage == LBM_F64 else np.float32
        self._ptrs = []
        self.d_store = self._up(st.astype(np.uint32))
        self.d_load = self._up(l0.astype(np.uint32))
        self.d_load1 = self._up(l1.astype(np.uint32))
        self.d_ca = self._up(ca.astype(ctype))
        self.d_cb = self._up(cb.astype(ctype))
        self.d_rhs = self._up(rhs.astype(ctype)) if np.any(rhs != 0) else None
        self.sign, self.d_sign = 1.0, None
        self.d_work = self.rt.malloc(max(1, self.n) * np.dtype(ctype).itemsize) if self.conflict else None
        if self.d_work:
            self._ptrs.append(self.d_work)

    def _up(self, a):
        p = self.rt.upload(a)
        self._ptrs.append(p)
        return p

    def update_rhs(self, rhs, wshift):
        """Re-upload rhs (time-dependent boundary values)."""
        if self.bouzidi:
            ctype = np.float64 if self.layout.storage == LBM_F64 else np.float32
            dev = (np.asarray(rhs, dtype=np.float64)[self.order] + self._shift).astype(ctype)
            if self.d_rhs is None:
                self.d_rhs = self._up(dev)
            else:
                self.rt.h2d(self.d_rhs, dev)
            return
        if self.conflict:
            raise NotImplementedError("time-dependent values on a self-overlapping link list")
        if self.layout.storage in DEVIATION_FORMATS:
            st = _flat(self.method._store_rows(), self.layout)
            ld = _flat(self.method.iload[0].T if self.method._fixed else self.method.iload[0], self.layout)
            rhs = (rhs + self.method.kind.sign * wshift[_population(ld, self.layout)]
                   - wshift[_population(st, self.layout)])
        if self.order is not None:
            rhs = np.asarray(rhs)[self.order]
        ctype = np.float64 if self.layout.storage == LBM_F64 else np.float32
        if self.d_rhs is None:
            self.d_rhs = self._up(rhs.astype(ctype))
        else:
            self.rt.h2d(self.d_rhs, rhs.astype(ctype))

    def link_set(self):
        s = LbmLinkSet()
        s.d_store, s.d_load = self.d_store, self.d_load
        s.d_rhs = self.d_rhs
        s.d_sign = self.d_sign
        s.sign = self.sign
        s.n = self.n
        s.two_phase = 1 if self.conflict else 0
        s.merge_prev = 0
        s.d_work = self.d_work
        s.d_load1, s.d_ca, s.d_cb = self.d_load1, self.d_ca, self.d_cb
        return s

    def apply(self, d_f):
        rt = self.rt
        if self.bouzidi:
            p = ctypes.c_void_p
            rt.check(rt.lib.lbm_links_apply_bouzidi(p(d_f), ctypes.byref(self.layout), p(self.d_store),
                                                    p(self.d_load), p(self.d_load1), p(self.d_ca), p(self.d_cb),
                                                    p(self.d_rhs), self.n, 1 if self.conflict else 0,
                                                    p(self.d_work), rt.stream), self.method.name)
            return
        rt.check(rt.lib.lbm_links_apply(ctypes.c_void_p(d_f), ctypes.byref(self.layout),
                                        ctypes.c_void_p(self.d_store), ctypes.c_void_p(self.d_load),
                                        ctypes.c_void_p(self.d_rhs), ctypes.c_void_p(self.d_sign),
                                        self.sign, self.n, 1 if self.conflict else 0,
                                        ctypes.c_void_p(self.d_work), rt.stream),
                 getattr(self.method, "name", str(self.method)))

    def __del__(self):
        try:
            for p in self._ptrs:
                self.rt.free(p)
        except Exception:
            pass


# ------------------------------------------------------ ghost slots outside the fill
def unwritten_ghost_reads(domain, stencil, methods, layout):
    """Ghost slots a pull or a Neumann-type list reads that the periodic-face fill does not write.

    pylbm wraps every halo slot before the link lists run (pylbm/storage.py:306-367).
    The device fill only wraps the slabs of periodic faces (a slot belongs to the
    slab of its first ghost axis, pylbm's rule for corner links), because the
    pulled slots of the other slabs are link stores -- except where pylbm dropped
    the link: cells inside a solid element lose their flags (pylbm/domain.py:
    578-589, 622-635), so a solid cell next to a wall pulls the wrapped value.
    Returns flat (store, load) indices of those slots and their wrapped sources."""
    dim = domain.dim
    shape = list(domain.shape_halo)
    h = [int(x) for x in stencil.vmax]
    v = stencil.get_all_velocities()
    labels = domain.box_label
    written = np.concatenate([_flat(m._store_rows(), layout) for m in methods]) if methods else np.zeros(0, np.int64)
    stores, loads = [], []
    for a in range(dim):
        for side in (0, 1):
            if labels[2 * a + side] == -1 or h[a] == 0:
                continue
            ranges = []
            for b in range(dim):
                if b < a:
                    ranges.append(np.arange(h[b], shape[b] - h[b]))
                elif b == a:
                    ranges.append(np.arange(0, h[a]) if side == 0 else np.arange(shape[a] - h[a], shape[a]))
                else:
                    ranges.append(np.arange(shape[b]))
            cells = np.stack([g.ravel() for g in np.meshgrid(*ranges, indexing="ij")])
            for k in range(v.shape[0]):
                x = cells + v[k][:, None]
                inside = np.all([(x[b] >= h[b]) & (x[b] < shape[b] - h[b]) for b in range(dim)], axis=0)
                if not inside.any():
                    continue
                s_ = cells[:, inside]
                wrapped = s_.copy()
                for b in range(dim):
                    nin = shape[b] - 2 * h[b]
                    wrapped[b] = np.where(s_[b] < h[b], s_[b] + nin, np.where(s_[b] >= shape[b] - h[b], s_[b] - nin, s_[b]))
                kk = np.full((1, s_.shape[1]), k)
                st = _flat(np.concatenate([kk, s_]), layout)
                keep = ~np.isin(st, written)
                if keep.any():
                    stores.append(st[keep])
                    loads.append(_flat(np.concatenate([kk, wrapped]), layout)[keep])
    # loads of the Neumann-type lists that fall on ghost slots: pylbm's wrap put the
    # fully wrapped value there before any list ran (its axis-by-axis wrap ends every
    # slot with a copy of an already wrapped one); a slot an earlier list stores to is
    # overwritten after this copy, as in pylbm
    for m in methods:
        if m.kind.shift_axis is None and m.kind.symmetric_load:
            continue
        rows = np.asarray(m.iload[0].T if m._fixed else m.iload[0])
        cells = rows[1:]
        ghost = np.any([(cells[b] < h[b]) | (cells[b] >= shape[b] - h[b]) for b in range(dim)], axis=0)
        if not ghost.any():
            continue
        s_ = cells[:, ghost]
        wrapped = s_.copy()
        for b in range(dim):
            nin = shape[b] - 2 * h[b]
            wrapped[b] = np.where(s_[b] < h[b], s_[b] + nin, np.where(s_[b] >= shape[b] - h[b], s_[b] - nin, s_[b]))
        kk = rows[:1, ghost]
        stores.append(_flat(np.concatenate([kk, s_]), layout))
        loads.append(_flat(np.concatenate([kk, wrapped]), layout))
    if not stores:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    st, ld = np.concatenate(stores), np.concatenate(loads)
    st, first = np.unique(st, return_index=True)
    return st, ld[first]
