"""Lattice velocities, their numbering and the stencil of a scheme.

The numbering is pylbm's convention (it is part of the user-facing dict
format, ``'velocities': list(range(19))``), so it has to be identical:

* 1D: 0, +1, -1, +2, -2, ...                          pylbm/stencil.py:290-294, 329-333
* 2D: rings of growing max-norm, table-driven         pylbm/stencil.py:178-185, 295-306, 334-350
* 3D: enumeration by (max, mid, min) magnitude, then
  unique permutations in lexicographic order, then
  signs + before -                                    pylbm/stencil.py:307-327, 351-372

``Stencil`` mirrors pylbm/stencil.py:643-835 (unique velocities, ``nv_ptr``,
``unum2index``, ``vmax``, ``get_all_velocities``, ``get_symmetric``).
"""
import itertools
import math

import numpy as np

__all__ = ["Velocity", "Stencil"]

_BIG = 1000
# 2D offset table indexed by (sign vx, sign vy, cmp(|vx|, |vy|)) + 1
_RING_OFFSET = np.array(
    [[[5, 6, 4], [_BIG, _BIG, 2], [2, 5, 3]],
     [[3, _BIG, _BIG], [_BIG, -1, _BIG], [1, _BIG, _BIG]],
     [[6, 7, 7], [_BIG, _BIG, 0], [1, 4, 0]]], dtype=np.int64)


def _sign(a):
    return (a > 0) - (a < 0)


def _build_3d_table(limit=7):
    out = []
    for k in range(limit):
        for i in range(k + 1):
            for j in range(i + 1):
                for perm in sorted(set(itertools.permutations((k, i, j)))):
                    signs = [(1, -1) if c else (1,) for c in perm]
                    for sx, sy, sz in itertools.product(*signs):
                        out.append((sx * perm[0], sy * perm[1], sz * perm[2]))
    return out


_TABLE3D = _build_3d_table()
_INDEX3D = {v: i for i, v in enumerate(_TABLE3D)}


def velocity_from_num(dim, num):
    """Components of velocity number ``num`` in dimension ``dim``."""
    if dim == 1:
        n = num + 1
        return (int((1 - 2 * (n % 2)) * (n // 2)),)
    if dim == 2:
        n = int((math.sqrt(num) + 1) / 2)
        p = num - (2 * n - 1) ** 2
        if p < 4:
            return ((n, 0, -n, 0)[p], (0, n, 0, -n)[p])
        if p < 8:
            return ((n, -n, -n, n)[p - 4], (n, n, -n, -n)[p - 4])
        j = p // 8
        xs = (n, j, -j, -n, -n, -j, j, n)
        ys = (j, n, n, j, -j, -n, -n, -j)
        return (xs[p % 8], ys[p % 8])
    if dim == 3:
        return _TABLE3D[num]
    raise ValueError("dimension must be 1, 2 or 3")


def num_from_velocity(v):
    """Velocity number of the integer vector ``v``."""
    v = tuple(int(c) for c in v)
    if len(v) == 1:
        return 2 * abs(v[0]) - (1 if v[0] > 0 else 0)
    if len(v) == 2:
        ax, ay = abs(v[0]), abs(v[1])
        c = _sign(ax - ay)
        ring = (2 * max(ax, ay) - 1) ** 2
        return int(ring + 8 * min(ax, ay) * abs(c)
                   + _RING_OFFSET[_sign(v[0]) + 1, _sign(v[1]) + 1, c + 1])
    if len(v) == 3:
        return _INDEX3D[v]
    raise ValueError("dimension must be 1, 2 or 3")


class Velocity:
    """One lattice velocity (pylbm/stencil.py:119-373)."""

    def __init__(self, dim=None, num=None, vx=None, vy=None, vz=None):
        if dim is None:
            dim = 3 if vz is not None else 2 if vy is not None else 1
        self.dim = dim
        if num is None:
            num = num_from_velocity([vx, vy, vz][:dim])
        self.num = int(num)
        comps = velocity_from_num(dim, self.num)
        self.vx = comps[0]
        self.vy = comps[1] if dim > 1 else None
        self.vz = comps[2] if dim > 2 else None

    @property
    def v(self):
        return [self.vx, self.vy, self.vz][: self.dim]

    @property
    def v_full(self):
        out = [0, 0, 0]
        out[: self.dim] = self.v
        return out

    def get_symmetric(self, axis=None):
        if axis is not None and not 0 <= axis < self.dim:
            raise ValueError("axis must be smaller than the dimension")
        v = [-c for c in self.v]
        if axis is not None:
            v[axis] = -v[axis]
        return Velocity(dim=self.dim, num=num_from_velocity(v))

    def __repr__(self):
        return "(%d: %s)" % (self.num, ", ".join(str(c) for c in self.v))


def _dim_from_dico(dico):
    if dico.get("dim"):
        return dico["dim"]
    box = dico["box"]
    dim = 1
    if box.get("y") is not None:
        dim += 1
        if box.get("z") is not None:
            dim += 1
    return dim


class Stencil:
    """Velocities of the elementary schemes of a dict (pylbm/stencil.py:643-835)."""

    def __init__(self, dico, need_validation=True):
        self.dim = _dim_from_dico(dico)
        self._nums = [np.asarray(s["velocities"], dtype=np.int64) for s in dico["schemes"]]
        self.nstencils = len(self._nums)
        unum = np.unique(np.concatenate(self._nums))
        self.unique_velocities = np.asarray([Velocity(dim=self.dim, num=int(n)) for n in unum])
        self._uv = np.array([velocity_from_num(self.dim, int(n)) for n in unum],
                            dtype=np.int64).reshape(-1, self.dim)
        self.nv = [int(a.size) for a in self._nums]
        self.nv_ptr = np.concatenate([[0], np.cumsum(self.nv)]).astype(np.int64)
        self.unum2index = np.full(int(unum.max()) + 1, -1000, dtype=np.int32)
        self.unum2index[unum] = np.arange(unum.size, dtype=np.int32)
        self.num2index = [int(n) for a in self._nums for n in a]
        self.v = [np.asarray([Velocity(dim=self.dim, num=int(n)) for n in a]) for a in self._nums]

    # unique-velocity views
    @property
    def unvtot(self):
        return int(self._uv.shape[0])

    @property
    def unum(self):
        return np.array([v.num for v in self.unique_velocities], dtype=np.int64)

    def _ucomp(self, d):
        if d >= self.dim:
            return np.array([None] * self.unvtot)
        return self._uv[:, d].copy()

    @property
    def uvx(self):
        return self._ucomp(0)

    @property
    def uvy(self):
        return self._ucomp(1)

    @property
    def uvz(self):
        return self._ucomp(2)

    @property
    def uv(self):
        """(unvtot, dim) integer components of the unique velocities."""
        return self._uv

    @property
    def vmax(self):
        return np.max(self._uv, axis=0)

    @property
    def vmin(self):
        return np.min(self._uv, axis=0)

    @property
    def vmax_full(self):
        out = np.zeros(3, dtype=np.int64)
        out[: self.dim] = self.vmax
        return out

    def num(self, k):
        return self._nums[k].copy()

    def get_all_velocities(self, scheme_id=None):
        ids = range(self.nstencils) if scheme_id is None else [scheme_id]
        rows = [velocity_from_num(self.dim, int(n)) for k in ids for n in self._nums[k]]
        return np.asarray(rows, dtype=np.int64).reshape(-1, self.dim)

    def get_symmetric(self, axis=None):
        out = np.empty(int(self.nv_ptr[-1]), dtype=np.int32)
        g = 0
        for k, nums in enumerate(self._nums):
            pos = {int(n): i for i, n in enumerate(nums)}
            for n in nums:
                sym = Velocity(dim=self.dim, num=int(n)).get_symmetric(axis).num
                out[g] = self.nv_ptr[k] + pos[sym]
                g += 1
        return out

    def is_symmetric(self):
        try:
            self.get_symmetric()
            return True
        except KeyError:
            return False
