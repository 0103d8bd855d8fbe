"""Velocity numbering and symmetry -- oracle restatement.

Follows pylbm/stencil.py:
  * 1D numbering         stencil.py:290-294 (num -> vx at 329-333)
  * 2D numbering (_R2)   stencil.py:178-185, 295-306, 334-350
  * 3D enumeration       stencil.py:307-327, 351-372 (permute_in_place 30-100)
  * Stencil fields       stencil.py:643-699 (unique velocities, nv_ptr, unum2index)
  * get_all_velocities   stencil.py:783-815
  * get_symmetric        stencil.py:817-835

Pinned by: the velocity tables of tests/conftest.py:19-165 (copied as data into
tests/golden/stencil_tables.json), the involution property of
tests/test_velocity.py, and brute-force inversion num -> v -> num.
"""
import itertools
import math

import numpy as np

_D = 1000
_R2 = np.array([[[5, 6, 4], [_D, _D, 2], [2, 5, 3]],
                [[3, _D, _D], [_D, -1, _D], [1, _D, _D]],
                [[6, 7, 7], [_D, _D, 0], [1, 4, 0]]], dtype=int)


def _cmp(a, b):
    a, b = int(a), int(b)
    return (a > b) - (a < b)


def _enum3d():
    """Yield (vx, vy, vz) in the reference's 3D enumeration order."""
    for k in itertools.count():
        for i in range(k + 1):
            for j in range(i + 1):
                # unique permutations in lexicographic order starting from the
                # ascending sort (what permute_in_place yields)
                perms = sorted(set(itertools.permutations([k, i, j])))
                for kk, ii, jj in perms:
                    for sk in ([1, -1] if kk else [1]):
                        for si in ([1, -1] if ii else [1]):
                            for sj in ([1, -1] if jj else [1]):
                                yield (sk * kk, si * ii, sj * jj)


_CACHE3 = []


def _table3d(n):
    if len(_CACHE3) < n:
        _CACHE3.clear()
        for t in _enum3d():
            _CACHE3.append(t)
            if len(_CACHE3) >= max(n, 400):
                break
    return _CACHE3


def num_to_v(dim, num):
    if dim == 1:
        n = num + 1
        return (int((1 - 2 * (n % 2)) * (n / 2)),)
    if dim == 2:
        n = int((math.sqrt(num) + 1) / 2)
        p = num - (2 * n - 1) ** 2
        if p < 4:
            return ([n, 0, -n, 0][p], [0, n, 0, -n][p])
        if p < 8:
            return ([n, -n, -n, n][p - 4], [n, n, -n, -n][p - 4])
        i, j = n, p // 8
        lx = [i, j, -j, -i, -i, -j, j, i]
        ly = [j, i, i, j, -j, -i, -i, -j]
        return (lx[p % 8], ly[p % 8])
    return _table3d(num + 1)[num]


def v_to_num(v):
    v = tuple(int(c) for c in v)
    dim = len(v)
    if dim == 1:
        vx = v[0]
        return 2 * abs(vx) - (1 if vx > 0 else 0)
    if dim == 2:
        vx, vy = v
        ax, ay = abs(vx), abs(vy)
        t3 = _cmp(ax, ay)
        p = (2 * max(ax, ay) - 1) ** 2
        q = 8 * min(ax, ay) * abs(t3)
        r = _R2[_cmp(vx, 0) + 1, _cmp(vy, 0) + 1, t3 + 1]
        return int(p + q + r)
    tab = _table3d(400)
    return tab.index(tuple(v))


class Stencil:
    """Velocities of one or several elementary schemes (stencil.py:643-699)."""

    def __init__(self, dim, schemes_velocities):
        self.dim = dim
        self.nums = [list(v) for v in schemes_velocities]
        unum = np.unique(np.concatenate([np.asarray(v, dtype=int) for v in self.nums]))
        self.unum = unum
        self.uv = np.array([num_to_v(dim, int(n)) for n in unum], dtype=int).reshape(-1, dim)
        self.nv = [len(v) for v in self.nums]
        self.nv_ptr = np.concatenate([[0], np.cumsum(self.nv)]).astype(int)
        self.unum2index = -1000 + np.zeros(int(unum.max()) + 1, dtype=int)
        self.unum2index[unum] = np.arange(unum.size)
        self.unvtot = unum.size
        self.vmax = np.max(self.uv, axis=0)

    def all_velocities(self):
        """Velocities of all schemes, concatenated (stencil.py:783-815)."""
        return np.array([num_to_v(self.dim, n) for v in self.nums for n in v],
                        dtype=int).reshape(-1, self.dim)

    def symmetric(self):
        """Global index of -v_k inside the same elementary scheme (stencil.py:817-835)."""
        out = []
        for s, nums in enumerate(self.nums):
            for n in nums:
                vs = tuple(-c for c in num_to_v(self.dim, n))
                out.append(self.nv_ptr[s] + nums.index(v_to_num(vs)))
        return np.array(out, dtype=np.int32)
