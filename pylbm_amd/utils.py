"""``itemproperty``: a property indexed with [] (pylbm/utils.py:26-70), used for
``sim.m[RHO]`` and ``sim.F[k]``."""

__all__ = ["itemproperty"]


class _Bound:
    def __init__(self, prop, obj):
        self._prop = prop
        self._obj = obj

    def __getitem__(self, key):
        return self._prop.fget(self._obj, key)

    def __setitem__(self, key, value):
        if self._prop.fset is None:
            raise AttributeError("read-only item property")
        self._prop.fset(self._obj, key, value)


class itemproperty:  # noqa: N801  (pylbm's name)
    def __init__(self, fget=None, fset=None, doc=None):
        self.fget = fget
        self.fset = fset
        self.__doc__ = doc or (fget.__doc__ if fget else None)

    def __get__(self, obj, objtype=None):
        if obj is None:
            return self
        return _Bound(self, obj)

    def setter(self, fset):
        return itemproperty(self.fget, fset, self.__doc__)
