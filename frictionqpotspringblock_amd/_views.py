"""Writable state arrays of a system.

The reference returns its state arrays as views of the C++ object
(python/main.cpp:80-127, ``py::return_value_policy::reference_internal``), so
that ``system.u[i] = x`` changes the system. Here the state lives in GPU
memory and the C ABI hands out copies; the ``u``, ``v`` and ``a`` properties
therefore return a ``FieldView``: a copy whose item assignments (including
those through slices of it and augmented assignments such as
``system.u[2:5] += 1``) are written back to the system at once. What a copy
cannot follow: a view kept across a step still holds the values it was read
with (read the property again), and writes that bypass ``__setitem__``
(``np.copyto``, a ufunc's ``out=``) stay in the copy (assign the property
instead: ``system.u = x``).
"""

from __future__ import annotations

import numpy as np


class FieldView(np.ndarray):
    """A copy of a system's state array that writes item assignments back."""

    def __new__(cls, data, owner, setter):
        obj = np.asarray(data).view(cls)
        obj._owner = (owner, setter)
        return obj

    def __array_finalize__(self, obj):
        # slices and views keep no owner of their own: they write back
        # through the array they were taken from (their .base)
        self._owner = None

    def __array_wrap__(self, array, context=None, return_scalar=False):
        # results of arithmetic are plain arrays, not views of the system
        out = np.asarray(array)
        return out[()] if return_scalar else out

    def _root(self):
        x = self
        while isinstance(x, FieldView):
            if x._owner is not None:
                return x
            x = x.base
        return None

    def __setitem__(self, key, value):
        super().__setitem__(key, value)
        root = self._root()
        if root is not None:
            owner, setter = root._owner
            setter(owner, np.asarray(root))


def install(cls, names=("u", "v", "a")):
    """Make the properties `names` of the pybind class `cls` (and so of its
    subclasses) return FieldViews."""
    for name in names:
        prop = cls.__dict__[name]
        getter, setter = prop.fget, prop.fset

        def fget(self, _get=getter, _set=setter):
            return FieldView(_get(self), self, _set)

        setattr(cls, name, property(fget, setter, doc=prop.__doc__))
