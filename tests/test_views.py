"""Writable state arrays (FieldView): item assignments, through slices and
reshaped views too, reach the owner's setter; arithmetic results are plain
arrays. Host logic only (a recording setter stands in for a system)."""

import numpy as np

from frictionqpotspringblock_amd._views import FieldView


def _view(n=6):
    rec = []
    return FieldView(np.arange(float(n)), object(), lambda owner, x: rec.append(x.copy())), rec


def test_item_and_slice_assignments_write_back():
    v, rec = _view()
    v[2] = 10.0
    assert rec[-1].tolist() == [0, 1, 10, 3, 4, 5]
    w = v[1:4]
    w[0] = -1.0
    assert rec[-1].tolist() == [0, -1, 10, 3, 4, 5]
    v[3:5] += 100.0
    assert rec[-1].tolist() == [0, -1, 10, 103, 104, 5]
    z = v.reshape(2, 3)
    z[1, 2] = 7.0
    assert rec[-1].tolist() == [0, -1, 10, 103, 104, 7]
    assert all(type(x) is np.ndarray for x in rec)


def test_arithmetic_gives_plain_arrays():
    v, rec = _view()
    y = v + 1.0
    assert type(y) is np.ndarray
    y[0] = 5.0
    assert rec == []
    assert type(v.sum()) is np.float64 and v.sum() == 15.0
    assert np.array_equal(v, np.arange(6.0))
