"""fqsb_batch_quasistatic_noalloc (the quasistatic protocol with its records
staged in the library's fixed device array instead of per-call buffers)
against fqsb_batch_quasistatic, through the C ABI: identical records and
states on two calls in a row, the second on a smaller batch with fewer
events, and the size limit."""

import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(R, N, seed):
    from frictionqpotspringblock_amd.batch import Batch

    return Batch(
        R,
        seed_stride=1,
        potential="Cuspy",
        interactions="Laplace1d",
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=seed,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=200,
    )


def _call(fn, b, first, nevents):
    from frictionqpotspringblock_amd.batch import _check

    R = b.replicas
    S = np.full((R, nevents), -7, dtype=np.int64)
    uf = np.full((R, nevents), np.nan)
    ff = np.full((R, nevents), np.nan)
    inc = np.full((R, nevents), -7, dtype=np.int64)
    ip = C.POINTER(C.c_int64)
    dp = C.POINTER(C.c_double)
    _check(
        fn(b._h, first, nevents, 1e-3, 1, 0, 1e-5, 10, 10**9, S.ctypes.data_as(ip), uf.ctypes.data_as(dp),
           ff.ctypes.data_as(dp), inc.ctypes.data_as(ip))
    )
    return S, uf, ff, inc


def _same(a, b, what):
    for x, y in zip(a, b):
        assert np.array_equal(x, y), what
    return True


def test_noalloc_equals_allocating_entry_point():
    from frictionqpotspringblock_amd.batch import capi

    lib = capi()
    print("noalloc: call 1, R=64 N=512, events 0..24", flush=True)
    a, b = _batch(64, 512, 3), _batch(64, 512, 3)
    ra = _call(lib.fqsb_batch_quasistatic, a, 0, 25)
    rb = _call(lib.fqsb_batch_quasistatic_noalloc, b, 0, 25)
    _same(ra, rb, "records, call 1")
    assert np.any(ra[0][:, 2::2] != 0)  # avalanches happened
    assert np.array_equal(a.u, b.u) and np.array_equal(a.index_at_align, b.index_at_align)
    assert np.array_equal(a.inc, b.inc) and np.array_equal(a.u_frame, b.u_frame)
    # second call in a row: smaller batch, fewer events, continuing the protocol
    print("noalloc: call 2, R=8 N=300, events 0..6", flush=True)
    c, d = _batch(8, 300, 11), _batch(8, 300, 11)
    rc = _call(lib.fqsb_batch_quasistatic, c, 0, 7)
    rd = _call(lib.fqsb_batch_quasistatic_noalloc, d, 0, 7)
    _same(rc, rd, "records, call 2")
    assert np.array_equal(c.u, d.u) and np.array_equal(c.inc, d.inc)
    # and the first batch continues from event 25 through the scratch again
    print("noalloc: call 3, R=64 N=512, events 25..28", flush=True)
    ra = _call(lib.fqsb_batch_quasistatic, a, 25, 4)
    rb = _call(lib.fqsb_batch_quasistatic_noalloc, b, 25, 4)
    _same(ra, rb, "records, call 3")
    assert np.array_equal(a.u, b.u)


def test_noalloc_limit_and_errors():
    from frictionqpotspringblock_amd.batch import Batch, capi

    lib = capi()
    b = _batch(64, 256, 0)
    u0 = b.u
    with pytest.raises(RuntimeError, match="at most 262144 records"):
        _call(lib.fqsb_batch_quasistatic_noalloc, b, 0, 262144 // 64 + 1)
    assert np.array_equal(b.u, u0) and np.all(b.inc == 0)
    # a batch beyond one resident workgroup has no device-side driver
    big = _batch(1, 20000, 0)
    with pytest.raises(RuntimeError, match="device-side driver"):
        _call(lib.fqsb_batch_quasistatic_noalloc, big, 0, 3)
    assert isinstance(Batch, type)
