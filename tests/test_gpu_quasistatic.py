"""GPU parity of the quasistatic driver (System.quasistatic, Batch.quasistatic,
Ragged.quasistatic) against the oracle running the examples' Python loop
(examples/Line1d_Cuspy_Laplace.py:39-61) event by event.

Avalanche sizes S, increments and frame positions are compared exactly, and
so is the state after the run (the driver evaluates the same expressions as
the single operations). The mean frame force is reduced in a fixed order on
the device and by numpy's pairwise sum on the host: relative tolerance 1e-12.
"""

import numpy as np
import pytest

from tests import protocols as P
from tests.test_gpu_parity import _base, _make, assert_same_state

pytestmark = pytest.mark.gpu

FF_RTOL = 1e-12


@pytest.fixture(scope="module")
def fq():
    import frictionqpotspringblock_amd as fq

    return fq


def assert_same_events(q, u_frame, f_frame, S, inc, what=""):
    n = len(S)
    assert np.array_equal(q["S"][-n:], S), (what, q["S"], S)
    assert np.array_equal(q["inc"][-n:], inc), what
    assert np.array_equal(q["u_frame"][-n:], u_frame), what
    np.testing.assert_allclose(q["f_frame"][-n:], f_frame, rtol=FF_RTOL, atol=1e-15, err_msg=what)


# (submodule, class, size, extra constructor arguments)
SYSTEMS = [
    ("Line1d", "System_Cuspy_Laplace", 3, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 67, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 1000, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 2000, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 3001, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 2000, dict(k_interactions=1.0, m=2.0)),
    ("Line1d", "System_SemiSmooth_Laplace", 700, dict(k_interactions=1.0, kappa=1.0)),
    ("Line1d", "System_Cuspy_Quartic", 900, dict(a1=1.0, a2=5.0)),
    ("Line1d", "System_Cuspy_QuarticGradient", 900, dict(k2=1.0, k4=2.0)),
    ("Line1d", "System_Cuspy_LongRange", 300, dict(k_interactions=1.0, alpha=1.5)),
    ("Line2d", "System_Cuspy_Laplace", (30, 40), dict(k_interactions=1.0)),
    ("Line2d", "System_Cuspy_QuarticGradient", (20, 25), dict(k2=1.0, k4=1.0)),
    ("Particles", "System_Cuspy", 900, {}),
    ("Particles", "System_SemiSmooth", 900, dict(kappa=1.0)),
]


@pytest.mark.parametrize(
    "sub,cls,n,extra", SYSTEMS, ids=[f"{s[1]}-{s[0]}-{s[2]}" + ("-m2" if "m" in s[3] else "") for s in SYSTEMS]
)
def test_driver_matches_protocol(fq, oracle, sub, cls, n, extra):
    """41 events on the device against the oracle's loop; nchunk 60 makes the
    driver refill landscapes both during minimisations and after kicks."""
    a = _make(fq, sub, cls, n, extra)
    b = _make(oracle, sub, cls, n, extra)
    q = a.quasistatic(1e-3, 41)
    ref = P.quasistatic(b, 41)
    assert_same_events(q, *ref, what=cls)
    assert np.sum(np.abs(ref[2])) > 0
    assert_same_state(a, b, "after the driver")


@pytest.mark.parametrize("budget", ["7", "1000"])
def test_driver_relaunches(fq, oracle, monkeypatch, budget):
    """Launches of a few steps (the run resumes inside minimisations, between
    events and after refills) and a run split over two calls give the same
    events and state as the oracle."""
    monkeypatch.setenv("FQSB_QS_BLOCK", budget)
    extra = dict(k_interactions=1.0)
    a = _make(fq, "Line1d", "System_Cuspy_Laplace", 500, extra, nchunk=40)
    b = _make(oracle, "Line1d", "System_Cuspy_Laplace", 500, extra, nchunk=40)
    q1 = a.quasistatic(1e-3, 12)
    q2 = a.quasistatic(1e-3, 17, first_event=12)
    ref = P.quasistatic(b, 29)
    q = {k: np.concatenate([q1[k], q2[k]]) for k in q1}
    assert_same_events(q, *ref)
    assert_same_state(a, b, "after the split run")


def test_direction_left(fq, oracle):
    extra = dict(k_interactions=1.0)
    a = _make(fq, "Line1d", "System_Cuspy_Laplace", 600, extra)
    b = _make(oracle, "Line1d", "System_Cuspy_Laplace", 600, extra)
    q = a.quasistatic(1e-3, 21, direction=-1)
    for step in range(21):
        if step == 0:
            b.u_frame = 0.0
        else:
            b.eventDrivenStep(1e-3, step % 2 == 0, -1)
        if step % 2 == 0:
            assert b.minimise() == 0
    assert q["u_frame"][-1] < 0
    assert_same_state(a, b, "direction -1")


def kicks_protocol(system, nstep, eps=1e-3):
    """The loop with a kick before every minimisation (the form available to
    the Smooth potential, whose maxUniformDisplacement raises)."""
    out = [np.empty(nstep), np.empty(nstep), np.empty(nstep, dtype=np.int64), np.empty(nstep, dtype=np.int64)]
    for step in range(nstep):
        i_n = np.copy(system.chunk.index_at_align)
        if step == 0:
            system.u_frame = 0.0
        else:
            system.eventDrivenStep(eps, True)
        assert system.minimise() == 0
        out[0][step] = system.u_frame
        out[1][step] = np.mean(system.f_frame)
        out[2][step] = np.sum(system.chunk.index_at_align - i_n)
        out[3][step] = system.inc
    return out


@pytest.mark.parametrize("cls,extra", [("System_Smooth_Laplace", {}), ("System_Cuspy_Laplace", {})])
@pytest.mark.parametrize("host", [False, True])
def test_kicks_only(fq, oracle, monkeypatch, cls, extra, host):
    if host:
        monkeypatch.setenv("FQSB_HOST_QUASISTATIC", "1")
    extra = dict(k_interactions=1.0, **extra)
    a = _make(fq, "Line1d", cls, 700, extra)
    b = _make(oracle, "Line1d", cls, 700, extra)
    q = a.quasistatic(1e-3, 12, kicks_only=True)
    ref = kicks_protocol(b, 12)
    assert_same_events(q, *ref, what=cls)
    assert_same_state(a, b, "kicks only")


@pytest.mark.parametrize("route", ["host", "tiled", "streaming"])
def test_host_driven_routes(fq, oracle, monkeypatch, route):
    """The event-by-event route (systems without a driver kernel: large
    systems on the tiled / streaming kernels) gives the oracle's events."""
    if route == "host":
        monkeypatch.setenv("FQSB_HOST_QUASISTATIC", "1")
    else:
        monkeypatch.setenv("FQSB_FORCE_STREAMING", "1")
        if route == "streaming":
            monkeypatch.setenv("FQSB_NO_TILES", "1")
    extra = dict(k_interactions=1.0)
    a = _make(fq, "Line1d", "System_Cuspy_Laplace", 800, extra)
    b = _make(oracle, "Line1d", "System_Cuspy_Laplace", 800, extra)
    q = a.quasistatic(1e-3, 15)
    ref = P.quasistatic(b, 15)
    assert_same_events(q, *ref, what=route)
    assert_same_state(a, b, route)


def test_nopassing(fq, oracle):
    par = _base(800, k_interactions=1.0)
    for k in ("m", "eta", "dt"):
        par.pop(k)
    a = fq.Line1d.System_Cuspy_Laplace_Nopassing(**par)
    b = oracle.Line1d.System_Cuspy_Laplace_Nopassing(**par)
    q = a.quasistatic(1e-3, 15)
    ref = P.quasistatic(b, 15)
    assert_same_events(q, *ref)
    assert_same_state(a, b, "no-passing")


def test_batch_replicas(fq, oracle):
    """Every replica runs its own event sequence at its own pace; each equals
    a single oracle system with its seed."""
    from frictionqpotspringblock_amd.batch import Batch

    R, N = 4, 512
    kw = dict(
        potential="Cuspy",
        interactions="Laplace1d",
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=3,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=60,
    )
    batch = Batch(R, seed_stride=1, **kw)
    q = batch.quasistatic(1e-3, 25)
    assert q["S"].shape == (R, 25)
    for r in range(R):
        o = oracle.Line1d.System_Cuspy_Laplace(**_base(N, k_interactions=1.0, seed=3 + r))
        ref = P.quasistatic(o, 25)
        assert_same_events({k: v[r] for k, v in q.items()}, *ref, what=f"replica {r}")
        assert np.array_equal(batch.u[r], o.u)
        assert np.array_equal(batch.index_at_align[r], o.chunk.index_at_align)


@pytest.mark.parametrize(
    "shape,m,block",
    [((8, 128), 1.0, None), ((4, 256), 1.0, None), ((16, 64), 1.0, None), ((8, 128), 2.0, None), ((4, 256), 1.0, "500")],
)
def test_batch_launch_shapes(fq, oracle, monkeypatch, shape, m, block):
    """Replicas of 1024 particles (every thread holds its full share) under
    the launch shapes the quasistatic workload tries, with unit and other
    mass, and in launches of 500 steps (the cached bound of the next well
    read again at every relaunch); each replica equals a single oracle system
    with its seed."""
    from frictionqpotspringblock_amd.batch import Batch

    if block is not None:
        monkeypatch.setenv("FQSB_QS_BLOCK", block)

    R, N = 2, 1024
    par = _base(N, k_interactions=1.0, m=m)
    for k in ("shape", "seed"):
        par.pop(k)
    batch = Batch(R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", shape=[N], seed=3, **par)
    batch.set_launch(*shape)
    assert batch.launch == shape
    q = batch.quasistatic(1e-3, 21)
    if block is not None:
        assert batch.quasistatic_stats[1] > 1  # relaunched
    for r in range(R):
        o = oracle.Line1d.System_Cuspy_Laplace(**_base(N, k_interactions=1.0, m=m, seed=3 + r))
        ref = P.quasistatic(o, 21)
        assert_same_events({k: v[r] for k, v in q.items()}, *ref, what=f"replica {r}")
        assert np.array_equal(batch.u[r], o.u)
        assert np.array_equal(batch.v[r], o.v)
        assert np.array_equal(batch.index_at_align[r], o.chunk.index_at_align)


def test_ragged(fq, oracle):
    from frictionqpotspringblock_amd.ragged import Ragged

    specs = [
        ("Line1d", "System_Cuspy_Laplace", 300, dict(k_interactions=1.0)),
        ("Line1d", "System_SemiSmooth_Laplace", 700, dict(k_interactions=1.0, kappa=1.0)),
        ("Line1d", "System_Cuspy_Quartic", 1500, dict(a1=1.0, a2=5.0)),
        ("Line2d", "System_Cuspy_Laplace", (16, 16), dict(k_interactions=1.0)),
        ("Particles", "System_SemiSmooth", 400, dict(kappa=1.0)),
    ]
    systems = []
    for i, (sub, cls, n, extra) in enumerate(specs):
        shape = list(n) if isinstance(n, tuple) else [n]
        systems.append((f"{sub}.{cls}", _base(int(np.prod(shape)), shape=shape, seed=10 + i, **extra)))
    rg = Ragged(systems)
    active = np.array([1, 1, 0, 1, 1], dtype=bool)
    q = rg.quasistatic(1e-3, 15, active=active)
    u = rg.u
    for i, (sub, cls, n, extra) in enumerate(specs):
        if not active[i]:
            assert np.all(q["S"][i] == 0) and np.all(q["inc"][i] == 0)
            continue
        shape = list(n) if isinstance(n, tuple) else [n]
        o = getattr(getattr(oracle, sub), cls)(**_base(int(np.prod(shape)), shape=shape, seed=10 + i, **extra))
        ref = P.quasistatic(o, 15)
        assert_same_events({k: v[i] for k, v in q.items()}, *ref, what=cls)
        assert np.array_equal(u[i], o.u)


def test_errors(fq):
    extra = dict(k_interactions=1.0)
    a = _make(fq, "Line1d", "System_Cuspy_Laplace", 200, extra)
    with pytest.raises(RuntimeError, match="eps"):
        a.quasistatic(0.0, 3)
    with pytest.raises(RuntimeError, match="direction"):
        a.quasistatic(1e-3, 3, direction=0)
    with pytest.raises(RuntimeError, match="No convergence found"):
        a.quasistatic(1e-3, 1, max_iter=5)
    s = _make(fq, "Line1d", "System_Smooth_Laplace", 200, extra)
    with pytest.raises(RuntimeError, match="Operation not possible"):
        s.quasistatic(1e-3, 3)
    assert s.quasistatic(1e-3, 1)["S"].shape == (1,)
    # systems with random forcing have no minimise, hence no driver
    assert not hasattr(fq.Line1d.System_Cuspy_Laplace_RandomForcing, "quasistatic")
    assert a.quasistatic(1e-3, 0)["S"].shape == (0,)


def test_batch_truncate_trigger_fixed_force(fq, oracle):
    """Batched minimise_truncate, trigger and advanceToFixedForce: replica r
    equals an oracle system with seed 3 + r doing the same operation with its
    own arguments."""
    from frictionqpotspringblock_amd.batch import Batch

    R, N = 3, 400
    kw = dict(
        potential="Cuspy",
        interactions="Laplace1d",
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=3,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=60,
    )
    b = Batch(R, seed_stride=1, **kw)
    o = [oracle.Line1d.System_Cuspy_Laplace(**_base(N, k_interactions=1.0, seed=3 + r)) for r in range(R)]

    def same(what):
        for r in range(R):
            assert np.array_equal(b.u[r], o[r].u), (what, r)
            assert np.array_equal(b.f[r], o[r].f), (what, r)
            assert np.array_equal(b.index_at_align[r], o[r].chunk.index_at_align), (what, r)
            assert b.inc[r] == o[r].inc and b.u_frame[r] == o[r].u_frame, (what, r)

    b.quasistatic(1e-3, 1)
    for s in o:
        s.u_frame = 0.0
        assert s.minimise() == 0
    same("quench")
    # advance to fixed mean frame forces (elastic)
    target = [float(np.mean(s.f_frame)) + 1e-4 * (r + 1) for r, s in enumerate(o)]
    b.advanceToFixedForce(target)
    for r, s in enumerate(o):
        s.advanceToFixedForce(target[r])
    same("advanceToFixedForce")
    # trigger a different particle in each replica, then a truncated minimise
    p = [5, 17, 123]
    i_n = b.index_at_align.copy()
    b.trigger(p, 1e-3)
    for r, s in enumerate(o):
        s.trigger(p[r], 1e-3)
    for r in range(R):
        assert np.array_equal(b.u[r], o[r].u)
    ret = b.minimise_truncate(i_n, A_truncate=3)
    for r, s in enumerate(o):
        assert ret[r] == s.minimise_truncate(i_n[r], A_truncate=3), r
    same("minimise_truncate")
    with pytest.raises(RuntimeError, match="out of range"):
        b.trigger([0, N, 0], 1e-3)


def _rf_kw(N):
    rng = np.random.default_rng(11)
    return dict(
        external="RandomNormalForcing",
        minimisation="None",
        mean=0.01,
        stddev=0.1,
        seed_forcing=5,
        dinc_init=rng.integers(0, 12, N),
        dinc=rng.integers(1, 12, N),
    )


@pytest.mark.parametrize("forcing", [False, True])
def test_batch_checkpoint_restart(fq, tmp_path, forcing):
    """A batch restored from a checkpoint (in place, and into a fresh batch
    from the .npz file) continues bit for bit: the same events, states and
    landscape windows (refills included: 40-well windows)."""
    from frictionqpotspringblock_amd.batch import Batch, save_events

    R, N = 3, 300
    kw = dict(
        potential="Cuspy",
        interactions="Laplace1d",
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=3,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=40,
    )
    if forcing:
        kw.update(_rf_kw(N))

    def advance(b):
        if forcing:
            b.flowSteps(700, 0.05)
            return None
        return b.quasistatic(1e-3, 16, first_event=11)

    b = Batch(R, seed_stride=1, **kw)
    if forcing:
        b.flowSteps(500, 0.05)
    else:
        b.quasistatic(1e-3, 11)
    ck = b.checkpoint()
    b.save(tmp_path / "ck.npz")
    qa = advance(b)
    after = b.checkpoint()
    b.restore(ck)
    qb = advance(b)
    fresh = Batch(R, seed_stride=1, **kw)
    fresh.load(tmp_path / "ck.npz")
    qc = advance(fresh)
    for other, q in ((b, qb), (fresh, qc)):
        got = other.checkpoint()
        for k, v in after.items():
            assert np.array_equal(got[k], v), k
        assert np.array_equal(other.f, b.f)
        if q is not None:
            for k in qa:
                assert np.array_equal(q[k], qa[k]), k
    assert not forcing or not np.array_equal(after["next"], ck["next"])
    assert forcing or np.any(after["start"] != ck["start"])  # windows moved
    if qa is not None:
        save_events(tmp_path / "events.npz", qa, system=1)
        with np.load(tmp_path / "events.npz") as f:
            assert np.array_equal(f["S"], qa["S"][1]) and np.array_equal(f["x_frame"], qa["u_frame"][1])


def test_restart_equals_uninterrupted_oracle(fq, oracle, tmp_path):
    """A batch that runs 11 events, is written to disk, and continues in a
    fresh process-level object from the file for 16 more events equals oracle
    systems (seeds 3, 4, 5) that ran the 27 events without stopping: the
    restart of the reference (chunk.restore and the state setters,
    test_Line1d.py:309-341) reproduces the uninterrupted run, refills of the
    40-well windows included; the event log is written as HDF5 in the
    examples' schema and read back."""
    from frictionqpotspringblock_amd import h5events
    from frictionqpotspringblock_amd.batch import Batch, load_events, save_events

    R, N = 3, 300
    par = dict(
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=3,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=40,
    )
    b = Batch(R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **par)
    q1 = b.quasistatic(1e-3, 11)
    b.save(tmp_path / "ck.npz")
    del b
    fresh = Batch(R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **par)
    fresh.load(tmp_path / "ck.npz")
    q2 = fresh.quasistatic(1e-3, 16, first_event=11)
    singles = [oracle.Line1d.System_Cuspy_Laplace(**dict(par, seed=3 + r)) for r in range(R)]
    u, idx = fresh.u, fresh.index_at_align
    for r, s in enumerate(singles):
        uf, ff, S, inc = P.quasistatic(s, 27)
        assert np.array_equal(np.concatenate([q1["S"][r], q2["S"][r]]), S), r
        assert np.array_equal(np.concatenate([q1["inc"][r], q2["inc"][r]]), inc), r
        assert np.array_equal(np.concatenate([q1["u_frame"][r], q2["u_frame"][r]]), uf), r
        assert np.array_equal(u[r], s.u), r
        assert np.array_equal(idx[r], s.chunk.index_at_align), r
    with np.load(tmp_path / "ck.npz") as f:
        assert np.any(fresh.checkpoint()["start"] != f["start"])  # windows moved after the restart
    try:
        h5events.library()
    except RuntimeError:
        return
    q = {k: np.concatenate([q1[k], q2[k]], axis=1) for k in q1}
    save_events(tmp_path / "events.h5", q, system=2)
    back = load_events(tmp_path / "events.h5")
    assert np.array_equal(back["S"], q["S"][2]) and np.array_equal(back["x_frame"], q["u_frame"][2])
