"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Per particle the device evaluates the reference's expressions in the same
order without FMA, so u, v, a, every force component and the landscape are
compared bit for bit (np.array_equal). The StopList decision uses norms reduced
in a different (fixed) order on the device; a decision could differ only if a
residual fell within rounding of tol, so increments are compared exactly too.
"""

import numpy as np
import pytest

from tests import protocols as P
from tests import reference_cases as RC

pytestmark = pytest.mark.gpu

FIELDS = ["u", "v", "a", "f", "f_potential", "f_interactions", "f_frame", "f_damping"]


@pytest.fixture(scope="module")
def fq():
    import frictionqpotspringblock_amd as fq

    return fq


@pytest.fixture(params=["resident", "tiled", "tiled_step", "streaming"])
def path(request, monkeypatch):
    """Run a test through the resident kernels, again with every system on
    the large-system path (FQSB_FORCE_STREAMING=1: the tiled kernels for
    timeSteps / flowSteps / minimise -- blocks of steps per launch for chains
    and lattices --, the one-workgroup streaming kernel otherwise), with the tiled
    path one step per launch (FQSB_TILE_K=0), and with the tiles switched off
    (FQSB_NO_TILES=1: the streaming kernel for everything)."""
    monkeypatch.delenv("FQSB_FORCE_STREAMING", raising=False)
    monkeypatch.delenv("FQSB_NO_TILES", raising=False)
    monkeypatch.delenv("FQSB_TILE_K", raising=False)
    monkeypatch.delenv("FQSB_NOPASSING_TILES", raising=False)
    if request.param in ("tiled", "tiled_step", "streaming"):
        monkeypatch.setenv("FQSB_FORCE_STREAMING", "1")
        monkeypatch.setenv("FQSB_NOPASSING_TILES", "1")  # no-passing systems: sweeps over tiles
    if request.param == "tiled_step":
        monkeypatch.setenv("FQSB_TILE_K", "0")
    if request.param == "streaming":
        monkeypatch.setenv("FQSB_NO_TILES", "1")
    return request.param


def assert_same_state(a, b, what=""):
    for name in FIELDS:
        x, y = getattr(a, name), getattr(b, name)
        assert np.array_equal(x, y), f"{what}: {name} differs, max |d| = {np.max(np.abs(x - y))}"
    assert np.array_equal(a.chunk.index_at_align, b.chunk.index_at_align), what
    assert a.inc == b.inc, (what, a.inc, b.inc)
    assert a.u_frame == b.u_frame, what


def _base(N, **kw):
    par = dict(
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=3,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=60,
    )
    par.update(kw)
    return par


# (submodule, class, size, extra constructor arguments)
SYSTEMS = [
    ("Line1d", "System_Cuspy_Laplace", 1000, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 64, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 1537, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 4096, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Laplace", 6001, dict(k_interactions=1.0)),
    ("Line1d", "System_SemiSmooth_Laplace", 700, dict(k_interactions=1.0, kappa=1.0)),
    ("Line1d", "System_Smooth_Laplace", 700, dict(k_interactions=1.0)),
    ("Line1d", "System_Cuspy_Quartic", 900, dict(a1=1.0, a2=5.0)),
    ("Line1d", "System_Cuspy_QuarticGradient", 900, dict(k2=1.0, k4=2.0)),
    ("Line1d", "System_Cuspy_LongRange", 300, dict(k_interactions=1.0, alpha=1.5)),
    ("Line2d", "System_Cuspy_Laplace", (30, 40), dict(k_interactions=1.0)),
    ("Line2d", "System_Cuspy_QuarticGradient", (20, 25), dict(k2=1.0, k4=1.0)),
    ("Particles", "System_Cuspy", 700, {}),
    ("Particles", "System_SemiSmooth", 700, dict(kappa=1.0)),
    ("Particles", "System_Smooth", 700, {}),
]


def _make(mod, sub, cls, n, extra, **kw):
    shape = list(n) if isinstance(n, tuple) else [n]
    size = int(np.prod(shape))
    par = _base(size, shape=shape, **extra)
    par.update(kw)
    return getattr(getattr(mod, sub), cls)(**par)


@pytest.mark.parametrize("sub,cls,n,extra", SYSTEMS, ids=[f"{s[1]}-{s[0]}-{s[2]}" for s in SYSTEMS])
def test_dynamics_and_quasistatics(fq, oracle, path, sub, cls, n, extra):
    a = _make(fq, sub, cls, n, extra)
    b = _make(oracle, sub, cls, n, extra)
    assert_same_state(a, b, "construction")
    assert np.array_equal(a.chunk.data, b.chunk.data)
    assert np.array_equal(a.chunk.start, b.chunk.start)
    # large initial velocities: particles cross tens of wells in both
    # directions, which exercises the chunk refills
    # (nonlinear interactions get a gentler kick: a cubic spring force at
    # dt = 0.1 is unstable for neighbour differences of several wells)
    scale = 1.0 if "Quartic" in cls else 10.0
    v0 = np.random.default_rng(7).normal(0.0, scale, a.shape)
    for s in (a, b):
        s.v = v0
    assert_same_state(a, b, "set_v")
    for s in (a, b):
        s.flowSteps(300, 0.2)
    assert_same_state(a, b, "flowSteps")
    assert np.array_equal(a.chunk.data, b.chunk.data), "landscape refill differs"
    for s in (a, b):
        s.timeSteps(7)
    assert_same_state(a, b, "timeSteps")
    ra = a.minimise()
    rb = b.minimise()
    assert ra == rb == 0
    assert_same_state(a, b, "minimise")
    if not cls.startswith("System_Smooth"):  # Smooth: maxUniformDisplacement raises
        for kick in (False, True):
            da = a.eventDrivenStep(1e-3, kick)
            db = b.eventDrivenStep(1e-3, kick)
            assert da == db
            assert_same_state(a, b, f"eventDrivenStep kick={kick}")
        assert a.minimise(time_activity=True) == b.minimise(time_activity=True) == 0
        assert a.quasistaticActivityFirst == b.quasistaticActivityFirst
        assert a.quasistaticActivityLast == b.quasistaticActivityLast
        assert_same_state(a, b, "minimise(time_activity)")
    assert np.array_equal(a.chunk.data, b.chunk.data)


def test_until_event_and_truncate(fq, oracle, path):
    N = 500
    a = _make(fq, "Line1d", "System_Cuspy_Laplace", N, dict(k_interactions=1.0))
    b = _make(oracle, "Line1d", "System_Cuspy_Laplace", N, dict(k_interactions=1.0))
    for s in (a, b):
        s.minimise()
        s.eventDrivenStep(1e-3, False)
        s.eventDrivenStep(1e-3, True)
    assert a.timeStepsUntilEvent() == b.timeStepsUntilEvent()
    assert_same_state(a, b, "timeStepsUntilEvent")
    i_n = a.chunk.index_at_align.copy()
    assert a.minimise_truncate(i_n=i_n, A_truncate=3) == b.minimise_truncate(i_n=i_n, A_truncate=3)
    assert_same_state(a, b, "minimise_truncate A")
    assert a.minimise_truncate(i_n=i_n, S_truncate=20) == b.minimise_truncate(i_n=i_n, S_truncate=20)
    assert_same_state(a, b, "minimise_truncate S")
    assert a.quasistaticActivityFirst == b.quasistaticActivityFirst
    assert a.quasistaticActivityLast == b.quasistaticActivityLast
    for s in (a, b):
        s.advanceToFixedForce(0.05, allow_plastic=True)
    assert_same_state(a, b, "advanceToFixedForce")
    for s in (a, b):
        s.trigger(7, 1e-3)
    assert np.array_equal(a.u, b.u)


@pytest.mark.parametrize("N,tiles", [(800, False), (800, True), (20000, False), (65536, False)])
def test_nopassing(fq, oracle, monkeypatch, N, tiles):
    """The no-passing minimiser (detail.h:1694-1753) on the resident kernel
    (N <= 8192) and on sweeps over tiles of 4096 particles (N = 20000 and
    65536, and N = 800 forced onto the tiles), bit for bit, through events
    with refills of the 60-well windows."""
    monkeypatch.delenv("FQSB_NOPASSING_TILES", raising=False)
    if tiles:
        monkeypatch.setenv("FQSB_NOPASSING_TILES", "1")
    print(f"no-passing N={N} tiles={tiles or N > 8192}", flush=True)
    par = _base(N, k_interactions=1.0)
    for k in ("m", "eta", "dt"):
        par.pop(k)
    a = fq.Line1d.System_Cuspy_Laplace_Nopassing(**par)
    b = oracle.Line1d.System_Cuspy_Laplace_Nopassing(**par)
    for step in range(8):
        for s in (a, b):
            if step:
                s.eventDrivenStep(1e-3 if step < 6 else 0.5, step % 2 == 0)
            assert s.minimise() == 0
        assert_same_state(a, b, f"no-passing step {step}")
        assert np.array_equal(a.chunk.data, b.chunk.data), step
        assert np.array_equal(a.chunk.start, b.chunk.start), step
    # max_iter is honoured (the reference's error and return value; the
    # StopList needs niter_tol = 10 residuals, so 3 sweeps never converge)
    for s in (a, b):
        s.eventDrivenStep(0.5, True)
        with pytest.raises(RuntimeError, match="No convergence found"):
            s.minimise(max_iter=3)
    assert_same_state(a, b, "no-passing after max_iter")
    assert a.minimise(max_iter=3, max_iter_is_error=False) == b.minimise(max_iter=3, max_iter_is_error=False) == 4
    assert_same_state(a, b, "no-passing after max_iter, no error")


@pytest.mark.parametrize("N,steps", [(2500, 100), (8192, 12)])
def test_longrange_tiled(fq, oracle, N, steps):
    """System_Cuspy_LongRange beyond 2048 particles steps over tiles of 256
    particles, each thread summing its particle's long-range force over the
    whole chain in the reference's order (detail.h:849-867): bit for bit
    through flowSteps, minimise and a kick + minimise (N = 2500), and through
    timeSteps at the long-range configuration's N = 8192."""
    print(f"long range tiled N={N}", flush=True)
    extra = dict(k_interactions=1.0, alpha=1.5)
    a = _make(fq, "Line1d", "System_Cuspy_LongRange", N, extra)
    b = _make(oracle, "Line1d", "System_Cuspy_LongRange", N, extra)
    v0 = np.random.default_rng(0).normal(size=N)
    for s in (a, b):
        s.v = v0
        s.flowSteps(steps, 0.05)
    assert_same_state(a, b, "long range flowSteps")
    if N > 3000:
        return
    for s in (a, b):
        assert s.minimise() == 0
    assert_same_state(a, b, "long range minimise")
    for s in (a, b):
        s.eventDrivenStep(1e-3, True)
        assert s.minimise() == 0
    assert_same_state(a, b, "long range kick + minimise")


@pytest.mark.parametrize("cls", ["System_Cuspy_Laplace_RandomForcing", "System_Cuspy_Quartic_RandomForcing"])
def test_random_forcing(fq, oracle, path, cls):
    N = 700
    extra = dict(k_interactions=1.0) if "Laplace" in cls else dict(a1=1.0, a2=5.0)
    rng = np.random.default_rng(0)
    kw = dict(
        mean=0.01,
        stddev=0.1,
        seed_forcing=5,
        dinc_init=rng.integers(0, 12, N),
        dinc=rng.integers(1, 12, N),
        **extra,
    )
    a = _make(fq, "Line1d", cls, N, kw)
    b = _make(oracle, "Line1d", cls, N, kw)
    assert_same_state(a, b, "construction")
    for s in (a, b):
        s.flowSteps(250, 0.1)
    assert_same_state(a, b, "flowSteps with random forcing")
    assert np.array_equal(a.external.f_thermal, b.external.f_thermal)
    assert np.array_equal(a.external.next, b.external.next)
    assert a.external.state == b.external.state
    for s in (a, b):
        s.inc = s.inc + 5
    assert_same_state(a, b, "set_inc")
    assert not hasattr(a, "minimise")


def test_particles_random_forcing(fq, oracle, path):
    N = 600
    rng = np.random.default_rng(3)
    kw = dict(mean=0.01, stddev=0.1, seed_forcing=5, dinc_init=rng.integers(0, 12, N), dinc=rng.integers(1, 12, N))
    a = _make(fq, "Particles", "System_Cuspy_RandomForcing", N, kw)
    b = _make(oracle, "Particles", "System_Cuspy_RandomForcing", N, kw)
    assert_same_state(a, b, "construction")
    for s in (a, b):
        s.flowSteps(250, 0.1)
    assert_same_state(a, b, "flowSteps with random forcing")
    assert np.array_equal(a.external.f_thermal, b.external.f_thermal)
    assert not hasattr(a, "minimise") and not hasattr(a, "eventDrivenStep")


def test_particles_example_driver(fq):
    """examples/Particles_Cuspy.py in full (2000 events) through the
    quasistatic driver against the reference's stored output."""
    ref = P.golden("Particles_Cuspy")
    a = P.example_system(fq, "Particles_Cuspy")
    q = a.quasistatic(1e-3, len(ref["S"]))
    assert np.array_equal(q["S"], ref["S"])
    assert np.allclose(q["u_frame"], ref["x_frame"])
    assert np.allclose(q["f_frame"], ref["f_frame"])


def test_reference_random_forcing(fq, oracle):
    RC.random_forcing_draws(fq, oracle)


@pytest.mark.parametrize("case", RC.ALL, ids=[f.__name__ for f in RC.ALL])
def test_reference_case(fq, case):
    case(fq)


@pytest.mark.parametrize(
    "name,nstep",
    [
        ("Line1d_Cuspy_Laplace", 200),
        ("Line1d_Cuspy_Quartic", 120),
        ("Line1d_SemiSmooth_Laplace", 200),
        ("Line1d_Cuspy_Laplace_Nopassing", 400),
        ("Line2d_Cuspy_Laplace", 60),
        ("Line1d_Cuspy_Laplace_LongRange", 6),
        ("Particles_Cuspy", 400),
    ],
)
def test_golden_examples(fq, oracle, path, name, nstep):
    ref = P.golden(name)
    a = P.example_system(fq, name)
    ua, fa, Sa, ia = P.quasistatic(a, nstep)
    assert np.array_equal(Sa, ref["S"][:nstep])
    assert np.allclose(ua, ref["x_frame"][:nstep])
    assert np.allclose(fa, ref["f_frame"][:nstep])
    b = P.example_system(oracle, name)
    ub, fb, Sb, ib = P.quasistatic(b, nstep)
    assert np.array_equal(ia, ib)
    assert np.array_equal(ua, ub)
    assert_same_state(a, b, name)


def test_landscape_far_moves(fq, oracle):
    for dist, par in [("weibull", [2.0, 1.1, 1e-3]), ("weibull", [3.0, 1.0]), ("random", [2.0]), ("exponential", [0.5])]:
        p = P.delta_par(4, distribution=dist, parameters=par, offset=-30.0, nchunk=40, seed=9)
        a = fq.Line1d.System_Cuspy_Laplace(**p)
        b = oracle.Line1d.System_Cuspy_Laplace(**p)
        assert np.array_equal(a.chunk.data, b.chunk.data), dist
        for u in [np.array([3.0, 50.0, 400.0, -70.0]), np.array([-100.0, 20.0, 10.0, 0.5]), np.zeros(4)]:
            a.u = u
            b.u = u
            assert np.array_equal(a.chunk.data, b.chunk.data), (dist, u)
            assert np.array_equal(a.chunk.start, b.chunk.start)
            assert np.array_equal(a.chunk.index_at_align, b.chunk.index_at_align)
            assert np.array_equal(a.f_potential, b.f_potential)
        st = a.chunk.start.copy() - 7
        state = b.chunk.state_at(st)
        assert np.array_equal(a.chunk.state_at(st), state)
        value = b.chunk.data[..., 0] - 1.0
        for s in (a, b):
            s.chunk.restore(state=state, value=value, index=st)
            s.u = np.array([-40.0, 3.0, 30.0, 0.0])
        assert np.array_equal(a.chunk.data, b.chunk.data)
        assert np.array_equal(a.chunk.index_at_align, b.chunk.index_at_align)


def test_batch_equals_single(fq, oracle):
    from frictionqpotspringblock_amd.batch import Batch

    R, N = 5, 1024
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
        seed=0,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=200,
    )
    batch = Batch(R, seed_stride=1, **kw)
    batch.flowSteps(2000, 1e-3, block=700)
    single = []
    for r in range(R):
        par = _base(N, k_interactions=1.0, seed=r, nchunk=200)
        s = fq.Line1d.System_Cuspy_Laplace(**par)
        s.flowSteps(2000, 1e-3)
        single.append(s)
        assert np.array_equal(batch.u[r], s.u)
        assert np.array_equal(batch.v[r], s.v)
        assert np.array_equal(batch.f[r], s.f)
        assert np.array_equal(batch.index_at_align[r], s.chunk.index_at_align)
        assert batch.u_frame[r] == s.u_frame
    o = oracle.Line1d.System_Cuspy_Laplace(**_base(N, k_interactions=1.0, seed=2, nchunk=200))
    o.flowSteps(2000, 1e-3)
    assert np.array_equal(batch.u[2], o.u)
    # batched minimise equals single minimise
    assert np.all(batch.minimise() == 0)
    for r in range(R):
        assert single[r].minimise() == 0
        assert np.array_equal(batch.u[r], single[r].u)
        assert batch.inc[r] == single[r].inc
    # batched event-driven step equals single
    du = batch.eventDrivenStep(1e-3, False)
    for r in range(R):
        assert du[r] == single[r].eventDrivenStep(1e-3, False)
        assert np.array_equal(batch.u[r], single[r].u)


def test_launch_shapes_agree(fq):
    from frictionqpotspringblock_amd.batch import Batch

    R, N = 3, 4096
    kw = dict(
        m=1.0,
        eta=0.3,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=0,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        nchunk=200,
    )
    ref = None
    for ppt, threads in [(4, 1024), (8, 512), (16, 256), (8, 1024), (16, 512)]:
        b = Batch(R, **kw)
        b.set_launch(ppt, threads)
        assert b.launch == (ppt, threads)
        b.flowSteps(500, 1e-2)
        if ref is None:
            ref = b.u
        assert np.array_equal(b.u, ref), (ppt, threads)


@pytest.mark.parametrize("N", [4096, 4032, 4090])
def test_flow_full_threads(fq, oracle, N):
    """timeSteps / flowSteps with 4 particles per thread on 1024 threads
    against the oracle: N = 4096 (every thread full: the compile-time last
    register), 4032 (full, with 16 threads that hold no particle), 4090 (a
    partly filled thread: the general form)."""
    from frictionqpotspringblock_amd.batch import Batch

    R = 2
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
        seed=0,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=200,
    )
    b = Batch(R, seed_stride=1, **kw)
    b.set_launch(4, 1024)
    v0 = np.random.default_rng(3).normal(0.0, 2.0, (R, N))
    b.v = v0
    b.flowSteps(600, 1e-2)
    b.timeSteps(100)
    for r in range(R):
        o = oracle.Line1d.System_Cuspy_Laplace(**_base(N, k_interactions=1.0, seed=r, nchunk=200))
        o.v = v0[r]
        o.flowSteps(600, 1e-2)
        o.timeSteps(100)
        assert np.array_equal(b.u[r], o.u)
        assert np.array_equal(b.v[r], o.v)
        assert np.array_equal(b.a[r], o.a)
        assert np.array_equal(b.f[r], o.f)
        assert np.array_equal(b.index_at_align[r], o.chunk.index_at_align)
        assert b.u_frame[r] == o.u_frame


def test_state_views_write_back(fq, oracle):
    """u, v, a behave like the reference's views for writes: item and slice
    assignments (also augmented, also through a kept reference) reach the
    system; the oracle gets the same values by whole-array assignment."""
    par = _base(300, k_interactions=1.0)
    a = fq.Line1d.System_Cuspy_Laplace(**par)
    b = oracle.Line1d.System_Cuspy_Laplace(**par)
    a.u[3] = a.u[3] + 0.7
    a.u[10:20] += 0.05
    v = a.v
    v[7] = 0.3
    v[100:110] -= 0.2
    a.a[5] = 0.01
    u, vb, ab = b.u, b.v, b.a
    u[3] += 0.7
    u[10:20] += 0.05
    vb[7] = 0.3
    vb[100:110] -= 0.2
    ab[5] = 0.01
    b.u, b.v, b.a = u, vb, ab
    assert np.array_equal(a.u, b.u) and np.array_equal(a.v, b.v) and np.array_equal(a.a, b.a)
    for s in (a, b):
        s.timeSteps(50)
    assert_same_state(a, b, "after writes through views")
    assert type(a.u + 1.0) is np.ndarray


def test_errors(fq):
    s = fq.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.full(5, np.nan)
    with pytest.raises(RuntimeError, match="NaN entries found"):
        s.timeStep()
    s = fq.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.array([0.3, 0, 0, 0, 0])
    with pytest.raises(RuntimeError, match="No convergence found"):
        s.minimise(max_iter=1)
    s = fq.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.array([0.3, 0, 0, 0, 0])
    assert s.minimise(max_iter=1, max_iter_is_error=False) == 2
    smooth = fq.Line1d.System_Smooth_Laplace(**P.delta_par(5))
    with pytest.raises(RuntimeError, match="Operation not possible."):
        smooth.maxUniformDisplacement()
    with pytest.raises(RuntimeError, match="Deprecated"):
        s.x
    with pytest.raises(RuntimeError, match="Deprecated"):
        s.f_neighbours


@pytest.mark.parametrize("start", ["nan", "inf_velocity"])
def test_nan_every_path(fq, oracle, path, start):
    """A NaN (in u from the start, or grown from an infinite velocity) stops
    timeSteps with "NaN entries found" after the same step as the oracle, on
    every path (for the blocked tiles: the committed prefix, then the NaN
    step)."""
    systems = [_make(m, "Line1d", "System_Cuspy_Laplace", 300, dict(k_interactions=1.0)) for m in (fq, oracle)]
    for s in systems:
        if start == "nan":
            u = s.u.copy()
            u[7] = np.nan
            s.u = u
        else:
            v = s.v.copy()
            v[150] = np.inf
            s.v = v
        with pytest.raises(RuntimeError, match="NaN entries found"):
            s.timeSteps(200)
    assert systems[0].inc == systems[1].inc > 0


# systems beyond the capacity of a resident workgroup (streaming path)
LARGE = [
    ("Line1d", "System_Cuspy_Laplace", 20000, dict(k_interactions=1.0), True),
    ("Line1d", "System_SemiSmooth_Laplace", 12000, dict(k_interactions=1.0, kappa=1.0), True),
    ("Line1d", "System_Cuspy_Quartic", 9000, dict(a1=1.0, a2=5.0), True),
    ("Line2d", "System_Cuspy_Laplace", (128, 128), dict(k_interactions=1.0), True),
    ("Line2d", "System_Cuspy_QuarticGradient", (100, 100), dict(k2=1.0, k4=1.0), False),
    ("Line1d", "System_Cuspy_LongRange", 7000, dict(k_interactions=1.0, alpha=1.5), False),
]


@pytest.mark.parametrize("sub,cls,n,extra,minimise", LARGE, ids=[f"{s[1]}-{s[0]}-{s[2]}" for s in LARGE])
def test_large_systems(fq, oracle, sub, cls, n, extra, minimise):
    a = _make(fq, sub, cls, n, extra)
    b = _make(oracle, sub, cls, n, extra)
    assert_same_state(a, b, "construction")
    scale = 1.0 if "Quartic" in cls else 10.0
    v0 = np.random.default_rng(7).normal(0.0, scale, a.shape)
    for s in (a, b):
        s.v = v0
    nsteps = 30 if "LongRange" in cls else 200
    for s in (a, b):
        s.flowSteps(nsteps, 0.2)
    assert_same_state(a, b, "flowSteps")
    assert np.array_equal(a.chunk.data, b.chunk.data)
    if minimise:
        assert a.minimise() == b.minimise() == 0
        assert_same_state(a, b, "minimise")
        for s in (a, b):
            s.eventDrivenStep(1e-3, True)
        assert a.minimise(time_activity=True) == b.minimise(time_activity=True) == 0
        assert a.quasistaticActivityFirst == b.quasistaticActivityFirst
        assert_same_state(a, b, "kick + minimise")


@pytest.mark.parametrize("k", ["1", "7", "256"])
def test_tile_blocks(fq, oracle, monkeypatch, k):
    """Blocks of K steps per launch on a large chain (halo recomputation,
    redo after the StopList or a refill inside a block, the commit of the
    prefix before a refill) against the oracle: flowSteps with crossings and
    refills (60-well windows), minimise, kick + minimise."""
    monkeypatch.setenv("FQSB_TILE_K", k)
    par = _base(20000, k_interactions=1.0, nchunk=60)
    a = fq.Line1d.System_Cuspy_Laplace(**par)
    b = oracle.Line1d.System_Cuspy_Laplace(**par)
    v0 = np.random.default_rng(5).normal(0.0, 10.0, a.shape)
    for s in (a, b):
        s.v = v0
        s.flowSteps(300, 0.2)
    assert_same_state(a, b, "flowSteps")
    assert np.array_equal(a.chunk.data, b.chunk.data)
    assert a.minimise() == b.minimise() == 0
    assert_same_state(a, b, "minimise")
    for s in (a, b):
        s.eventDrivenStep(1e-3, False)
        s.eventDrivenStep(1e-3, True)
    assert a.minimise() == b.minimise() == 0
    assert_same_state(a, b, "kick + minimise")
    for s in (a, b):
        s.timeSteps(45)
    assert_same_state(a, b, "timeSteps")


LATTICE_BLOCKS = [
    ("System_Cuspy_Laplace", (96, 80), dict(k_interactions=1.0), "4", None),
    ("System_Cuspy_Laplace", (96, 80), dict(k_interactions=1.0), "1", None),
    ("System_Cuspy_Laplace", (70, 130), dict(k_interactions=1.0), "9", None),
    ("System_Cuspy_Laplace", (64, 72), dict(k_interactions=1.0), "8", "3"),  # blocks shorter than the halos
    ("System_Cuspy_Laplace", (5, 7), dict(k_interactions=1.0), "4", None),  # halos repeat sites
    ("System_Cuspy_QuarticGradient", (64, 72), dict(k2=1.0, k4=1.0), "4", None),
]


@pytest.mark.parametrize("cls,shape,extra,halo,k", LATTICE_BLOCKS,
                         ids=[f"{c[0]}-{c[1][0]}x{c[1][1]}-K{c[3]}-{c[4]}" for c in LATTICE_BLOCKS])
def test_lattice_tile_blocks(fq, oracle, monkeypatch, cls, shape, extra, halo, k):
    """Blocks of steps over 2-d tiles of lattices (halo rows and columns of
    `halo` sites, blocks of min(`k`, halo) steps) against the oracle:
    flowSteps with crossings and refills, minimise, kick + minimise,
    timeSteps."""
    monkeypatch.setenv("FQSB_FORCE_STREAMING", "1")
    monkeypatch.setenv("FQSB_LATTICE_K", halo)
    if k is not None:
        monkeypatch.setenv("FQSB_TILE_K", k)
    n = int(np.prod(shape))
    print(f"lattice blocks: {cls} {shape} halo {halo} block {k}", flush=True)
    par = _base(n, shape=list(shape), **extra)
    a = getattr(fq.Line2d, cls)(**par)
    b = getattr(oracle.Line2d, cls)(**par)
    scale = 1.0 if "Quartic" in cls else 10.0
    v0 = np.random.default_rng(5).normal(0.0, scale, a.shape)
    for s in (a, b):
        s.v = v0
        s.flowSteps(150, 0.2)
    assert_same_state(a, b, "flowSteps")
    assert np.array_equal(a.chunk.data, b.chunk.data)
    assert a.minimise() == b.minimise() == 0
    assert_same_state(a, b, "minimise")
    for s in (a, b):
        s.eventDrivenStep(1e-3, False)
        s.eventDrivenStep(1e-3, True)
    assert a.minimise() == b.minimise() == 0
    assert_same_state(a, b, "kick + minimise")
    for s in (a, b):
        s.timeSteps(45)
    assert_same_state(a, b, "timeSteps")


def test_tiled_graphs_match_eager(fq, monkeypatch):
    """Blocks of tiled steps replayed from captured graphs give the results of
    the kernels launched one by one, call after call (with host work and
    synchronisations between the replays)."""
    par = _base(20000, k_interactions=1.0, nchunk=60)
    monkeypatch.setenv("FQSB_NO_GRAPHS", "1")
    eager = fq.Line1d.System_Cuspy_Laplace(**par)
    monkeypatch.delenv("FQSB_NO_GRAPHS")
    graph = fq.Line1d.System_Cuspy_Laplace(**par)
    v0 = np.random.default_rng(3).normal(0.0, 5.0, 20000)
    for s in (eager, graph):
        s.v = v0
    for rep in range(3):
        for s in (eager, graph):
            s.flowSteps(700, 0.05)  # blocks of 256, 256, 188 steps: two graphs, replayed
        assert_same_state(eager, graph, f"flowSteps {rep}")
    for s in (eager, graph):
        s.eventDrivenStep(1e-3, True)
        assert s.minimise() == 0
    assert_same_state(eager, graph, "minimise")


def test_c0_quasistatic_protocol(fq, oracle):
    """The reference's CPU configuration: a 1024-particle Cuspy-Laplace chain
    (Weibull k = 2 yield increments, offset -100, chunks of 200 wells), 100
    quasistatic events of the examples' protocol (examples/Line1d_Cuspy_
    Laplace.py:39-61). Avalanche sizes, increments and frame positions are
    exact; the final state is bit-identical."""
    par = dict(
        m=1.0,
        eta=2.0 * np.sqrt(3.0) * 0.1,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / 1024,
        dt=0.1,
        shape=[1024],
        seed=0,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=200,
    )
    a = fq.Line1d.System_Cuspy_Laplace(**par)
    b = oracle.Line1d.System_Cuspy_Laplace(**par)
    ua, fa, Sa, ia = P.quasistatic(a, 201)
    ub, fb, Sb, ib = P.quasistatic(b, 201)
    assert np.array_equal(Sa, Sb)
    assert np.array_equal(ia, ib)
    assert np.array_equal(ua, ub)
    np.testing.assert_allclose(fa, fb, rtol=1e-12, atol=1e-14)
    assert np.sum(np.abs(Sa)) > 0
    assert_same_state(a, b, "after 100 events")
