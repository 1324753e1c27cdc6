"""The reference's unit tests (tests/test_Line1d.py, tests/test_Line2d.py),
restated as functions of a module exposing the reference's API, so they run
against the oracle on the CPU and against the product on the GPU.

They use the reference's delta landscape (yield positions at half integers)
so that every expected value is a closed form. Random coefficients come from
a seeded generator instead of the reference's time-based seed.
"""

import numpy as np

from tests.protocols import delta_par


def _rng(seed=0):
    return np.random.default_rng(seed)


def uniform_init(mod):
    """test_Line1d.py:25-86: every class starts at zero force."""
    rng = _rng(1)
    N = 5
    par = delta_par(
        N,
        eta=float(rng.random()),
        mu=float(rng.random()),
        k_interactions=float(rng.random()),
        k_frame=float(rng.random()),
    )
    rand = dict(
        mean=0, stddev=1, seed_forcing=0, dinc_init=np.ones(N, dtype=np.int64), dinc=np.ones(N, dtype=np.int64)
    )
    L = mod.Line1d
    systems = [
        L.System_Cuspy_Laplace(**par),
        L.System_SemiSmooth_Laplace(kappa=1, **par),
        L.System_Smooth_Laplace(**par),
        L.System_Cuspy_LongRange(alpha=1, **par),
        L.System_Cuspy_Laplace_RandomForcing(**par, **rand),
    ]
    p2 = dict(par)
    p2.pop("m")
    p2.pop("dt")
    eta = p2.pop("eta")
    systems += [L.System_Cuspy_Laplace_Nopassing(eta=eta, **p2)]
    p3 = dict(par)
    p3.pop("k_interactions")
    systems += [L.System_Cuspy_Quartic(a1=1, a2=1, **p3)]
    systems += [L.System_Cuspy_QuarticGradient(k2=1, k4=1, **p3)]
    for system in systems:
        assert system.residual < 1e-5
        assert np.allclose(system.f, 0.0)
        assert np.allclose(system.f_potential, 0.0)
        assert np.allclose(system.f_frame, 0.0)
        assert np.allclose(system.f_interactions, 0.0)
        assert np.allclose(system.f_damping, 0.0)
        assert np.all(system.chunk.index_at_align + 1 == np.argmax(system.chunk.data[0, :] > 0))
        assert np.all(system.chunk.right_of_align > 0)
        assert np.all(system.chunk.left_of_align <= 0)


def cuspy_laplace_forces(mod):
    """test_Line1d.py:89-176"""
    rng = _rng(2)
    N = 5
    eta, mu, k_interactions, k_frame = (float(x) for x in rng.random(4))
    system = mod.Line1d.System_Cuspy_Laplace(
        **delta_par(N, eta=eta, mu=mu, k_interactions=k_interactions, k_frame=k_frame)
    )
    du = np.zeros(N)
    dv = np.zeros(N)
    du[0] = float(rng.random())
    dv[2] = float(rng.random())
    system.u = system.u + du
    system.v = system.v + dv
    umin = np.floor(du[0] + 0.5)
    f_potential = mu * np.array([umin - du[0], 0, 0, 0, 0])
    f_interactions = k_interactions * np.array([-2 * du[0], du[0], 0, 0, du[0]])
    f_frame = k_frame * np.array([-du[0], 0, 0, 0, 0])
    f_damping = eta * np.array([0, 0, -dv[2], 0, 0])
    assert np.all(system.chunk.right_of_align > system.u)
    assert np.all(system.chunk.left_of_align <= system.u)
    assert np.allclose(system.f_potential, f_potential)
    assert np.allclose(system.f_frame, f_frame)
    assert np.allclose(system.f_interactions, f_interactions)
    assert np.allclose(system.f_damping, f_damping)
    assert np.allclose(system.f, f_potential + f_frame + f_interactions + f_damping)

    du = np.zeros(N)
    dv = np.zeros(N)
    du[1] = 2.0 * float(rng.random())
    dv[3] = 2.0 * float(rng.random())
    system.u = system.u + du
    system.v = system.v + dv
    u = system.u
    v = system.v
    f_potential = mu * np.array([np.floor(u[0] + 0.5) - u[0], np.floor(u[1] + 0.5) - u[1], 0, 0, 0])
    f_interactions = k_interactions * np.array(
        [u[-1] - 2 * u[0] + u[1], u[0] - 2 * u[1] + u[2], u[1] - 2 * u[2] + u[3], 0, u[-2] - 2 * u[-1] + u[0]]
    )
    f_frame = k_frame * np.array([-u[0], -u[1], 0, 0, 0])
    f_damping = eta * np.array([0, 0, -v[2], -v[3], 0])
    assert np.allclose(system.f_potential, f_potential)
    assert np.allclose(system.f_frame, f_frame)
    assert np.allclose(system.f_interactions, f_interactions)
    assert np.allclose(system.f_damping, f_damping)
    assert np.allclose(system.f, f_potential + f_frame + f_interactions + f_damping)


def _event_steps(system, N):
    assert system.residual < 1e-5
    i_n = system.chunk.index_at_align
    system.eventDrivenStep(0.2, False)
    assert system.residual < 1e-5
    assert np.allclose(system.u, (0.5 - 0.1) * np.ones(N))
    assert np.all(system.chunk.index_at_align == i_n)
    assert np.isclose(system.u_frame, (0.5 - 0.1) * (1.0 + 0.1) / 0.1)
    i_n = system.chunk.index_at_align
    system.eventDrivenStep(0.2, True)
    assert np.allclose(system.u, (0.5 + 0.1) * np.ones(N))
    assert not np.all(system.chunk.index_at_align == i_n)
    assert np.isclose(system.u_frame, (0.5 + 0.1) * (1.0 + 0.1) / 0.1)
    i_n = system.chunk.index_at_align
    system.eventDrivenStep(0.2, False)
    assert np.allclose(system.u, (1.5 - 0.1) * np.ones(N))
    assert np.all(system.chunk.index_at_align == i_n)
    assert np.isclose(system.u_frame, (1.5 - 0.1) * (1.0 + 0.1) / 0.1)
    i_n = system.chunk.index_at_align
    system.eventDrivenStep(0.2, True)
    assert np.allclose(system.u, (1.5 + 0.1) * np.ones(N))
    assert not np.all(system.chunk.index_at_align == i_n)
    assert np.isclose(system.u_frame, (1.5 + 0.1) * (1.0 + 0.1) / 0.1)


def cuspy_laplace_event_driven(mod):
    """test_Line1d.py:178-219"""
    N = 3
    _event_steps(mod.Line1d.System_Cuspy_Laplace(**delta_par(N)), N)


def longrange_event_driven(mod):
    """test_Line1d.py:520-563"""
    N = 3
    _event_steps(mod.Line1d.System_Cuspy_LongRange(alpha=1, **delta_par(N)), N)


def cuspy_laplace_trigger(mod):
    """test_Line1d.py:221-245"""
    N = 3
    system = mod.Line1d.System_Cuspy_Laplace(**delta_par(N))
    i_n = system.chunk.index_at_align.copy()
    system.trigger(0, 0.2)
    u = np.zeros(N)
    u[0] = 0.5 + 0.1
    assert np.allclose(system.u, u)
    ret = system.minimise_truncate(i_n=i_n, A_truncate=1)
    assert np.sum(system.chunk.index_at_align != i_n) >= 1
    assert ret > 0
    assert system.residual > 1e-5


def cuspy_laplace_fixed_force(mod):
    """test_Line1d.py:247-275"""
    N = 3
    system = mod.Line1d.System_Cuspy_Laplace(**delta_par(N))
    assert system.residual < 1e-5
    system.advanceToFixedForce(0.1)
    assert np.isclose(np.mean(system.f_frame), 0.1)
    assert system.residual < 1e-5
    system.advanceToFixedForce(0.0)
    assert np.isclose(np.mean(system.f_frame), 0.0)
    assert np.allclose(system.u, 0.0)
    assert np.allclose(system.u_frame, 0.0)


def semismooth_event_driven(mod):
    """test_Line1d.py:344-399"""
    N = 3
    mu = 1
    kappa = 0.1
    system = mod.Line1d.System_SemiSmooth_Laplace(kappa=kappa, **delta_par(N, mu=mu))
    assert system.residual < 1e-5
    u0 = system.u.copy()
    uf0 = system.u_frame
    left = system.chunk.left_of_align
    right = system.chunk.right_of_align
    mid = 0.5 * (left + right)
    upper = (mu * mid + kappa * right) / (mu + kappa)
    lower = (mu * mid + kappa * left) / (mu + kappa)
    eps = 0.001
    assert np.isclose(system.maxUniformDisplacement(), np.min(upper - system.u))
    system.eventDrivenStep(eps=eps, kick=False)
    assert system.residual < 1e-5
    assert np.allclose(system.u, upper - 0.5 * eps)
    assert np.isclose(system.maxUniformDisplacement(), 0.5 * eps)
    system.eventDrivenStep(eps=eps, kick=True)
    assert system.residual > 1e-5
    assert np.allclose(system.u, upper + 0.5 * eps)
    assert np.isclose(system.maxUniformDisplacement(), 0)
    system.u = u0
    system.u_frame = uf0
    assert np.isclose(system.maxUniformDisplacement(-1), np.min(system.u - lower))
    system.eventDrivenStep(eps=eps, kick=False, direction=-1)
    assert system.residual < 1e-5
    assert np.allclose(system.u, lower + 0.5 * eps)
    assert np.isclose(system.maxUniformDisplacement(-1), 0.5 * eps)
    system.eventDrivenStep(eps=eps, kick=True, direction=-1)
    assert system.residual > 1e-5
    assert np.allclose(system.u, lower - 0.5 * eps)
    assert np.isclose(system.maxUniformDisplacement(), 0)


def quartic_interactions(mod):
    """test_Line1d.py:402-437"""
    rng = _rng(3)
    N = 10
    a1, a2 = (float(x) for x in rng.random(2))
    par = delta_par(N)
    par.pop("k_interactions")
    system = mod.Line1d.System_Cuspy_Quartic(a1=a1, a2=a2, **par)
    assert system.residual < 1e-5
    du = float(rng.random())
    u0 = np.array([du, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    laplace = np.array([-2 * du, du, 0, 0, 0, 0, 0, 0, 0, du])
    du_p = np.array([-du, 0, 0, 0, 0, 0, 0, 0, 0, du])
    du_n = np.array([-du, du, 0, 0, 0, 0, 0, 0, 0, 0])
    f0 = a1 * laplace + a2 * (du_p**3 + du_n**3)
    for i in range(N):
        u = np.roll(u0, i)
        system.u = u
        assert np.allclose(system.f_interactions, np.roll(f0, i))
        assert np.allclose(system.u, u)


def quartic_gradient_interactions(mod):
    """test_Line1d.py:440-471"""
    rng = _rng(4)
    N = 10
    k2, k4, du = (float(x) for x in rng.random(3))
    par = delta_par(N)
    par.pop("k_interactions")
    system = mod.Line1d.System_Cuspy_QuarticGradient(k2=k2, k4=k4, **par)
    assert system.residual < 1e-5
    u0 = np.array([du, 0, 0, 0, 0, 0, 0, 0, 0, 0])
    laplace = np.array([-2 * du, du, 0, 0, 0, 0, 0, 0, 0, du])
    gradient = np.array([0, 0.5 * du, 0, 0, 0, 0, 0, 0, 0, 0.5 * du])
    f0 = k2 * laplace + k4 * laplace * gradient**2
    for i in range(N):
        u = np.roll(u0, i)
        system.u = u
        assert np.allclose(system.f_interactions, np.roll(f0, i))
        assert np.allclose(system.u, u)


def longrange_interactions(mod):
    """test_Line1d.py:474-518"""
    N = 10
    k_interactions = 0.12
    system = mod.Line1d.System_Cuspy_LongRange(alpha=1, **delta_par(N, k_interactions=k_interactions))
    dp = np.arange(N)
    dn = np.arange(N)[::-1] + 1
    d = np.where(dp < dn, dp, dn)
    x = np.zeros(N)
    x[0] = 1
    system.u = x
    f = np.zeros_like(x)
    for j in range(1, N):
        f[j] = k_interactions * (x[0] - x[j]) / (d[j] ** 2)
    f[0] = -np.sum(f)
    for i in range(N):
        system.u = np.roll(x, i)
        assert np.allclose(np.roll(f, i), system.f_interactions)


def random_forcing_draws(mod, oracle):
    """test_Line1d.py:566-600 with the draw order checked against the oracle
    (the reference checks against prrng.pcg32(0).normal, which is absent)."""
    N = 10
    kw = dict(
        mean=0,
        stddev=1,
        seed_forcing=0,
        dinc_init=np.ones(N, dtype=np.int64),
        dinc=np.ones(N, dtype=np.int64),
    )
    a = mod.Line1d.System_Cuspy_Laplace_RandomForcing(**delta_par(N), **kw)
    b = oracle.Line1d.System_Cuspy_Laplace_RandomForcing(**delta_par(N), **kw)
    for s in (a, b):
        assert s.residual < 1e-5
    for _ in range(2):
        for s in (a, b):
            s.inc = s.inc + 1
            s.refresh()
            assert s.residual > 1e-5
        assert np.array_equal(a.external.f_thermal, b.external.f_thermal)
        assert a.external.state == b.external.state
        assert np.array_equal(a.external.next, b.external.next)


def line2d_laplace(mod):
    """test_Line2d.py:21-57"""
    rows, cols = 5, 4
    k_interactions = float(_rng(5).random())
    par = delta_par(1, k_interactions=k_interactions)
    par["shape"] = [rows, cols]
    system = mod.Line2d.System_Cuspy_Laplace(**par)
    assert system.residual < 1e-5
    assert list(system.shape) == [rows, cols]
    assert system.size == rows * cols
    c = -4
    f0 = np.array([[c, 1, 0, 1], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]]) * k_interactions
    u0 = np.array([[1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0]], dtype=float)
    for i in range(rows):
        for j in range(cols):
            x = np.roll(np.roll(u0, i, axis=0), j, axis=1)
            system.u = x
            f = np.roll(np.roll(f0, i, axis=0), j, axis=1)
            assert np.allclose(system.f_interactions, f)
            assert np.allclose(system.u, x)


def line2d_quartic_gradient(mod):
    """test_Line2d.py:60-173 (k4 = 0 reduces to the Laplacian; a random field
    is checked against the discretisation detail.h:735-748 evaluated in numpy)"""
    rows, cols = 5, 4
    par = delta_par(1)
    par.pop("k_interactions")
    par["shape"] = [rows, cols]
    k2 = 0.12
    system = mod.Line2d.System_Cuspy_QuarticGradient(k2=k2, k4=0.0, **par)
    u0 = np.zeros((rows, cols))
    u0[0, 0] = 1
    f0 = np.array([[-4, 1, 0, 1], [1, 0, 0, 0], [0, 0, 0, 0], [0, 0, 0, 0], [1, 0, 0, 0]]) * k2
    for i in range(rows):
        for j in range(cols):
            system.u = np.roll(np.roll(u0, i, axis=0), j, axis=1)
            assert np.allclose(system.f_interactions, np.roll(np.roll(f0, i, axis=0), j, axis=1))
    k4 = 0.7
    system = mod.Line2d.System_Cuspy_QuarticGradient(k2=k2, k4=k4, **par)
    u = _rng(6).uniform(-0.4, 0.4, (rows, cols))
    system.u = u
    up, um = np.roll(u, -1, 0), np.roll(u, 1, 0)
    ur, ul = np.roll(u, -1, 1), np.roll(u, 1, 1)
    lap = up + um + ur + ul - 4 * u
    dx = 0.5 * (up - um)
    dy = 0.5 * (ur - ul)
    dxy = 0.25 * (np.roll(up, -1, 1) - np.roll(up, 1, 1) - np.roll(um, -1, 1) + np.roll(um, 1, 1))
    dxx = up - 2 * u + um
    dyy = ur - 2 * u + ul
    f = lap * (k2 + k4 / 3) + 2 * k4 / 3 * (dx * dx * dxx + dy * dy * dyy + 2 * dx * dy * dxy)
    assert np.allclose(system.f_interactions, f)


ALL = [
    uniform_init,
    cuspy_laplace_forces,
    cuspy_laplace_event_driven,
    longrange_event_driven,
    cuspy_laplace_trigger,
    cuspy_laplace_fixed_force,
    semismooth_event_driven,
    quartic_interactions,
    quartic_gradient_interactions,
    longrange_interactions,
    line2d_laplace,
    line2d_quartic_gradient,
]
