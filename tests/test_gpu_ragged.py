"""Ragged sets on the GPU against separate systems of the CPU oracle.

A ragged set groups systems by kernel class and size class and runs the groups
concurrently; each operation must still apply to every system exactly as the
single-system call does (System::minimise, eventDrivenStep, timeSteps,
detail.h:1577-1960), so the comparison with the oracle is bit for bit.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CLASS = {
    # name -> (rank, potential, interactions, minimisation)
    "System_Cuspy_Laplace": (1, "Cuspy", "Laplace1d", 0),
    "System_Cuspy_Laplace_Nopassing": (1, "Cuspy", "Laplace1d", 1),
    "System_SemiSmooth_Laplace": (1, "SemiSmooth", "Laplace1d", 0),
    "System_Smooth_Laplace": (1, "Smooth", "Laplace1d", 0),
    "System_Cuspy_Quartic": (1, "Cuspy", "Quartic1d", 0),
    "System_SemiSmooth_Quartic": (1, "SemiSmooth", "Quartic1d", 0),
    "System_Smooth_Quartic": (1, "Smooth", "Quartic1d", 0),
    "Line2d.System_Cuspy_Laplace": (2, "Cuspy", "Laplace2d", 0),
}


def _par(N, seed, **kw):
    par = dict(
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
        nchunk=100,
    )
    par.update(kw)
    return par


def _systems(large=True):
    s = [
        ("System_Cuspy_Laplace", _par(700, 1)),
        ("System_Cuspy_Laplace", _par(1000, 2)),  # same group as the one above, other size
        ("System_Cuspy_Laplace", _par(130, 3, distribution="random", parameters=[2.0], offset=-50.0)),
        ("System_SemiSmooth_Laplace", _par(500, 4, kappa=1.0)),
        ("System_Smooth_Laplace", _par(600, 5)),
        ("System_Cuspy_Quartic", _par(900, 6, a1=1.0, a2=5.0)),
        ("System_SemiSmooth_Quartic", _par(300, 7, a1=1.0, a2=5.0, kappa=1.0)),
        ("System_Smooth_Quartic", _par(400, 8, a1=1.0, a2=5.0)),
        ("Line2d.System_Cuspy_Laplace", _par(32 * 48, 9, shape=[32, 48])),
        ("System_Cuspy_Laplace_Nopassing", dict(_par(800, 10), eta=0.0, dt=0.0)),
    ]
    if large:
        s.append(("System_Cuspy_Laplace", _par(17000, 11)))  # beyond a resident workgroup: streaming
    return s


def _oracle_system(oracle, name, kw):
    rank, pot, inter, minim = CLASS[name]
    return oracle._System(rank=rank, potential=pot, interactions=inter, minimisation=minim, **kw)


def _assert_same(rg, singles, what):
    u, idx = rg.u, rg.index_at_align
    for i, s in enumerate(singles):
        assert np.array_equal(u[i].ravel(), np.ravel(s.u)), (what, i, np.max(np.abs(u[i].ravel() - np.ravel(s.u))))
        assert np.array_equal(idx[i].ravel(), np.ravel(s.chunk.index_at_align)), (what, i)
        assert rg.inc[i] == s.inc, (what, i)
        assert rg.u_frame[i] == s.u_frame, (what, i)


def test_ragged_quasistatic_protocol(oracle):
    """The examples' protocol (examples/Line1d_Cuspy_Laplace.py:39-61) on a
    ragged set: elastic step (not for Smooth, whose maxUniformDisplacement
    raises), kick, minimise; three events."""
    from frictionqpotspringblock_amd.ragged import Ragged

    systems = _systems()
    rg = Ragged(systems)
    assert len(rg) == len(systems)
    assert rg.groups >= 9
    singles = [_oracle_system(oracle, n, kw) for n, kw in systems]
    eps = 1e-3
    smooth = np.array(["Smooth" in n for n, _ in systems])
    ret = rg.minimise()
    for i, s in enumerate(singles):
        assert ret[i] == s.minimise()
    _assert_same(rg, singles, "initial minimise")
    for ev in range(3):
        du = rg.eventDrivenStep(eps, False, active=~smooth)
        for i, s in enumerate(singles):
            if not smooth[i]:
                assert du[i] == s.eventDrivenStep(eps, False), (ev, i)
            else:
                assert du[i] == 0.0
        _assert_same(rg, singles, f"elastic {ev}")
        du = rg.eventDrivenStep(eps, True)
        for i, s in enumerate(singles):
            assert du[i] == s.eventDrivenStep(eps, True)
        ret = rg.minimise()
        for i, s in enumerate(singles):
            assert ret[i] == s.minimise(), (ev, i)
        _assert_same(rg, singles, f"event {ev}")
    assert np.all(rg.inc[~np.array(["Nopassing" in n for n, _ in systems])] > 0)


def test_ragged_dynamics_and_fields(oracle):
    from frictionqpotspringblock_amd.ragged import Ragged

    systems = [x for x in _systems(large=False) if "Nopassing" not in x[0]]
    rg = Ragged(systems)
    singles = [_oracle_system(oracle, n, kw) for n, kw in systems]
    rng = np.random.default_rng(0)
    v0 = [rng.normal(size=kw["shape"]) * (0.5 if "Quartic" in n else 2.0) for n, kw in systems]
    rg.v = v0
    for s, v in zip(singles, v0):
        s.v = v
    rg.flowSteps(300, 1e-2)
    for s in singles:
        s.flowSteps(300, 1e-2)
    rg.timeSteps(50)
    for s in singles:
        s.timeSteps(50)
    _assert_same(rg, singles, "dynamics")
    for name in ["v", "a", "f", "f_potential", "f_interactions", "f_frame", "f_damping"]:
        got = getattr(rg, name)
        for i, s in enumerate(singles):
            assert np.array_equal(got[i].ravel(), np.ravel(getattr(s, name))), (name, i)


def test_ragged_masks_and_errors(oracle):
    from frictionqpotspringblock_amd.ragged import Ragged

    systems = [x for x in _systems(large=False) if x[0] in ("System_Cuspy_Laplace", "System_Smooth_Laplace")]
    rg = Ragged(systems)
    smooth = np.array(["Smooth" in n for n, _ in systems])
    with pytest.raises(RuntimeError, match="Operation not possible."):
        rg.eventDrivenStep(1e-3, False)
    # masked-out systems stay untouched
    u0 = rg.u
    active = np.zeros(len(systems), dtype=bool)
    active[0] = True
    du = rg.eventDrivenStep(1e-3, True, active=active)
    assert du[0] != 0 and np.all(du[1:] == 0)
    u1 = rg.u
    assert not np.array_equal(u1[0], u0[0])
    for i in range(1, len(systems)):
        assert np.array_equal(u1[i], u0[i])
    ret = rg.minimise(active=active)
    assert np.all(ret == 0)
    assert rg.inc[0] > 0 and np.all(rg.inc[1:] == 0)
    with pytest.raises(RuntimeError, match="No convergence found"):
        rg.eventDrivenStep(1e-1, True, active=~smooth)
        rg.minimise(max_iter=3, active=~smooth)
    with pytest.raises(RuntimeError, match="unknown system class"):
        Ragged([("System_Nothing", _par(10, 0))])
    # no-passing chains on both sides of the resident kernel's limit (8192
    # particles, 16 per thread on 512 threads; longer ones sweep over tiles)
    for n in (8192, 9000):
        kw = dict(_par(n, 0), eta=0.0, dt=0.0)
        rgn = Ragged([("System_Cuspy_Laplace_Nopassing", kw)])
        o = _oracle_system(oracle, "System_Cuspy_Laplace_Nopassing", kw)
        for s in (rgn, o):
            s.eventDrivenStep(1e-2, True)
        assert rgn.minimise()[0] == o.minimise() == 0
        _assert_same(rgn, [o], f"no-passing {n}")
        assert rgn.last_steps[0] > 10


def test_ragged_autotune(oracle, tmp_path):
    """The tuner picks a launch shape per resident group and a block length
    per tiled group (chains and a lattice on 2-d tiles, whose blocks are at
    most its halo width) by timing trial events on a copy of the set; the
    choice changes no result (a tuned set and an untuned one whose groups run
    one after the other stay bit-identical), and a cache file applies the same
    choices without trials."""
    from frictionqpotspringblock_amd.ragged import Ragged

    systems = _systems() + [("System_Cuspy_Laplace", _par(5000, 12)),
                            ("Line2d.System_Cuspy_Laplace", _par(48 * 64, 13, shape=[48, 64]))]
    systems = [x for x in systems if "Nopassing" not in x[0]]
    smooth = np.array(["Smooth" in n for n, _ in systems])
    a, b = Ragged(systems), Ragged(systems)
    cache = str(tmp_path / "tune.json")
    res = a.autotune(events=1, blocks=(8, 32), cache=cache, smooth=smooth)
    assert res and all(not r["cached"] for r in res.values())
    assert any(r["path"] == "tiles" for r in res.values())
    lattice = 0
    for g, r in res.items():
        info = a.group_info(g)
        if r["choice"][0] == "launch":
            assert list(info["launch"]) == list(r["choice"][1])
        else:
            assert info["block"] == r["choice"][1]
            if info["key"].startswith("2-"):
                lattice += 1
                assert info["block_max"] == 8 and r["choice"][1] in (2, 4, 8), info
                assert sorted(r["trials_ms"]) == ["block:2", "block:4", "block:8"]
    assert lattice == 1
    b.serial = True
    for rg in (a, b):
        rg.quasistatic(1e-3, 5, active=~smooth)
        rg.quasistatic(1e-3, 3, kicks_only=True, active=smooth)
    for x, y in zip(a.u, b.u):
        assert np.array_equal(x, y)
    assert np.array_equal(a.inc, b.inc)
    c = Ragged(systems)
    res2 = c.autotune(cache=cache, smooth=smooth)
    assert all(r["cached"] for r in res2.values())
    assert [r["choice"] for r in res2.values()] == [r["choice"] for r in res.values()]
