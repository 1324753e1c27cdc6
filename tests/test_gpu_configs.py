"""Each benchmark configuration checked against the oracle at the size it runs
(the other parity tests use reduced sizes so that the oracle finishes fast).

* batched flow (1024 Cuspy-Laplace replicas x 4096, seed r, nchunk 200):
  1000 flowSteps of the whole batch in one resident launch; replicas 0, 511
  and 1023 (the last one's landscape starts 6.7e9 bytes into the batch's
  array, beyond 2^31) equal oracle systems of seed r bit for bit;
* ragged set: members of the ragged workload at its largest sizes -- a
  65536-particle Cuspy-Quartic chain (a1 = 1, a2 = 5), the workload's first
  256 x 256 Line2d lattice and its largest Cuspy-Laplace, SemiSmooth-Laplace
  and Smooth-Quartic chains -- through the examples' protocol, every event's
  avalanche size, frame position, mean frame force and increment equal to the
  oracle's, and the final states bit for bit;
* adjoint at the adjoint workload's size per replica: N = 4096 (the 8 x 512
  launch shape exactly full), T = 2000 steps, checkpoints every 100 steps,
  dot-product test against the FP64 tangent-linear model.
"""

import numpy as np
import pytest

import bench
from tests import adjoint_tlm as TLM
from tests import protocols as P

pytestmark = pytest.mark.gpu


def _flow_oracle(oracle, r):
    s = oracle.Line1d.System_Cuspy_Laplace(**dict(bench.PARAMS, seed=r))
    s.flowSteps(bench.BLOCK, bench.V_FRAME)
    return s


def test_batched_flow_full_size(oracle):
    from frictionqpotspringblock_amd.batch import Batch

    print(f"batched flow: {bench.R} x {bench.N}, {bench.BLOCK} flowSteps", flush=True)
    b = Batch(bench.R, seed_stride=1, potential="Cuspy", interactions="Laplace1d", **bench.PARAMS)
    b.flowSteps(bench.BLOCK, bench.V_FRAME, block=bench.BLOCK)
    u, v, a = b.u, b.v, b.a
    idx, inc, uf = b.index_at_align, b.inc, b.u_frame
    f = {k: getattr(b, k) for k in ("f", "f_potential", "f_interactions", "f_frame", "f_damping")}
    assert np.all(inc == bench.BLOCK)
    for r in (0, 511, bench.R - 1):
        s = _flow_oracle(oracle, r)
        assert np.array_equal(u[r], s.u), r
        assert np.array_equal(v[r], s.v), r
        assert np.array_equal(a[r], s.a), r
        assert np.array_equal(idx[r], s.chunk.index_at_align), r
        assert inc[r] == s.inc and uf[r] == s.u_frame, r
        for k, x in f.items():
            assert np.array_equal(x[r], getattr(s, k)), (r, k)
    # the particles moved and crossed wells during the run
    assert np.any(idx[bench.R - 1] != oracle.Line1d.System_Cuspy_Laplace(**dict(bench.PARAMS, seed=1023)).chunk.index_at_align)


CLASS = {
    "System_Cuspy_Laplace": (1, "Cuspy", "Laplace1d"),
    "System_SemiSmooth_Laplace": (1, "SemiSmooth", "Laplace1d"),
    "System_Smooth_Quartic": (1, "Smooth", "Quartic1d"),
    "System_Cuspy_Quartic": (1, "Cuspy", "Quartic1d"),
    "Line2d.System_Cuspy_Laplace": (2, "Cuspy", "Laplace2d"),
}


def _ragged_members():
    """The largest members of the ragged workload (bench.ragged_systems) of
    each class tested here, plus a Cuspy-Quartic chain of the workload's
    largest size."""
    systems = bench.ragged_systems()
    out = []
    for cls in ("System_Cuspy_Laplace", "System_SemiSmooth_Laplace", "System_Smooth_Quartic"):
        best = max((s for s in systems if s[0] == cls), key=lambda s: s[1]["shape"][0])
        out.append(best)
    out.append(next(s for s in systems if s[0] == "Line2d.System_Cuspy_Laplace"))
    kw = dict(bench.PARAMS, k_frame=1.0 / 65536, shape=[65536], seed=10**7, nchunk=200, a1=1.0, a2=5.0)
    kw.pop("k_interactions")
    out.append(("System_Cuspy_Quartic", kw))
    return out


def test_ragged_largest_members(oracle):
    from frictionqpotspringblock_amd.ragged import Ragged

    systems = _ragged_members()
    sizes = [int(np.prod(kw["shape"])) for _, kw in systems]
    print("ragged members:", [(n, kw["shape"]) for n, kw in systems], flush=True)
    assert max(sizes) == 65536 and sum(s > 60000 for s in sizes) == len(sizes)
    rg = Ragged(systems)
    smooth = np.array(["Smooth" in n for n, _ in systems])
    nev = 5
    q = rg.quasistatic(bench.RG_EPS, nev, active=~smooth)
    qk = rg.quasistatic(bench.RG_EPS, 3, kicks_only=True, active=smooth)
    u, idx = rg.u, rg.index_at_align
    for i, (name, kw) in enumerate(systems):
        rank, pot, inter = CLASS[name]
        s = oracle._System(rank=rank, potential=pot, interactions=inter, **kw)
        if smooth[i]:
            s.u_frame = 0.0
            s.minimise()
            for e in range(1, 3):
                i0 = s.chunk.index_at_align.copy()
                s.eventDrivenStep(bench.RG_EPS, True)
                s.minimise()
                assert qk["S"][i, e] == np.sum(s.chunk.index_at_align - i0), (name, e)
                assert qk["inc"][i, e] == s.inc, (name, e)
        else:
            uf, ff, S, inc = P.quasistatic(s, nev)
            assert np.array_equal(q["S"][i], S), (name, q["S"][i], S)
            assert np.array_equal(q["inc"][i], inc), name
            assert np.array_equal(q["u_frame"][i], uf), name
            assert np.allclose(q["f_frame"][i], ff, rtol=1e-12, atol=1e-15), name
        assert np.array_equal(np.ravel(u[i]), np.ravel(s.u)), name
        assert np.array_equal(np.ravel(idx[i]), np.ravel(s.chunk.index_at_align)), name
        assert rg.inc[i] == s.inc, name


def test_ragged_nopassing_largest(oracle):
    """The no-passing pass of the ragged workload on its three largest
    Cuspy-Laplace chains (about 60000 particles: sweeps over tiles), a few
    events of the examples' protocol, bit for bit."""
    from frictionqpotspringblock_amd.ragged import Ragged

    nps = sorted(bench.ragged_nopassing(bench.ragged_systems()), key=lambda s: -s[1]["shape"][0])[:3]
    print("no-passing members:", [kw["shape"] for _, kw in nps], flush=True)
    assert all(kw["shape"][0] > 8192 for _, kw in nps)
    rg = Ragged(nps)
    singles = [oracle._System(rank=1, potential="Cuspy", interactions="Laplace1d", minimisation=1, **kw)
               for _, kw in nps]
    q = rg.quasistatic(bench.RG_EPS, 5)
    for i, s in enumerate(singles):
        uf, ff, S, inc = P.quasistatic(s, 5)
        assert np.array_equal(q["S"][i], S), (i, q["S"][i], S)
        assert np.array_equal(q["u_frame"][i], uf), i
    u, idx = rg.u, rg.index_at_align
    for i, s in enumerate(singles):
        assert np.array_equal(u[i], s.u), i
        assert np.array_equal(idx[i], s.chunk.index_at_align), i


def test_adjoint_full_size_dot_product():
    """N = 4096 (8 particles per thread on 512 threads: every slot of the
    launch holds a particle), T = 2000 in windows of 100 steps."""
    import torch

    from frictionqpotspringblock_amd import adjoint as A
    from frictionqpotspringblock_amd.batch import Batch

    R, N, T, K = 2, bench.AD_N, bench.AD_T, bench.AD_K
    print(f"adjoint: R={R} N={N} T={T} K={K}", flush=True)
    kw = dict(bench.PARAMS, k_frame=1.0 / N, shape=[N], nchunk=400)
    b = Batch(R, seed_stride=N, potential="Cuspy", interactions="Laplace1d", **kw)
    rng = np.random.default_rng(11)
    dev = "cuda"
    u0 = A.get_field(b, "u")
    v0 = torch.tensor(rng.normal(size=(R, N)) * 0.5, device=dev)
    a0 = torch.zeros((R, N), dtype=torch.float64, device=dev)
    par = A.parameters(b)
    y = A.chunk_data(b)
    leaves = [x.clone().requires_grad_(True) for x in (u0, v0, a0, par, y)]
    uT, vT, aT = A.run(b, T, K, *leaves)
    args = [x.detach().cpu().numpy() for x in (u0, v0, a0)]
    i0 = TLM.wells(y.cpu().numpy(), args[0])
    iT = TLM.wells(y.cpu().numpy(), uT.detach().cpu().numpy())
    assert np.sum(i0 != iT) > 100
    d = [rng.normal(size=x.shape) for x in args] + [rng.normal(size=(R, 4)) * 1e-2, rng.normal(size=tuple(y.shape))]
    uf0 = np.zeros(R)
    (u, v, a), (du, dv, da) = TLM.run(*args, y.cpu().numpy(), par.numpy(), uf0, T, bench.PARAMS["dt"], 1.0,
                                      tangent=d)
    assert np.array_equal(u, uT.detach().cpu().numpy())
    assert np.array_equal(v, vT.detach().cpu().numpy())
    c = [rng.normal(size=(R, N)) for _ in range(3)]
    lhs = np.sum(c[0] * du) + np.sum(c[1] * dv) + np.sum(c[2] * da)
    loss = sum((torch.tensor(ci, device=dev) * o).sum() for ci, o in zip(c, (uT, vT, aT)))
    loss.backward()
    grads = [x.grad.detach().cpu().numpy() for x in leaves]
    rhs = sum(np.sum(g * di) for g, di in zip(grads, d))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs)), (lhs, rhs)
