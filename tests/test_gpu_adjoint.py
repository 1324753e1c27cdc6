"""Adjoint kernels (through torch.autograd) against finite differences and an
FP64 CPU tangent-linear model.

* dot-product test: <c, J d> from the CPU tangent-linear model (tests/
  adjoint_tlm.py) equals <J^T c, d> from the adjoint, for random c and d over
  every input (u0, v0, a0, parameters, yield positions), across many well
  crossings; relative tolerance 1e-10 (both sides are FP64 evaluations of the
  same bilinear form; their rounding differs only in summation order).
* central finite differences on a crossing-free window (no well changes under
  the perturbation): relative error <= 1e-6 per direction.
"""

import numpy as np
import pytest

from tests import adjoint_tlm as TLM

pytestmark = pytest.mark.gpu

DT = 0.1
M = 1.0


def _batch(R, N, nchunk=300, seed=0, **kw):
    from frictionqpotspringblock_amd.batch import Batch

    par = dict(
        potential="Cuspy",
        interactions="Laplace1d",
        m=M,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        k_frame=1.0 / N,
        dt=DT,
        shape=[N],
        seed=seed,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=nchunk,
    )
    par.update(kw)
    return Batch(R, seed_stride=N, **par)


def _setup(b, rng, vscale):
    import torch

    from frictionqpotspringblock_amd import adjoint as A

    R, N = b.replicas, b.size
    u0 = A.get_field(b, "u").cpu().numpy()
    v0 = rng.normal(size=(R, N)) * vscale
    a0 = rng.normal(size=(R, N)) * vscale * 0.1
    par = A.parameters(b).numpy()
    par = par * (1.0 + 0.05 * rng.uniform(-1, 1, size=par.shape))  # replicas differ
    y = A.chunk_data(b)
    t = lambda x: torch.tensor(x, dtype=torch.float64, device="cuda")  # noqa: E731
    return t(u0), t(v0), t(a0), torch.tensor(par, dtype=torch.float64), y


@pytest.mark.parametrize("flow", [False, True])
@pytest.mark.parametrize("N", [200, 1500, 2100, 3500, 6000, 20000])
def test_dot_product_with_crossings(flow, N):
    """N = 200, 1500, 2100, 3500: the whole chain on one workgroup, 1, 2, 4
    and 4 particles per thread, each with a partly filled last wavefront;
    N = 6000 and 20000: the record kernel over 2 and 6 tiles with halos of
    K particles, the reverse kernel over tiles that a single workgroup could
    not hold (and at N = 20000 a forward pass on the blocked tiled path)."""
    import torch

    from frictionqpotspringblock_amd import adjoint as A

    rng = np.random.default_rng(1)
    R, n, K = 3, 260, 64  # last window partial
    b = _batch(R, N)
    u0, v0, a0, par, y = _setup(b, rng, vscale=1.0)
    uf0 = torch.tensor(b.u_frame)
    vf = 0.05 if flow else 0.0
    leaves = [x.clone().requires_grad_(True) for x in (u0, v0, a0, par, y)]
    uT, vT, aT = A.run(b, n, K, *leaves, flow=flow, v_frame=vf)
    i0 = TLM.wells(y.cpu().numpy(), u0.cpu().numpy())
    iT = TLM.wells(y.cpu().numpy(), uT.detach().cpu().numpy())
    assert np.sum(i0 != iT) > 50  # the trajectory crosses many wells
    # CPU forward: bit-identical trajectory
    args = [x.detach().cpu().numpy() for x in (u0, v0, a0)]
    d = [rng.normal(size=x.shape) for x in args] + [
        rng.normal(size=(R, 4)) * 1e-2,
        rng.normal(size=tuple(y.shape)),
    ]
    (u, v, a), (du, dv, da) = TLM.run(
        *args, y.cpu().numpy(), par.numpy(), uf0.numpy(), n, DT, M, flow=flow, v_frame=vf, tangent=d
    )
    assert np.array_equal(u, uT.detach().cpu().numpy())
    assert np.array_equal(v, vT.detach().cpu().numpy())
    assert np.array_equal(a, aT.detach().cpu().numpy())
    c = [rng.normal(size=(R, N)) for _ in range(3)]
    lhs = np.sum(c[0] * du) + np.sum(c[1] * dv) + np.sum(c[2] * da)
    loss = sum((torch.tensor(ci, device="cuda") * o).sum() for ci, o in zip(c, (uT, vT, aT)))
    loss.backward()
    grads = [x.grad.detach().cpu().numpy() for x in leaves]
    rhs = sum(np.sum(g * di) for g, di in zip(grads, d))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs)), (lhs, rhs)
    # each input on its own (catches errors that cancel in the sum)
    for q in range(5):
        dq = [np.zeros_like(x) for x in d]
        dq[q] = d[q]
        _, (du, dv, da) = TLM.run(
            *args, y.cpu().numpy(), par.numpy(), uf0.numpy(), n, DT, M, flow=flow, v_frame=vf, tangent=dq
        )
        l_q = np.sum(c[0] * du) + np.sum(c[1] * dv) + np.sum(c[2] * da)
        r_q = np.sum(grads[q] * d[q])
        assert abs(l_q - r_q) <= 1e-10 * (abs(l_q) + abs(r_q)) + 1e-300, (q, l_q, r_q)


def test_finite_differences_crossing_free():
    import torch

    from frictionqpotspringblock_amd import adjoint as A

    rng = np.random.default_rng(2)
    R, N, n, K = 2, 128, 120, 50
    b = _batch(R, N, seed=5)
    b.minimise()
    u0, v0, a0, par, y = _setup(b, rng, vscale=1e-3)
    ut = torch.tensor(rng.normal(size=(R, N)), device="cuda")

    def loss(u0_, v0_, a0_, par_, y_):
        uT, vT, aT = A.run(b, n, K, u0_, v0_, a0_, par_, y_)
        return ((uT - ut) ** 2).sum() + 0.3 * (vT**2).sum() + 0.1 * (aT * ut).sum(), uT

    leaves = [x.clone().requires_grad_(True) for x in (u0, v0, a0, par, y)]
    L, uT = loss(*leaves)
    assert np.array_equal(
        TLM.wells(y.cpu().numpy(), u0.cpu().numpy()), TLM.wells(y.cpu().numpy(), uT.detach().cpu().numpy())
    )
    L.backward()
    base = [x.detach() for x in (u0, v0, a0, par, y)]
    for q, scale in enumerate([1e-6, 1e-6, 1e-6, 1e-6, 1e-6]):
        g = leaves[q].grad
        for trial in range(2):
            dq = torch.tensor(rng.normal(size=tuple(base[q].shape)), dtype=torch.float64, device=base[q].device)
            if q == 4:  # only the yield positions of the current wells matter: keep the landscape ordered
                dq = dq * 0.1
            h = scale
            plus = [x.clone() for x in base]
            minus = [x.clone() for x in base]
            plus[q] = plus[q] + h * dq
            minus[q] = minus[q] - h * dq
            with torch.no_grad():
                Lp, _ = loss(*plus)
                Lm, _ = loss(*minus)
            fd = (Lp - Lm).item() / (2 * h)
            ad = float((g.to(dq.device) * dq).sum())
            assert abs(fd - ad) <= 1e-6 * max(abs(ad), 1e-8), (q, trial, fd, ad)


def test_adjoint_errors():
    from frictionqpotspringblock_amd import adjoint as A
    from frictionqpotspringblock_amd.batch import Batch, _check, capi

    b = _batch(2, 64, interactions="Quartic1d", a1=1.0, a2=1.0)
    with pytest.raises(RuntimeError, match="adjoint needs"):
        _check(capi().fqsb_batch_adjoint_forward(b._h, 10, 5, 0, 0.0))
    b = _batch(2, 64)
    with pytest.raises(RuntimeError, match="no adjoint forward run"):
        _check(capi().fqsb_batch_adjoint_backward_device(b._h, *([None] * 8)))
    # chains over tiles: the halos of a window must leave room for the tile
    b = _batch(1, 5000)
    with pytest.raises(RuntimeError, match="tiles do not fit"):
        _check(capi().fqsb_batch_adjoint_forward(b._h, 4000, 2048, 0, 0.0))
    # a trajectory that needs a landscape refill is refused
    b = _batch(1, 64, nchunk=40, k_frame=0.5)
    with pytest.raises(RuntimeError, match="refill"):
        _check(capi().fqsb_batch_adjoint_forward(b._h, 2000, 100, 1, 0.5))
    assert isinstance(Batch, type) and A.parameters(b).shape == (1, 4)


def test_backward_of_a_replaced_tape_raises():
    """A batch holds the tape of its last forward pass: the backward pass of an
    earlier run on the same batch, or of a run after which the batch was
    stepped (a refill could have changed the landscape), raises; the last
    run still differentiates and matches the tangent-linear model."""
    import torch

    from frictionqpotspringblock_amd import adjoint as A

    rng = np.random.default_rng(3)
    R, N, n, K = 2, 256, 60, 20
    b = _batch(R, N)
    u0, v0, a0, par, y = _setup(b, rng, vscale=0.5)
    l1 = [x.clone().requires_grad_(True) for x in (u0, v0, a0)]
    u1, v1, a1 = A.run(b, n, K, *l1, par)
    g1 = A.tape_generation(b)
    assert g1 > 0
    # second run on the same batch, chained on the first one's output
    u2, _, _ = A.run(b, n, K, u1, v1, a1, par)
    assert A.tape_generation(b) == g1 + 1
    with pytest.raises(RuntimeError, match="tape of this forward pass is gone"):
        u1.sum().backward(retain_graph=True)
    # the last run differentiates; through it, the first one's backward raises
    l2 = [x.detach().clone().requires_grad_(True) for x in (u1, v1, a1)]
    u3, _, _ = A.run(b, n, K, *l2, par)
    u3.sum().backward()
    assert all(x.grad is not None and torch.all(torch.isfinite(x.grad)) for x in l2)
    # stepping the batch after a forward pass invalidates its tape
    l3 = [x.detach().clone().requires_grad_(True) for x in (u0, v0, a0)]
    u4, _, _ = A.run(b, n, K, *l3, par)
    b.timeSteps(5)
    assert A.tape_generation(b) == 0
    with pytest.raises(RuntimeError, match="tape of this forward pass is gone"):
        u4.sum().backward()
    del u2
