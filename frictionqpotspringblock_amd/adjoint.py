"""Differentiable batched dynamics: torch.autograd over the adjoint kernels.

No reference counterpart. ``run(batch, n, window, u0, v0, a0, params, y)``
sets the batch's state, per-replica parameters and (optionally) its
yield-position chunks from torch tensors on the GPU, runs ``n`` timeSteps (or
flowSteps, detail.h:1539-1583, 1637-1645) keeping a checkpoint every
``window`` steps, and returns the final (u, v, a). Its backward pass runs the
hand-written adjoint kernels (libfqsb.so, ``fqsb_batch_adjoint_*``) and gives
the gradients with respect to u0, v0, a0, params [R, 4] = (mu,
k_interactions, k_frame, eta) and y [R, N, nchunk]. Well indices are held at
the values the trajectory used and are not differentiated.

Cuspy + Laplace1d replicas (``Batch(..., potential="Cuspy",
interactions="Laplace1d")``) of any length (chains above 4096 particles run
over tiles with halos of about ``window`` particles, so a window must be
shorter than 2048 steps there), no external forcing. A batch holds
the tape of its last forward pass only: the backward pass of an earlier
``run`` on the same batch (or one after any other call that changed the
batch) raises instead of using another trajectory; chain runs on separate
batches.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from .batch import FIELD, _check, _dp, capi


def _ptr(t):
    return C.c_void_p(t.data_ptr())


def _f64(t, device):
    return t.detach().to(device=device, dtype=torch.float64).contiguous()


def parameters(batch):
    """Per-replica (mu, k_interactions, k_frame, eta) as a [R, 4] float64 tensor."""
    out = np.empty((batch.replicas, 4))
    _check(capi().fqsb_batch_get_parameters(batch._h, out.ctypes.data_as(_dp)))
    return torch.from_numpy(out)


def set_parameters(batch, params):
    p = np.ascontiguousarray(torch.as_tensor(params, dtype=torch.float64).detach().cpu().numpy())
    if p.shape != (batch.replicas, 4):
        raise RuntimeError("parameters must have shape [replicas, 4]")
    _check(capi().fqsb_batch_set_parameters(batch._h, p.ctypes.data_as(_dp)))


def chunk_data(batch, device="cuda"):
    """The yield-position chunks of all replicas, [R, N, nchunk] on the GPU."""
    nch = int(capi().fqsb_batch_nchunk(batch._h))
    out = torch.empty((batch.replicas, batch.size, nch), dtype=torch.float64, device=device)
    torch.cuda.current_stream(out.device).synchronize()
    _check(capi().fqsb_batch_chunk_data_device(batch._h, _ptr(out)))
    return out


def tape_generation(batch):
    """Which tape the batch holds: increases with every forward pass of
    `run`; 0 once the batch's state, landscape or parameters changed after
    the last one (fqsb_batch_adjoint_generation)."""
    g = C.c_uint64()
    _check(capi().fqsb_batch_adjoint_generation(batch._h, C.byref(g)))
    return g.value


def get_field(batch, name, device="cuda"):
    out = torch.empty((batch.replicas, batch.size), dtype=torch.float64, device=device)
    torch.cuda.current_stream(out.device).synchronize()
    _check(capi().fqsb_batch_get_device(batch._h, FIELD[name], _ptr(out)))
    return out


class _VerletAdjoint(torch.autograd.Function):
    @staticmethod
    def forward(ctx, batch, n, window, flow, v_frame, u0, v0, a0, params, y):
        lib = capi()
        h = batch._h
        dev = u0.device
        R, N = batch.replicas, batch.size
        u0c, v0c, a0c = (_f64(x, dev).reshape(R, N) for x in (u0, v0, a0))
        torch.cuda.current_stream(dev).synchronize()
        if y is not None:
            yc = _f64(y, dev)
            torch.cuda.current_stream(dev).synchronize()
            _check(lib.fqsb_batch_set_chunk_data_device(h, _ptr(yc)))
        set_parameters(batch, params)
        _check(lib.fqsb_batch_set_device(h, FIELD["u"], _ptr(u0c)))
        _check(lib.fqsb_batch_set_device(h, FIELD["v"], _ptr(v0c)))
        _check(lib.fqsb_batch_set_device(h, FIELD["a"], _ptr(a0c)))
        _check(lib.fqsb_batch_adjoint_forward(h, int(n), int(window), int(bool(flow)), float(v_frame)))
        outs = []
        for name in ("u", "v", "a"):
            t = torch.empty((R, N), dtype=torch.float64, device=dev)
            _check(lib.fqsb_batch_get_device(h, FIELD[name], _ptr(t)))
            outs.append(t)
        ctx.batch = batch
        ctx.generation = tape_generation(batch)
        ctx.has_y = y is not None
        ctx.y_shape = None if y is None else tuple(y.shape)
        ctx.par_device = torch.as_tensor(params).device
        ctx.shape = tuple(u0.shape)
        ctx.dev = dev
        return tuple(outs)

    @staticmethod
    def backward(ctx, gu, gv, ga):
        batch, dev = ctx.batch, ctx.dev
        R, N = batch.replicas, batch.size
        # the batch holds one tape (the last forward pass), recomputed from its
        # current landscape and parameters: refuse a backward pass whose
        # forward pass is no longer the one the batch holds
        if tape_generation(batch) != ctx.generation:
            raise RuntimeError(
                "the adjoint tape of this forward pass is gone: the batch ran another forward pass or changed "
                "its state, landscape or parameters since; a batch differentiates its last run() only"
            )
        g = []
        for x in (gu, gv, ga):
            g.append(torch.zeros((R, N), dtype=torch.float64, device=dev) if x is None else _f64(x, dev))
        gu0, gv0, ga0 = (torch.empty((R, N), dtype=torch.float64, device=dev) for _ in range(3))
        gpar = torch.empty((R, 4), dtype=torch.float64, device=dev)
        want_y = ctx.has_y and ctx.needs_input_grad[9]
        gy = torch.empty(ctx.y_shape, dtype=torch.float64, device=dev) if want_y else None
        torch.cuda.current_stream(dev).synchronize()
        _check(
            capi().fqsb_batch_adjoint_backward_device(
                batch._h, _ptr(g[0]), _ptr(g[1]), _ptr(g[2]), _ptr(gu0), _ptr(gv0), _ptr(ga0), _ptr(gpar),
                None if gy is None else _ptr(gy),
            )
        )
        return (
            None, None, None, None, None,
            gu0.reshape(ctx.shape), gv0.reshape(ctx.shape), ga0.reshape(ctx.shape),
            gpar.to(ctx.par_device), gy,
        )


def kernel_ms(batch):
    """(record, reverse): average duration in ms of one window's recompute and
    reverse kernel in the last backward pass."""
    a, b = C.c_float(), C.c_float()
    _check(capi().fqsb_batch_adjoint_kernel_ms(batch._h, C.byref(a), C.byref(b)))
    return a.value, b.value


def run(batch, n, window, u0, v0, a0, params, y=None, flow=False, v_frame=0.0):
    """n steps of the batch from (u0, v0, a0) with parameters params [R, 4]
    (and yield-position chunks y [R, N, nchunk] if given); differentiable.
    Returns (u, v, a) after the steps."""
    return _VerletAdjoint.apply(batch, n, window, flow, v_frame, u0, v0, a0, params, y)
