"""FP64 CPU forward and tangent-linear model of batched Cuspy + Laplace1d
timeSteps / flowSteps (detail.h:1539-1570, 1637-1645), used to check the
adjoint kernels.

The forward evaluates every expression in the order of the reference (numpy
does not fuse multiply-adds), so its trajectory and the wells it visits are
bit-identical to the device's. The tangent-linear model propagates a
perturbation of (u, v, a), of the parameters (mu, k_interactions, k_frame, eta)
and of the yield positions through the same steps, with every well index held
at the value of the trajectory. A dot-product test  <c, J d> = <J^T c, d>
then pins the adjoint exactly, also across well crossings.
"""

import numpy as np


def wells(y, x):
    """Chunk-local index j with y[j] <= x < y[j + 1] for every (r, p)."""
    R, N, n = y.shape
    j = np.empty((R, N), dtype=np.int64)
    for r in range(R):
        for p in range(N):
            j[r, p] = np.searchsorted(y[r, p], x[r, p], side="right") - 1
    assert np.all(j >= 0) and np.all(j < n - 1)
    return j


def walk(y, x, j):
    """The same wells as `wells`, found by walking from the indices j of the
    previous step (a step moves a particle by a few wells at most)."""
    R, N, n = y.shape
    rows = np.arange(R)[:, None]
    cols = np.arange(N)[None, :]
    j = j.copy()
    while True:
        up = (j < n - 2) & (x >= y[rows, cols, np.minimum(j + 1, n - 1)])
        down = (j > 0) & (x < y[rows, cols, j])
        if not (np.any(up) or np.any(down)):
            break
        j += up
        j -= down
    assert np.all(y[rows, cols, j] <= x) and np.all(x < y[rows, cols, j + 1])
    return j


def lap(x):
    """(x[p-1] - 2 x[p] + x[p+1]), periodic, in the reference's order."""
    return np.roll(x, 1, axis=1) - 2 * x + np.roll(x, -1, axis=1)


def run(u, v, a, y, par, uf, n, dt, m, flow=False, v_frame=0.0, tangent=None):
    """n steps from (u, v, a) [R, N] with chunks y [R, N, nchunk], parameters
    par [R, 4] and frame positions uf [R]. With `tangent` = (du, dv, da, dpar,
    dy), also the tangent-linear model. Returns (u, v, a) and, if asked, the
    final tangent (du, dv, da)."""
    u, v, a = u.copy(), v.copy(), a.copy()
    mu, k, kf, eta = (par[:, q : q + 1] for q in range(4))
    meta = -eta
    inv_m = 1.0 / m
    c2 = 0.5 * dt * dt
    hdt = 0.5 * dt
    uf = uf.reshape(-1, 1).copy()
    R, N = u.shape
    rows = np.arange(R)[:, None]
    cols = np.arange(N)[None, :]
    if tangent is not None:
        du, dv, da, dpar, dy = (np.array(x, dtype=float) for x in tangent)
        dmu, dk, dkf, deta = (dpar[:, q : q + 1] for q in range(4))
    j = None
    for _ in range(n):
        if flow:
            uf = uf + v_frame * dt
        u1 = u + dt * v + c2 * a
        j = wells(y, u1) if j is None else walk(y, u1, j)
        yl = y[rows, cols, j]
        yr = y[rows, cols, j + 1]
        lu = lap(u1)
        fi = lu * k
        fp = (0.5 * (yl + yr) - u1) * mu
        ff = kf * (uf - u1)
        sum3 = ff + fp + fi
        vv1 = v + dt * a
        a1 = (sum3 + meta * vv1) * inv_m
        vv2 = v + hdt * (a + a1)
        a2 = (sum3 + meta * vv2) * inv_m
        vv3 = v + hdt * (a + a2)
        a3 = (sum3 + meta * vv3) * inv_m
        if tangent is not None:
            du1 = du + dt * dv + c2 * da
            dym = 0.5 * (dy[rows, cols, j] + dy[rows, cols, j + 1])
            ym = 0.5 * (yl + yr)
            dg = dkf * (uf - u1) - kf * du1 + dmu * (ym - u1) + mu * (dym - du1) + dk * lu + k * lap(du1)
            dv1 = dv + dt * da
            da1 = (dg - deta * vv1 - eta * dv1) * inv_m
            dv2 = dv + hdt * (da + da1)
            da2 = (dg - deta * vv2 - eta * dv2) * inv_m
            dv3 = dv + hdt * (da + da2)
            da3 = (dg - deta * vv3 - eta * dv3) * inv_m
            du, dv, da = du1, dv3, da3
        u, v, a = u1, vv3, a3
    if tangent is None:
        return u, v, a
    return (u, v, a), (du, dv, da)
