"""Batched LongRange1d in matrix form (FP64 matrix cores) against the exact path.

The exact path (the reference's sequential sum, detail.h:846-867) is pinned
bit for bit against the oracle elsewhere; here the matrix form is compared to
it. One force evaluation of either form is a sum of N products, so each
differs from the true value by at most gamma_N * sum_i |term_i| with
gamma_N = N eps / (1 - N eps) (the standard bound for recursive and blocked
summation, fused multiply-add included); the single-step test uses twice that
bound per particle. Over a trajectory the deviations are damped (eta > 0) and
driven only by per-step rounding; a well index could only differ if a slip fell
within rounding distance of a yield position, so indices are compared exactly
and u within rtol 1e-9 after a few hundred steps.
"""

import numpy as np
import pytest

from tests.test_oracle import circulant_longrange

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _kw(N, **extra):
    kw = dict(
        potential="Cuspy",
        interactions="LongRange1d",
        m=1.0,
        eta=2.0 * np.sqrt(3.0) / 10.0,
        mu=1.0,
        k_interactions=1.0,
        alpha=1.5,
        k_frame=1.0 / N,
        dt=0.1,
        shape=[N],
        seed=0,
        distribution="weibull",
        parameters=[2.0, 1.0, 0.0],
        offset=-100.0,
        nchunk=200,
    )
    kw.update(extra)
    return kw


def _pair(R, N, **extra):
    from frictionqpotspringblock_amd.batch import Batch

    mat = Batch(R, **_kw(N, **extra))
    ext = Batch(R, **_kw(N, **extra))
    assert mat.longrange_method == "matrix"
    ext.longrange_method = "exact"
    assert ext.longrange_method == "exact"
    return mat, ext


def _bound(u, k, alpha):
    """gamma_{N+2} * (sum_i K_pi |u_i - u_p| + sum_i K_pi |u_i - u_0| + c |u_p - u_0|) per particle."""
    R, N = u.shape
    m = np.arange(N)
    d = np.minimum(m, N - m).astype(float)
    g = np.zeros(N)
    g[1:] = k / d[1:] ** (alpha + 1.0)
    idx = np.arange(N)
    K = g[(idx[:, None] - idx[None, :]) % N]
    out = np.empty_like(u)
    for r in range(R):
        x = u[r]
        s_exact = np.sum(K * np.abs(x[None, :] - x[:, None]), axis=1)
        s_mat = K @ np.abs(x - x[0]) + g.sum() * np.abs(x - x[0])
        out[r] = s_exact + s_mat
    gamma = (N + 2) * EPS / (1 - (N + 2) * EPS)
    return 2.0 * gamma * out


@pytest.mark.parametrize("R,N", [(5, 1000), (3, 129), (66, 300)])
def test_single_step_forces(R, N):
    """One step from the same random state: every field within the bound."""
    mat, ext = _pair(R, N)
    rng = np.random.default_rng(N + R)
    u = rng.normal(size=(R, N)) * 2.0
    v = rng.normal(size=(R, N)) * 0.1
    for b in (mat, ext):
        b.u = u
        b.v = v
    mat.timeSteps(1)
    ext.timeSteps(1)
    un = ext.u
    bound = _bound(un, 1.0, 1.5)
    assert np.all(np.abs(mat.f_interactions - ext.f_interactions) <= bound)
    # the FFT eigen-form of the same product (independent of both paths)
    f_fft, _ = circulant_longrange(un, 1.0, 1.5)
    assert np.all(np.abs(mat.f_interactions - f_fft) <= bound * np.log2(N) + 1e-14 * np.max(bound / EPS))
    # u after one step does not depend on the interaction force of that step
    assert np.array_equal(mat.u, ext.u)
    assert np.array_equal(mat.f_potential, ext.f_potential)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)
    # v, a: one multiplication by dt/2 / m of a force within the bound
    np.testing.assert_allclose(mat.v, ext.v, rtol=0, atol=float(np.max(bound)))
    np.testing.assert_allclose(mat.f, ext.f, rtol=0, atol=4 * float(np.max(bound)))
    assert np.array_equal(mat.inc, ext.inc)
    assert np.array_equal(mat.u_frame, ext.u_frame)


@pytest.mark.parametrize("R,N,block", [(4, 1000, 128), (3, 257, 7)])
def test_trajectory(R, N, block):
    mat, ext = _pair(R, N)
    mat.flowSteps(400, 1e-2, block=block)
    ext.flowSteps(400, 1e-2)
    np.testing.assert_allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(mat.v, ext.v, rtol=0, atol=1e-9)
    np.testing.assert_allclose(mat.f_interactions, ext.f_interactions, rtol=0, atol=1e-8)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)
    assert np.array_equal(mat.inc, ext.inc)
    assert np.array_equal(mat.u_frame, ext.u_frame)
    # time steps after flow steps continue from the same state
    mat.timeSteps(50)
    ext.timeSteps(50)
    np.testing.assert_allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)


def test_landscape_refills():
    """A fast drive through a short landscape window: replicas stop for
    refills at different steps and catch up."""
    R, N = 3, 200
    mat, ext = _pair(R, N, nchunk=40, k_frame=0.1)
    i0 = mat.index_at_align
    mat.flowSteps(1000, 0.5, block=50)
    ext.flowSteps(1000, 0.5)
    assert np.all(mat.index_at_align - i0 > 40)  # moved beyond the window: refills
    np.testing.assert_allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)
    assert np.array_equal(mat.inc, ext.inc)


def test_random_forcing():
    """LongRange + RandomNormalForcing (dinc = dinc_init = 10): the thermal
    draws depend on the increment only, so they agree bit for bit."""
    R, N = 4, 500
    extra = dict(external="RandomNormalForcing", mean=0.0, stddev=0.1, seed_forcing=7, dinc_init=10, dinc=10)
    mat, ext = _pair(R, N, **extra)
    mat.timeSteps(205, block=64)
    ext.timeSteps(205)
    assert np.array_equal(mat.f_thermal, ext.f_thermal)
    assert np.any(mat.f_thermal != 0)
    np.testing.assert_allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)


def test_large_shape():
    """The benchmark's system size (N = 8192) with a replica count that is not
    a multiple of the matrix tile; the exact path runs the streaming kernel."""
    R, N = 67, 8192
    mat, ext = _pair(R, N)
    mat.flowSteps(20, 1e-3)
    ext.flowSteps(20, 1e-3)
    np.testing.assert_allclose(mat.u, ext.u, rtol=1e-10, atol=1e-10)
    np.testing.assert_allclose(mat.f_interactions, ext.f_interactions, rtol=0, atol=1e-9)
    f_fft, _ = circulant_longrange(mat.u[:2], 1.0, 1.5)
    np.testing.assert_allclose(mat.f_interactions[:2], f_fft, rtol=0, atol=1e-9)


def test_errors():
    from frictionqpotspringblock_amd.batch import Batch

    b = Batch(2, **_kw(100, interactions="Laplace1d"))
    with pytest.raises(RuntimeError, match="no long-range interactions"):
        b.longrange_method = "matrix"
    mat, _ = _pair(2, 100)
    with pytest.raises(RuntimeError, match="unknown long-range method"):
        mat.longrange_method = "fft"
    mat.u = np.full((2, 100), np.nan)
    with pytest.raises(RuntimeError, match="NaN entries found"):
        mat.timeSteps(3)


def test_matrix_minimise_and_events():
    """minimise (detail.h:1754-1785) and the examples' quasistatic protocol
    of a batch in matrix form against the exact sequential sum: the same stop
    steps, avalanche sizes and wells, u within 1e-9 (the forces of the two
    forms differ by a few ulp per step, see the module docstring)."""
    mat, ext = _pair(4, 1000, nchunk=100)
    rng = np.random.default_rng(3)
    v0 = rng.normal(size=(4, 1000))
    for b in (mat, ext):
        b.v = v0
    rm, re = mat.minimise(), ext.minimise()
    assert np.all(rm == 0) and np.all(re == 0)
    assert np.array_equal(mat.inc, ext.inc)
    assert np.array_equal(mat.index_at_align, ext.index_at_align)
    assert np.allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    assert np.all(mat.v == 0) and np.all(mat.a == 0)  # quenched
    qm = mat.quasistatic(1e-3, 9, first_event=1)
    qe = ext.quasistatic(1e-3, 9, first_event=1)
    assert np.array_equal(qm["S"], qe["S"])
    assert np.array_equal(qm["inc"], qe["inc"])
    assert np.allclose(qm["u_frame"], qe["u_frame"], rtol=1e-12)
    assert np.any(qm["S"] != 0)
    assert np.allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
    assert np.allclose(mat.f, ext.f, rtol=1e-6, atol=1e-9)


def test_matrix_quasistatic_full_size():
    """The long-range configuration's chain (N = 8192, alpha 1.5, without the
    forcing, whose class has no minimise): a few events of the examples'
    protocol in matrix form equal the exact sum (tiled, one particle per
    thread) in avalanche sizes and wells."""
    print("long range quasistatic: R=2 N=8192", flush=True)
    mat, ext = _pair(2, 8192)
    qm = mat.quasistatic(1e-3, 5)
    qe = ext.quasistatic(1e-3, 5)
    assert np.array_equal(qm["S"], qe["S"])
    assert np.array_equal(qm["inc"], qe["inc"])
    assert np.array_equal(mat.index_at_align, ext.index_at_align)
    assert np.allclose(mat.u, ext.u, rtol=1e-9, atol=1e-9)
