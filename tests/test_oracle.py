"""CPU checks that pin the oracle (oracle/fqsb_oracle.c) before it is trusted.

* golden: the reference's own example outputs (examples/*.h5 -> tests/golden),
  exact avalanche sizes S and allclose u_frame / f_frame, as asserted by the
  reference's CI (.github/workflows/historic.yml:85-94);
* known answers: pcg32 (O'Neill's published pcg32 demo output), uint32->double;
* mathematics: log/exp/sin/cos against numpy within 1 ulp, pcg32 jump-ahead
  against stepping, landscape against a monolithic cumulative sum, the
  StopList against GooseFEM's documented ring semantics.
"""

import os

import numpy as np
import pytest

from tests import protocols as P

FULL = os.environ.get("FQSB_FULL_GOLDEN", "0") == "1"

# number of quasistatic steps checked by default (all of them with FQSB_FULL_GOLDEN=1)
PREFIX = {
    "Line1d_Cuspy_Laplace": 200,
    "Line1d_Cuspy_Quartic": 200,
    "Line1d_SemiSmooth_Laplace": 300,
    "Line1d_Cuspy_Laplace_LongRange": 16,
    "Line1d_Cuspy_Laplace_Nopassing": 1000,
    "Line2d_Cuspy_Laplace": 100,
    "Particles_Cuspy": 400,
}


@pytest.mark.parametrize("name", sorted(P.EXAMPLES))
def test_golden_examples(oracle, name):
    ref = P.golden(name)
    n = len(ref["S"]) if FULL else min(PREFIX[name], len(ref["S"]))
    system = P.example_system(oracle, name)
    u_frame, f_frame, S, _ = P.quasistatic(system, n)
    assert np.array_equal(S, ref["S"][:n])
    assert np.allclose(u_frame, ref["x_frame"][:n])
    assert np.allclose(f_frame, ref["f_frame"][:n])


def test_pcg32_known_answer(oracle):
    # pcg32-demo (pcg-c-basic) with pcg32_srandom_r(&rng, 42u, 54u)
    st, inc = oracle.pcg32_seed(42, 54)
    out, _ = oracle.pcg32_next(st, inc, 6)
    assert [int(x) for x in out] == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_pcg32_advance(oracle):
    st, inc = oracle.pcg32_seed(12345, 0)
    out, st_after = oracle.pcg32_next(st, inc, 1000)
    assert oracle.pcg32_advance(st, inc, 1000) == st_after
    assert oracle.pcg32_advance(st_after, inc, -1000) == st
    # stepping back once reproduces the last output
    prev = oracle.pcg32_advance(st_after, inc, -1)
    again, _ = oracle.pcg32_next(prev, inc, 1)
    assert again[0] == out[-1]


def test_uint32_to_double(oracle):
    lib = oracle.lib()
    assert lib.fqo_uint32_to_double(0) == 0.0
    assert lib.fqo_uint32_to_double(0xFFFFFFFF) == 1.0 - 2.0**-32
    assert lib.fqo_uint32_to_double(1 << 31) == 0.5
    x = np.random.default_rng(0).integers(0, 2**32, 1000, dtype=np.uint64)
    assert np.array_equal([lib.fqo_uint32_to_double(int(v)) for v in x], x.astype(float) / 2.0**32)


def test_uniform_draws(oracle):
    st, inc = oracle.pcg32_seed(7, 0)
    raw, _ = oracle.pcg32_next(st, inc, 100)
    d, _ = oracle.draw("random", [2.0], st, inc, 100)
    assert np.array_equal(d, 2.0 * (raw.astype(float) / 2.0**32))
    d, _ = oracle.draw("delta", [1.5], st, inc, 10)
    assert np.all(d == 1.5)


def _ulp_err(a, b):
    return np.max(np.abs(a - b) / np.spacing(np.abs(b)))


def test_log_exp(oracle):
    rng = np.random.default_rng(1)
    x = np.concatenate([rng.random(20000), rng.random(5000) * 1e3, np.exp(rng.uniform(-700, 700, 5000))])
    x = x[x > 0]
    assert _ulp_err(oracle.vec_math("log", x), np.log(x)) <= 1.0
    y = rng.uniform(-700, 700, 20000)
    assert _ulp_err(oracle.vec_math("exp", y), np.exp(y)) <= 1.0
    assert oracle.vec_math("log", [1.0])[0] == 0.0
    assert oracle.vec_math("log", [0.0])[0] == -np.inf


def test_sin_cos(oracle):
    rng = np.random.default_rng(2)
    x = rng.uniform(-2 * np.pi, 2 * np.pi, 40000)
    assert _ulp_err(oracle.vec_math("sin", x), np.sin(x)) <= 1.0
    assert _ulp_err(oracle.vec_math("cos", x), np.cos(x)) <= 1.0
    assert oracle.vec_math("sin", [0.0])[0] == 0.0


def test_weibull_transform(oracle):
    # weibull [k, scale, offset] = offset + scale * (-log(1 - r))^(1/k) (test_Line1d.py:289, 302)
    st, inc = oracle.pcg32_seed(3, 0)
    raw, _ = oracle.pcg32_next(st, inc, 5000)
    r = raw.astype(float) / 2.0**32
    d, _ = oracle.draw("weibull", [2.0, 1.1, 1e-3], st, inc, 5000)
    ref = 1e-3 + 1.1 * np.sqrt(-np.log(1.0 - r))
    assert np.allclose(d, ref, rtol=2e-16 * 4, atol=0)
    d3, _ = oracle.draw("weibull", [3.0, 1.0], st, inc, 5000)
    assert np.allclose(d3, np.power(-np.log(1.0 - r), 1.0 / 3.0), rtol=1e-15, atol=1e-300)


def _stoplist_ring(residuals, n, tol):
    """GooseFEM Iterate::StopList: ring initialised to +inf; descending() is
    strictly decreasing; all_less(x) is 'no entry >= x' (detail.h:1613-1615)."""
    ring = [np.inf] * n
    for i, r in enumerate(residuals):
        ring = ring[1:] + [r]
        desc = all(not (ring[k] >= ring[k - 1]) for k in range(1, n))
        if (desc and all(not (x >= tol) for x in ring)) or all(not (x >= tol * tol) for x in ring):
            return i
    return -1


def _stoplist_counters(residuals, n, tol):
    """Counter form used by the resident kernel (fq_resident.hpp)."""
    below = below2 = dec = 0
    prev = np.inf
    for i, r in enumerate(residuals):
        below = below + 1 if not (r >= tol) else 0
        below2 = below2 + 1 if not (r >= tol * tol) else 0
        dec = dec + 1 if not (r >= prev) else 0
        prev = r
        if n <= 0 or (dec >= n - 1 and below >= n) or below2 >= n:
            return i
    return -1


def test_stoplist(oracle):
    rng = np.random.default_rng(4)
    for trial in range(300):
        n = int(rng.integers(1, 12))
        m = int(rng.integers(1, 60))
        tol = 1e-2
        # decaying noisy sequences with plateaus and exact ties
        r = np.exp(-rng.uniform(0, 0.4, m).cumsum()) * rng.choice([1.0, 1.0, 1.5], m)
        if trial % 5 == 0:
            r[m // 2 :] = r[m // 2]
        ref = _stoplist_ring(list(r), n, tol)
        assert oracle.stoplist_run(r, n, tol) == ref
        assert _stoplist_counters(list(r), n, tol) == ref


def _monolithic_landscape(oracle, seed, n_first, n_total, offset, distribution, parameters):
    """Yield positions of one particle computed in one pass: the first window is
    cumsum(increments) + offset, beyond it the running sum continues."""
    st, inc = oracle.pcg32_seed(seed, 0)
    d, _ = oracle.draw(distribution, parameters, st, inc, n_total)
    y = np.empty(n_total)
    y[:n_first] = np.cumsum(d[:n_first]) + offset
    for g in range(n_first, n_total):
        y[g] = y[g - 1] + d[g]
    return y


def test_landscape_forward(oracle):
    # forward sweeps through many chunk refills (test_Line1d.py:277-341, forward part)
    N, nchunk = 3, 50
    par = P.delta_par(N, distribution="weibull", parameters=[2.0, 1.1, 1e-3], offset=-50.0, nchunk=nchunk, seed=11)
    s = oracle.Line1d.System_Cuspy_Laplace(**par)
    ys = [_monolithic_landscape(oracle, 11 + p, nchunk, 4000, -50.0, "weibull", [2.0, 1.1, 1e-3]) for p in range(N)]
    du = np.array([5.0, 7.0, 10.0])
    for i in range(300):
        s.u = i * du
        idx = s.chunk.index_at_align
        for p in range(N):
            j = np.searchsorted(ys[p], s.u[p], side="right") - 1
            assert idx[p] == j
            assert s.chunk.left_of_align[p] == ys[p][j]
            assert s.chunk.right_of_align[p] == ys[p][j + 1]
        umin = 0.5 * (s.chunk.left_of_align + s.chunk.right_of_align)
        assert np.allclose(s.f_potential, par["mu"] * (umin - s.u))


def test_landscape_backward_and_restore(oracle):
    # backward sweep and chunk restore (test_Line1d.py:309-341)
    N = 3
    par = P.delta_par(N, distribution="weibull", parameters=[2.0, 1.1, 1e-3], offset=-50.0, nchunk=100, seed=5)
    s = oracle.Line1d.System_Cuspy_Laplace(**par)
    ys = [_monolithic_landscape(oracle, 5 + p, 100, 20000, -50.0, "weibull", [2.0, 1.1, 1e-3]) for p in range(N)]
    du = np.array([5.0, 7.0, 10.0])
    saved = None
    for i in list(range(400)) + list(range(400))[::-1]:
        s.u = i * du
        idx = s.chunk.index_at_align
        for p in range(N):
            j = np.searchsorted(ys[p], s.u[p], side="right") - 1
            assert idx[p] == j
            assert np.isclose(s.chunk.left_of_align[p], ys[p][j])
            assert np.isclose(s.chunk.right_of_align[p], ys[p][j + 1])
        if i == 250 and saved is None:
            start = s.chunk.start.copy()
            saved = (s.u.copy(), start, s.chunk.state_at(start).copy(), s.chunk.data[..., 0].copy(), idx.copy())
    u, start, state, value, idx = saved
    s.chunk.restore(state=state, value=value, index=start)
    s.u = u
    assert np.array_equal(s.chunk.index_at_align, idx)


def test_errors(oracle):
    s = oracle.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.full(5, np.nan)
    with pytest.raises(RuntimeError, match="NaN entries found"):
        s.timeStep()
    s = oracle.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.array([0.3, 0, 0, 0, 0])
    with pytest.raises(RuntimeError, match="No convergence found"):
        s.minimise(max_iter=1)
    s = oracle.Line1d.System_Cuspy_Laplace(**P.delta_par(5))
    s.u = np.array([0.3, 0, 0, 0, 0])
    assert s.minimise(max_iter=1, max_iter_is_error=False) == 2
    smooth = oracle.Line1d.System_Smooth_Laplace(**P.delta_par(5))
    with pytest.raises(RuntimeError, match="Operation not possible."):
        smooth.maxUniformDisplacement()
    with pytest.raises(RuntimeError, match="Unknown distribution: foo"):
        oracle.Line1d.System_Cuspy_Laplace(**P.delta_par(5, distribution="foo"))


def circulant_longrange(u, k, alpha):
    """LongRange1d forces (detail.h:829-867) through the eigen-form of the
    circulant matrix: f = g (*) u - c u, with g[m] = k / fold(m)^(alpha+1)
    (g[0] = 0), fold(m) = min(m, N - m), c = sum(g), (*) a circular
    convolution evaluated by FFT. Rows of u are independent systems."""
    u = np.atleast_2d(u)
    N = u.shape[-1]
    m = np.arange(N)
    d = np.minimum(m, N - m).astype(float)
    g = np.zeros(N)
    g[1:] = k / d[1:] ** (alpha + 1.0)
    conv = np.fft.irfft(np.fft.rfft(g) * np.fft.rfft(u, axis=-1), n=N, axis=-1)
    return conv - g.sum() * u, g


@pytest.mark.parametrize("N", [7, 10, 64, 301])
def test_longrange_force_eigen_form(oracle, N):
    """The oracle's sequential LongRange1d sum equals the FFT eigen-form and a
    dense-matrix product, within the summation error bound gamma_N * sum|term|."""
    rng = np.random.default_rng(N)
    k, alpha = 0.7, 1.5
    s = oracle.Line1d.System_Cuspy_LongRange(
        m=1.0, eta=0.1, mu=1.0, k_interactions=k, alpha=alpha, k_frame=0.1, dt=0.1, shape=[N], seed=0,
        distribution="random", parameters=[2.0], offset=-5.0, nchunk=100,
    )
    u = rng.normal(size=N) * 3.0
    s.u = u
    f_oracle = s.f_interactions
    f_fft, g = circulant_longrange(u, k, alpha)
    idx = np.arange(N)
    K = g[(idx[:, None] - idx[None, :]) % N]
    f_dense = K @ u - g.sum() * u
    scale = np.abs(K) @ np.abs(u) + g.sum() * np.abs(u)
    gamma = 4 * N * np.finfo(float).eps
    assert np.all(np.abs(f_oracle - f_dense) <= gamma * scale)
    assert np.all(np.abs(f_oracle - f_fft[0]) <= gamma * scale * np.log2(N) + 1e-15 * np.max(scale))
    # a dropped self term or a wrong fold would fail: perturb and recheck
    g_bad = g.copy()
    g_bad[1] *= 1.5
    K_bad = g_bad[(idx[:, None] - idx[None, :]) % N]
    assert not np.allclose(K_bad @ u - g_bad.sum() * u, f_oracle)
