"""Pins of the oracle's random draws that the golden files leave open, and of
the StopList decision margin used by every GPU test.

The reference's landscapes and thermal forces come from prrng (absent here,
see the oracle's header). What a plausible mistake in the restatement would
change is checked against mathematics that does not depend on the oracle's
own formulas:

* thermal draw (RandomNormalForcing::force, detail.h:930-943): the values are
  N(mean, stddev) -- moments, Kolmogorov-Smirnov statistic and the absence of
  serial correlation over 10^6 draws (a wrong sigma, a dropped sqrt(2), a
  reused uniform or a cos/sin with the wrong period all fail);
* draw order (the reference's own test, test_Line1d.py:566-600, with
  prrng.pcg32(0).normal replaced by the oracle's stream-level draw): one
  shared generator seeded with seed_forcing, the due particles drawing in
  particle order, a second refresh continuing the stream, particles not due
  drawing nothing;
* Weibull increments (prrng's weibull [k, scale, offset] as used by every
  configuration, test_Line1d.py:289): KS and moments against scipy's
  Weibull law, for the stand-alone draw and for a system's landscape
  (differences of chunk.data), with the parameter order checked by a
  non-unit scale;
* StopList margin (tests/conftest.py): checked against perturbed sequences.
"""

import numpy as np
import pytest
from scipy import special, stats

from tests import protocols as P


def _ks_pvalue(x, cdf):
    return stats.kstest(x, cdf).pvalue


@pytest.mark.parametrize("mean,stddev", [(0.0, 1.0), (0.3, 0.1), (-2.0, 2.5)])
def test_thermal_draw_distribution(oracle, mean, stddev):
    n = 1_000_000
    st, inc = oracle.forcing_stream(7)
    x, _ = oracle.draw_normal(st, inc, mean, stddev, n)
    z = (x - mean) / stddev
    se = 1.0 / np.sqrt(n)
    assert abs(z.mean()) < 5 * se
    assert abs(z.std() - 1.0) < 5 * se / np.sqrt(2.0)
    assert abs(stats.skew(z)) < 5 * np.sqrt(6.0) * se
    assert abs(stats.kurtosis(z)) < 5 * np.sqrt(24.0) * se
    assert _ks_pvalue(z, "norm") > 1e-4
    # tails: the fraction beyond 3 sigma (0.27 %)
    tail = np.mean(np.abs(z) > 3.0)
    p3 = 2.0 * stats.norm.sf(3.0)
    assert abs(tail - p3) < 5 * np.sqrt(p3 / n)
    # successive draws are independent
    assert abs(np.corrcoef(z[:-1], z[1:])[0, 1]) < 5 * se
    assert abs(np.corrcoef(z[:-2], z[2:])[0, 1]) < 5 * se


def test_thermal_draw_consumes_two_uniforms(oracle):
    # each normal takes two values of the stream: after n draws the generator
    # stands 2n values further (jump-ahead, O'Neill's published LCG)
    st, inc = oracle.forcing_stream(3)
    _, st_after = oracle.draw_normal(st, inc, 0.0, 1.0, 1000)
    assert st_after == oracle.pcg32_advance(st, inc, 2000)


def _forcing_system(oracle, N, dinc_init, dinc, mean=0.0, stddev=1.0, seed_forcing=0):
    kw = dict(
        mean=mean,
        stddev=stddev,
        seed_forcing=seed_forcing,
        dinc_init=np.asarray(dinc_init, dtype=np.int64),
        dinc=np.asarray(dinc, dtype=np.int64),
    )
    return oracle.Line1d.System_Cuspy_Laplace_RandomForcing(**P.delta_par(N), **kw)


def test_thermal_draw_order_reference_case(oracle):
    """test_Line1d.py:566-600: after each `inc += 1; refresh()` every particle
    holds the next N normals of one generator seeded with seed_forcing = 0."""
    N = 10
    s = _forcing_system(oracle, N, np.ones(N), np.ones(N))
    assert s.residual < 1e-5
    st, inc = oracle.forcing_stream(0)
    for _ in range(2):
        s.inc = s.inc + 1
        s.refresh()
        assert s.residual > 1e-5
        ref, st = oracle.draw_normal(st, inc, 0.0, 1.0, N)
        assert np.array_equal(s.external.f_thermal, ref)
        assert s.external.state == st


def test_thermal_draw_order_partial(oracle):
    """Uneven dinc: at every increment only the due particles draw, in
    particle order, from the one shared stream; the others keep their force
    and their next increment."""
    rng = np.random.default_rng(5)
    N = 37
    dinc_init = rng.integers(1, 4, N)
    dinc = rng.integers(1, 5, N)
    mean, stddev = 0.2, 0.7
    s = _forcing_system(oracle, N, dinc_init, dinc, mean, stddev, seed_forcing=11)
    st, inc = oracle.forcing_stream(11)
    nxt = dinc_init.copy()
    f = np.zeros(N)
    for i in range(1, 25):
        s.inc = i
        s.refresh()
        due = np.flatnonzero(i >= nxt)
        vals, st = oracle.draw_normal(st, inc, mean, stddev, due.size)
        f[due] = vals
        nxt[due] += dinc[due]
        assert np.array_equal(s.external.f_thermal, f), i
        assert np.array_equal(s.external.next, nxt), i
        assert s.external.state == st, i


def test_thermal_forces_of_a_system_are_normal(oracle):
    # system-level: the forces a chain receives over many increments follow
    # N(mean, stddev) (catches mean / stddev swapped or ignored by the system)
    N = 2000
    mean, stddev = 0.25, 0.1
    s = _forcing_system(oracle, N, np.ones(N), np.ones(N), mean, stddev, seed_forcing=2)
    draws = []
    for i in range(1, 51):
        s.inc = i
        s.refresh()
        draws.append(s.external.f_thermal.copy())
    z = (np.concatenate(draws) - mean) / stddev
    assert abs(z.mean()) < 5 / np.sqrt(z.size)
    assert abs(z.std() - 1.0) < 5 / np.sqrt(2 * z.size)
    assert _ks_pvalue(z, "norm") > 1e-4


@pytest.mark.parametrize("k,scale,offset", [(2.0, 1.0, 0.0), (3.0, 1.7, 0.0), (2.0, 0.5, 1e-3)])
def test_weibull_increments_distribution(oracle, k, scale, offset):
    n = 400_000
    st, inc = oracle.pcg32_seed(13, 0)
    d, _ = oracle.draw("weibull", [k, scale, offset], st, inc, n)
    law = stats.weibull_min(c=k, scale=scale, loc=offset)
    assert _ks_pvalue(d, law.cdf) > 1e-4
    mean = offset + scale * special.gamma(1.0 + 1.0 / k)
    var = scale**2 * (special.gamma(1.0 + 2.0 / k) - special.gamma(1.0 + 1.0 / k) ** 2)
    assert abs(d.mean() - mean) < 5 * np.sqrt(var / n)
    assert abs(d.var() / var - 1.0) < 0.02
    assert d.min() >= offset


@pytest.mark.parametrize("k,scale", [(2.0, 1.0), (3.0, 1.7)])
def test_weibull_landscape_of_a_system(oracle, k, scale):
    """The increments of a system's yield positions (parameters [k, scale,
    offset], the order of test_Line1d.py:289) are Weibull(k, scale)."""
    N, nchunk = 400, 600
    par = P.delta_par(N, distribution="weibull", parameters=[k, scale, 0.0], offset=-100.0, nchunk=nchunk, seed=3)
    s = oracle.Line1d.System_Cuspy_Laplace(**par)
    d = np.diff(s.chunk.data, axis=1).ravel()
    law = stats.weibull_min(c=k, scale=scale)
    assert _ks_pvalue(d, law.cdf) > 1e-4
    # the swapped parameter order is rejected by the same test
    wrong = stats.weibull_min(c=scale, scale=k)
    assert _ks_pvalue(d, wrong.cdf) < 1e-10


def test_stoplist_margin_near_tie(oracle):
    tol, n = 1e-2, 3
    r = np.array([1.0, 0.5, 0.1, 0.05, 0.02, 0.009, 0.008, 0.007])
    assert oracle.stoplist_run(r, n, tol) == 7
    assert oracle.stoplist_margin(r, n, tol) > 1e-2
    # one residual a hair below tol: the stop step hangs on it
    r2 = r.copy()
    r2[5] = tol * (1.0 - 1e-14)
    assert oracle.stoplist_run(r2, n, tol) == 7
    assert oracle.stoplist_margin(r2, n, tol) < 1e-13
    # a near-tie of two successive residuals below tol
    r3 = r.copy()
    r3[6] = r3[5] * (1.0 - 1e-15)
    assert oracle.stoplist_margin(r3, n, tol) < 1e-14


def test_stoplist_margin_is_sound(oracle):
    """Perturbing every residual by a relative amount below half the margin
    never changes the stop step (random decaying sequences with plateaus)."""
    rng = np.random.default_rng(9)
    tol = 1e-2
    for trial in range(400):
        n = int(rng.integers(1, 6))
        m = int(rng.integers(4, 40))
        r = np.exp(-rng.uniform(0, 0.5, m).cumsum()) * rng.choice([1.0, 1.0, 1.3], m)
        stop = oracle.stoplist_run(r, n, tol)
        margin = oracle.stoplist_margin(r, n, tol)
        if not np.isfinite(margin) or margin > 0.5:
            continue
        for _ in range(20):
            d = rng.uniform(-1, 1, m) * 0.45 * margin
            assert oracle.stoplist_run(r * (1.0 + d), n, tol) == stop, (trial, margin)
