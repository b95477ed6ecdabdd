"""Stopping rules of the probability analysis (host code, no GPU).

Pins:
* the Beta quantile against scipy.stats.beta.ppf, which evaluates Boost.Math's
  inverse incomplete beta, the function the reference's Bayes rule calls
  (src/mc/scheduler/convergence_policy.h:97-99), from a handful of trials to
  the 1e12 trials of a long run, and against the closed form of Beta(1, 1);
* the Wald update against its closed forms (src/mc/stats/ci_utils.h:118-143;
  convergence_policy.h:28-61);
* the Bayes update against the same formulas evaluated with SciPy
  (convergence_policy.h:66-143).
"""
import math

import numpy as np
import pytest
from scipy import stats

import mcscram_amd as mc

Z99 = 2.5758293035489004
LN10 = math.log(10.0)


@pytest.fixture(scope="module")
def lib():
    return mc.load_library()


CASES = [(0.5, 0.5), (1.5, 3.5), (10.5, 1000.5), (0.5, 1e6 + 0.5), (3.5, 2**30),
         (1234.5, 2**32), (4.1e6, 2**40), (2**31, 2**31), (7.5, 2.5), (5e8, 5e11)]


@pytest.mark.parametrize("a, b", CASES)
@pytest.mark.parametrize("q", [0.005, 0.025, 0.5, 0.975, 0.995])
def test_beta_quantile_matches_boost(lib, a, b, q):
    got = mc.beta_quantile(a, b, q)
    want = stats.beta.ppf(q, a, b)
    sd = math.sqrt(a * b / ((a + b) ** 2 * (a + b + 1)))
    # Agreement to 2e-7 of the distribution's spread (or 1e-12 relative), so
    # the half-widths the stopping rule compares agree to ~1e-7 relative. At
    # a + b above 1e9 Boost's own result moves by ~1e-7 spreads: its median
    # of Beta(2^31, 2^31) is 0.5 + 9.3e-13, where symmetry fixes 0.5 exactly
    # (see the next test).
    assert abs(got - want) <= max(2e-7 * sd, 1e-12 * want), (got, want, sd)


@pytest.mark.parametrize("a", [0.5, 3.0, 1e4 + 0.5, 2.0**31])
def test_beta_quantile_symmetries(lib, a):
    # Beta(a, a) has median 1/2; Q(a, b, q) = 1 - Q(b, a, 1 - q).
    assert mc.beta_quantile(a, a, 0.5) == pytest.approx(0.5, abs=1e-15)
    for b, q in ((a * 7 + 1, 0.005), (a + 2, 0.975)):
        assert mc.beta_quantile(a, b, q) == pytest.approx(1 - mc.beta_quantile(b, a, 1 - q),
                                                            abs=1e-15)


def test_beta_quantile_closed_form_and_edges(lib):
    for q in (0.0, 0.1, 0.5, 0.9, 1.0):
        assert mc.beta_quantile(1.0, 1.0, q) == pytest.approx(q, abs=1e-15)
    # Beta(a, 1): F(x) = x^a.
    assert mc.beta_quantile(3.0, 1.0, 0.2) == pytest.approx(0.2 ** (1 / 3), rel=1e-13)
    for bad in ((0.0, 1.0, 0.5), (1.0, -1.0, 0.5), (1.0, 1.0, 1.5)):
        assert math.isnan(mc.beta_quantile(*bad))


@pytest.mark.parametrize("ones, total", [(0, 1 << 20), (17, 1 << 20), (5000, 1 << 24),
                                         (1 << 29, 1 << 30), ((1 << 30) - 3, 1 << 30)])
def test_wald_update_closed_form(lib, ones, total):
    c = mc.convergence_update("wald", ones, total, 0.001, Z99)
    mean = ones / total
    p = max(mean, 1e-12)
    se = math.sqrt(mean * (1 - mean) / total)
    assert c["epsilon"] == pytest.approx(Z99 * se, rel=1e-14, abs=1e-300)
    assert c["target_epsilon"] == pytest.approx(0.001 * p, rel=1e-15)
    assert c["epsilon_log10"] == pytest.approx(Z99 * se / (p * LN10), rel=1e-14, abs=1e-300)
    assert c["target_epsilon_log10"] == 0.001
    n_lin = Z99 ** 2 * p * (1 - p) / (0.001 * p) ** 2
    n_log = Z99 ** 2 * (1 - p) / (p * 0.001 ** 2 * LN10 ** 2)
    assert abs(c["target_trials"] - math.ceil(n_lin)) <= max(1, 1e-15 * n_lin)
    assert abs(c["target_trials_log10"] - math.ceil(n_log)) <= max(1, 1e-15 * n_log)


@pytest.mark.parametrize("ones, total", [(0, 1 << 20), (17, 1 << 20), (5000, 1 << 24),
                                         (1 << 29, 1 << 30), (123456, 1 << 36)])
def test_bayes_update_matches_boost_posterior(lib, ones, total):
    c = mc.convergence_update("bayes", ones, total, 0.001, Z99)
    a, b = 0.5 + ones, 0.5 + (total - ones)
    cred = 2 * stats.norm.cdf(Z99) - 1
    tail = (1 - cred) / 2
    lo, hi = stats.beta.ppf([tail, 1 - tail], a, b)
    p = a / (a + b)
    eps = (hi - lo) / 2
    assert c["epsilon"] == pytest.approx(eps, rel=1e-8)
    assert c["epsilon_log10"] == pytest.approx(eps / (p * LN10), rel=1e-8)
    assert c["target_epsilon"] == pytest.approx(0.001 * p, rel=1e-15)
    n_lin = Z99 ** 2 * p * (1 - p) / (0.001 * p) ** 2 - 2
    n_log = Z99 ** 2 * (1 - p) / (p * 0.001 ** 2 * LN10 ** 2) - 2
    for got, n in ((c["target_trials"], n_lin), (c["target_trials_log10"], n_log)):
        assert abs(got - max(math.ceil(max(n, 0)), total)) <= max(1, 1e-15 * n)


def test_bayes_interval_approaches_wald_at_large_counts(lib):
    # Both are central intervals at the same level: for np(1-p) large they agree.
    w = mc.convergence_update("wald", 1 << 28, 1 << 34, 0.001, Z99)
    b = mc.convergence_update("bayes", 1 << 28, 1 << 34, 0.001, Z99)
    assert b["epsilon"] == pytest.approx(w["epsilon"], rel=1e-4)


def test_bad_policy_is_rejected(lib):
    with pytest.raises(ValueError):
        mc.Settings(ci_policy="frequentist")._c()
    with pytest.raises(mc.engine.EngineError):
        from mcscram_amd.engine import _Convergence, _check
        import ctypes as C
        out = _Convergence()
        _check(lib.mcs_convergence_update(7, 1, 2, 0.001, Z99, C.byref(out)))
    with pytest.raises(mc.engine.EngineError):
        mc.convergence_update("wald", 3, 2, 0.001, Z99)


def test_beta_quantile_is_fast_at_long_run_counts(lib):
    import time
    t = time.perf_counter()
    for k in range(20):
        mc.convergence_update("bayes", 1000 + k, 1 << 40, 0.001, Z99)
    assert time.perf_counter() - t < 2.0
