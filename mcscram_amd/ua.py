"""Uncertainty analysis over basic-event probabilities.

The reference's outer loop (UncertaintyAnalyzer::Sample,
src/uncertainty_analysis.h:154-169) draws every basic event's deviate, clamps
it to [0, 1] (src/uncertainty_analysis.cc:108-120) and runs the probability
calculator once per sample; the statistics of the samples are computed by
UncertaintyAnalysis::CalculateStatistics (src/uncertainty_analysis.cc:
123-157). Here the calculator is the GPU Monte Carlo engine, run for many
parameter samples at once (``UncertaintyEngine``, the C ABI's mcs_ua_*).

Lognormal deviates follow src/expression/random_deviate.cc:91-108 with the
median parametrisation used by the workloads: ln p = ln(median) + sigma*Z,
sigma = ln(EF) / z(level), z(0.95) = 1.6448536... The normal draws come from
numpy's PCG64 with a fixed seed (the reference's std::mt19937 stream is not
reproduced: parameter sampling is not part of the bit-exact path).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import _check, _p, _PdagView, _shape, load_library, MCS_REFERENCE_KEY
from .pdag import Pdag, SplitMix64, large_tree, random_layered_tree

Z95 = 1.6448536269514722  # standard normal quantile of 0.95


class _UaConfig(C.Structure):
    _fields_ = [("key0", C.c_uint32), ("key1", C.c_uint32),
                ("common_random_numbers", C.c_uint32)]


class UncertaintyEngine:
    """K parameter samples of one tree on the GPU, each an independent plain
    Monte Carlo run (Philox key (key0 + sample_id, key1)).

    ``p_samples``: [K, n_vars] probabilities in the order of ``pdag.var_index``.
    ``tally``: PDAG indices counted in every sample (default: the root).
    """

    def __init__(self, pdag: Pdag, shape, p_samples, sample_ids=None, tally=None,
                 key=MCS_REFERENCE_KEY, common_random_numbers=False):
        L = load_library()
        self._view = _PdagView(pdag)
        self._shape = _shape(shape)
        self.p = np.ascontiguousarray(p_samples, dtype=np.float64)
        if self.p.ndim != 2 or self.p.shape[1] != pdag.n_vars:
            raise ValueError("p_samples must be [n_samples, n_vars]")
        self.n_samples = int(self.p.shape[0])
        ids = None if sample_ids is None else np.ascontiguousarray(sample_ids, dtype=np.uint32)
        self._ids = ids
        self.tally = list(tally) if tally else [pdag.root]
        t = np.ascontiguousarray(self.tally, dtype=np.int32)
        self._t = t
        cfg = _UaConfig(key[0] & 0xFFFFFFFF, key[1] & 0xFFFFFFFF, 1 if common_random_numbers else 0)
        h = C.c_void_p()
        _check(L.mcs_ua_create(C.byref(self._view.struct), C.byref(self._shape), self.n_samples,
                               _p(self.p), _p(ids) if ids is not None else None, _p(t), t.shape[0],
                               C.byref(cfg), C.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None):
            load_library().mcs_ua_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, n_passes: int = 1) -> None:
        _check(load_library().mcs_ua_run(self._h, n_passes))

    def time_run(self, n_passes: int = 1) -> float:
        ms = C.c_double()
        _check(load_library().mcs_ua_time_run(self._h, n_passes, C.byref(ms)))
        return ms.value

    def collect(self):
        """-> (ones [n_samples, n_tally] uint64, total trials per sample)."""
        out = np.zeros((self.n_samples, len(self.tally)), dtype=np.uint64)
        total = C.c_uint64()
        _check(load_library().mcs_ua_collect(self._h, _p(out), out.size, C.byref(total)))
        return out, total.value


@dataclass
class LognormalModel:
    """Per basic event: ln p ~ N(ln median, sigma^2), sigma = ln(EF) / z95."""

    median: np.ndarray
    error_factor: np.ndarray

    @property
    def mu(self):
        return np.log(self.median)

    @property
    def sigma(self):
        return np.log(self.error_factor) / Z95

    def sample(self, n: int, seed: int):
        """-> (p [n, n_vars] clamped to [0, 1], z [n, n_vars] standard normals)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        z = rng.standard_normal((n, self.median.shape[0]))
        p = np.exp(self.mu[None, :] + self.sigma[None, :] * z)
        return np.clip(p, 0.0, 1.0), z


def uncertainty_tree(seed: int = 4, n_be: int = 2000, n_gates: int = 3000):
    """The small tree's family scaled up (2,000 events, 3,000 gates in 8
    layers, 45 % AND / 45 % OR / 10 % ATLEAST, fan-in 2-4, single OR top
    gate over the last layer), with lognormal uncertainty: median
    log-uniform in [1e-5, 1e-2], error factor uniform in [3, 10].
    -> (pdag with the medians as p, LognormalModel)."""
    pd = random_layered_tree(seed, n_be, n_gates, 8, p_range=(1e-5, 1e-2), fan_in=(2, 4),
                             mix=(0.45, 0.45, 0.10), atleast_n=(3, 6), top="last_layer")
    rng = SplitMix64(seed ^ 0x5EED)
    ef = np.array([3.0 + 7.0 * rng.uniform() for _ in range(n_be)])
    return pd, LognormalModel(pd.var_p.copy(), ef)


# The large tree's own top gate is an OR of 11,015 gates, 3,563 of which have
# probability above 0.9 (complemented rare inputs make 11,492 of its 40,000
# gates that likely). Its top event held in every one of 2^18 trials at the
# medians and at four sampled parameter sets, so every Monte Carlo estimate
# is (trials + 1/2) / (trials + 1) and there is nothing to analyse. Gate
# 64134 (an OR; 210 gates and 956 basic events below it; about 6e-3 at the
# medians) has the most basic events below it among the gates of that tree
# whose probability at the medians lies in [3e-3, 0.15].
LARGE_TREE_UA_TOP = 64134


def large_tree_uncertainty(top: int = LARGE_TREE_UA_TOP, seed: int = 2):
    """The large tree (25,000 events, 40,000 gates) with lognormal
    uncertainty on every basic event (median = the tree's probability,
    log-uniform in [1e-6, 1e-2]; error factor uniform in [3, 10]), analysed
    for gate ``top``: -> (that gate's PDAG with the medians as p,
    LognormalModel of its basic events)."""
    full = large_tree(seed)
    rng = SplitMix64(seed ^ 0x5EED)
    ef_all = np.array([3.0 + 7.0 * rng.uniform() for _ in range(full.n_vars)])
    pd = full.subgraph(top)
    ef = ef_all[np.isin(full.var_index, pd.var_index)]
    return pd, LognormalModel(pd.var_p.copy(), ef)


def statistics(samples, n_quantiles: int = 20) -> dict:
    """Statistics of the sampled top-event probabilities, as the reference's
    CalculateStatistics (src/uncertainty_analysis.cc:123-157): mean, sigma
    with the n/(n-1) correction, error factor exp(1.96 sigma), the 95 %
    confidence interval of the mean, and quantiles at i/n_quantiles. The
    reference estimates quantiles with Boost's extended P^2 algorithm; these
    are the exact empirical quantiles (numpy 'linear')."""
    x = np.asarray(samples, dtype=np.float64)
    n = x.shape[0]
    mean = float(x.mean())
    sigma = float(np.sqrt(n * x.var() / (n - 1))) if n > 1 else 0.0
    probs = [(i + 1) / n_quantiles for i in range(n_quantiles)]
    return {"mean": mean, "sigma": sigma, "error_factor": float(np.exp(1.96 * sigma)),
            "ci95": (mean - 1.96 * sigma / np.sqrt(n), mean + 1.96 * sigma / np.sqrt(n)),
            "quantiles": dict(zip(probs, np.quantile(x, probs).tolist())),
            "p05": float(np.quantile(x, 0.05)), "p50": float(np.quantile(x, 0.50)),
            "p95": float(np.quantile(x, 0.95)), "n": n}
