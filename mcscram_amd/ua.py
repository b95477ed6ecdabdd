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
sigma = ln(EF) / z(level), z(0.95) = 1.6448536... The workloads draw them on
the GPU (``LognormalModel.draw``: Philox4x32-10 and Box-Muller, counter-based
per row); ``LognormalModel.sample`` draws them on the host from numpy's PCG64.
The reference's std::mt19937 stream is not reproduced: parameter sampling is
not part of the bit-exact path.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from .engine import _check, _p, _PdagView, _shape, load_library, MCS_REFERENCE_KEY
from .pdag import Pdag, SplitMix64, random_layered_tree

Z95 = 1.6448536269514722  # standard normal quantile of 0.95


class _UaConfig(C.Structure):
    _fields_ = [("key0", C.c_uint32), ("key1", C.c_uint32),
                ("common_random_numbers", C.c_uint32), ("samples_per_launch", C.c_uint32),
                ("pipelines", C.c_uint32), ("chunk_words", C.c_uint32)]


class UncertaintyEngine:
    """K parameter samples of one tree on the GPU, each an independent plain
    Monte Carlo run (Philox key (key0 + sample_id, key1)).

    ``p_samples``: [K, n_vars] probabilities in the order of ``pdag.var_index``.
    ``tally``: PDAG indices counted in every sample (default: the root).
    ``samples_per_launch``, ``pipelines``, ``chunk_words``: the grouping
    and the words per task (0: the engine's choice); they change no bit of
    any sample's tallies.
    """

    def __init__(self, pdag: Pdag, shape, p_samples, sample_ids=None, tally=None,
                 key=MCS_REFERENCE_KEY, common_random_numbers=False, samples_per_launch=0,
                 pipelines=0, chunk_words=0):
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
        cfg = _UaConfig(key[0] & 0xFFFFFFFF, key[1] & 0xFFFFFFFF, 1 if common_random_numbers else 0,
                        int(samples_per_launch), int(pipelines), int(chunk_words))
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

    def set_samples(self, p_samples, sample_ids=None) -> None:
        """Replace the samples (mcs_ua_set_samples): the engine keeps its
        device arenas and behaves as one created with these samples."""
        p = np.ascontiguousarray(p_samples, dtype=np.float64)
        if p.ndim != 2 or p.shape[1] != self.p.shape[1]:
            raise ValueError("p_samples must be [n_samples, n_vars]")
        ids = None if sample_ids is None else np.ascontiguousarray(sample_ids, dtype=np.uint32)
        if ids is not None and ids.shape != (p.shape[0],):
            raise ValueError("sample_ids must hold one id per sample")
        self.p, self._ids = p, ids
        self.n_samples = int(p.shape[0])
        _check(load_library().mcs_ua_set_samples(self._h, self.n_samples, _p(p),
                                                 _p(ids) if ids is not None else None))

    def run(self, n_passes: int = 1) -> None:
        _check(load_library().mcs_ua_run(self._h, n_passes))

    def time_run(self, n_passes: int = 1) -> float:
        ms = C.c_double()
        _check(load_library().mcs_ua_time_run(self._h, n_passes, C.byref(ms)))
        return ms.value

    @property
    def info(self) -> dict:
        """Samples per launch and pipelines the engine chose, device bytes."""
        g, p, c, b = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        _check(load_library().mcs_ua_info(self._h, C.byref(g), C.byref(p), C.byref(c), C.byref(b)))
        return {"samples_per_launch": g.value, "pipelines": p.value, "chunk_words": c.value,
                "device_bytes": b.value}

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

    def sample_chunks(self, n: int, seed: int, chunk: int):
        """The samples of ``sample(n, seed)`` in consecutive row blocks of at
        most ``chunk`` rows (one generator, drawn in order, so the rows are
        the same). Yields (first row, p, z)."""
        rng = np.random.Generator(np.random.PCG64(seed))
        for a in range(0, n, chunk):
            z = rng.standard_normal((min(chunk, n - a), self.median.shape[0]))
            p = np.exp(self.mu[None, :] + self.sigma[None, :] * z)
            yield a, np.clip(p, 0.0, 1.0), z

    def draw(self, n: int, seed: int, first_row: int = 0):
        """Rows first_row .. first_row + n - 1 of the GPU's counter-based
        stream (mcs_lognormal_samples: Philox4x32-10 and Box-Muller; a row is
        the same whichever call draws it). -> (p [n, n_vars] float64 clamped
        to 1, z [n, n_vars] float32)."""
        nv = int(self.median.shape[0])
        mu = np.ascontiguousarray(self.mu, dtype=np.float64)
        sigma = np.ascontiguousarray(self.sigma, dtype=np.float64)
        z = np.empty((n, nv), dtype=np.float32)
        p = np.empty((n, nv), dtype=np.float64)
        _check(load_library().mcs_lognormal_samples(_p(mu), _p(sigma), nv, int(seed), int(first_row),
                                                    int(n), _p(z), _p(p)))
        return p, z

    def draw_chunks(self, n: int, seed: int, chunk: int):
        """``draw(n, seed)`` in consecutive row blocks of at most ``chunk``
        rows. Yields (first row, p, z)."""
        for a in range(0, n, chunk):
            p, z = self.draw(min(chunk, n - a), seed, a)
            yield a, p, z


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


def independence_probabilities(pd: Pdag) -> dict:
    """Each gate's probability under the independence approximation (inputs
    treated as independent: AND = product, OR = 1 - product of complements,
    ATLEAST by the Poisson-binomial distribution of the count, XOR by its
    parity recursion, NOT/NULL the last input), from the basic events' p.
    Deterministic and exact for trees without shared inputs; used here only
    to pick a workload's top structure. -> {pdag index: p}."""
    from .pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR
    prob = {int(v): float(p) for v, p in zip(pd.var_index, pd.var_p)}
    gates = {int(g): s for s, g in enumerate(pd.gate_index)}

    def value(s):
        ps = [prob[int(a)] if a > 0 else 1.0 - prob[int(-a)] for a in pd.gate_args(s)]
        t = int(pd.gate_type[s])
        if t in (AND, NAND):
            v = float(np.prod(ps))
        elif t in (OR, NOR):
            v = 1.0 - float(np.prod([1.0 - q for q in ps]))
        elif t == XOR:
            v = 0.0
            for q in ps:
                v = v * (1.0 - q) + (1.0 - v) * q
        elif t in (NOT, NULL):
            v = ps[-1] if ps else 0.0
        else:  # ATLEAST: P(count >= k)
            dist = np.zeros(len(ps) + 1)
            dist[0] = 1.0
            for q in ps:
                dist[1:] = dist[1:] * (1.0 - q) + dist[:-1] * q
                dist[0] *= 1.0 - q
            v = float(dist[int(pd.gate_k[s]):].sum())
        return 1.0 - v if t in (NAND, NOR, NOT) else v

    # Gates in an order where every gate follows its gate inputs.
    order, state = [], {}
    for g in gates:
        stack = [(g, False)]
        while stack:
            x, done = stack.pop()
            if done:
                if state.get(x) != 2:
                    state[x] = 2
                    order.append(x)
                continue
            if state.get(x):
                continue
            state[x] = 1
            stack.append((x, True))
            for a in pd.gate_args(gates[x]):
                a = abs(int(a))
                if a in gates and not state.get(a):
                    stack.append((a, False))
    for g in order:
        prob[g] = value(gates[g])
    return prob


# The large tree's generator with the top OR over the last generator layer
# (1,379 gates) is certain: a third of those gates are near-certain, since a
# complemented rare input makes an OR likely. The other last-layer gates
# are rare; together their cones hold 24,974 of the 25,000 basic events.
AT_SCALE_LOW_P = 1e-3


def at_scale_tree(seed: int = 2, low_p: float = AT_SCALE_LOW_P) -> Pdag:
    """The large tree's family (25,000 basic events with p log-uniform in
    [1e-6, 1e-2]; 40,000 gates in 30 generator layers, fan-in 2-8, 48 % AND /
    47 % OR / 5 % ATLEAST with n up to 16, 10 % of inputs complemented;
    ``top="last_layer"``) with a top event that is neither certain nor
    negligible: the generator's top OR over the last layer becomes
    AND(A, B), with A the OR of the last-layer gates whose independence-
    approximation probability at the medians is below ``low_p`` and B the OR
    of the others (near-certain). Every generated gate stays reachable;
    the tree has 40,002 gates."""
    from .pdag import AND, OR
    full = random_layered_tree(seed, 25_000, 40_000, 30, p_range=(1e-6, 1e-2), fan_in=(2, 8),
                               mix=(0.48, 0.47, 0.05), atleast_n=(3, 16), neg_frac=0.10,
                               top="last_layer")
    prob = independence_probabilities(full)
    gates = {int(g): s for s, g in enumerate(full.gate_index)}
    top_args = [int(a) for a in full.gate_args(gates[full.root])]
    low = [a for a in top_args if prob[a] < low_p]
    high = [a for a in top_args if prob[a] >= low_p]
    if len(low) < 2 or len(high) < 2:
        raise ValueError("the last layer does not split at low_p")
    variables = list(zip(full.var_index.tolist(), full.var_p.tolist()))
    body = [(int(g), int(full.gate_type[s]), int(full.gate_k[s]),
             [int(a) for a in full.gate_args(s)])
            for s, g in enumerate(full.gate_index) if int(g) != full.root]
    a_idx, b_idx, top = full.root, full.root + 1, full.root + 2
    body += [(a_idx, OR, 0, low), (b_idx, OR, 0, high), (top, AND, 0, [a_idx, b_idx])]
    pd = Pdag.from_lists(variables, body, top)
    pd.meta = {"seed": seed, "generator": "at_scale_tree", "low_p": low_p,
               "low_gates": len(low), "high_gates": len(high)}
    return pd


def at_scale_uncertainty(seed: int = 2):
    """``at_scale_tree`` with lognormal uncertainty on every basic event:
    median = the tree's probability (log-uniform in [1e-6, 1e-2]), error
    factor uniform in [3, 10]. -> (pdag with the medians as p, LognormalModel)."""
    pd = at_scale_tree(seed)
    rng = SplitMix64(seed ^ 0x5EED)
    ef = np.array([3.0 + 7.0 * rng.uniform() for _ in range(pd.n_vars)])
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
