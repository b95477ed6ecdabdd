"""Surrogate-screened uncertainty analysis.

1. Training pairs: lognormal parameter samples of the tree, drawn on the GPU
   (``LognormalModel.draw``; their standard-normal scores z, the normalised
   log probabilities, are the network's input) and the GPU Monte Carlo
   estimate of the top event for each (its logit is the target).
2. A two-member ensemble of surrogate networks fitted to the pairs. Each
   network first maps z_i to p_i / q_i, with q_i the event's 99th percentile
   (``lognormal_input_map``): the top event is a sum of products of event
   probabilities, nearly linear in p and far from linear in ln p, and on the
   2,000-event uncertainty tree a linear fit of the logit explains 35 % of its
   variance on p against 1 % on ln p.
3. Screening: every outer sample is predicted by both members; samples whose
   members' logits differ by more than ``disagreement`` (an absolute bound in
   logit space: 0.05 is a 5 % ratio of the two predicted odds, whatever the
   probability), plus a random ``audit`` fraction, go to plain Monte Carlo.
   Routed samples take the Monte Carlo value; the audit measures the
   surrogate's error on samples it was trusted with.

The networks learn the residual of the top event's logit over a closed-form
baseline computed from the same sampled probabilities: the independence
approximation of the top event (``independence_logits``, on the GPU). On the
2,000-event tree it follows the Monte Carlo logit with correlation 0.98;
over it the two members disagree by more than 0.05 on about 1 % of
held-out samples instead of 4 %, and their 99th-percentile logit error
halves.

Plain Monte Carlo over every outer sample is the fallback and the yardstick of
the result (``plain_uncertainty``). Monte Carlo estimates use
p = (ones + 1/2) / (trials + 1) so that the logit stays finite.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

import ctypes as C

from .engine import _check, _p, _PdagView, load_library
from .pdag import Pdag
from .surrogate import Ensemble
from .ua import LognormalModel, UncertaintyEngine, statistics


def logit(p):
    p = np.asarray(p, dtype=np.float64)
    return np.log(p) - np.log1p(-p)


def sigmoid(x):
    return 1.0 / (1.0 + np.exp(-np.asarray(x, dtype=np.float64)))


def independence_logits(pd: Pdag, p_samples) -> np.ndarray:
    """Logit of the top event under the independence approximation, one per
    row of p_samples ([n, n_vars], the order of pd.var_index), computed on
    the GPU (mcs_independence_logits)."""
    p = np.ascontiguousarray(p_samples, dtype=np.float32)
    if p.ndim != 2 or p.shape[1] != pd.n_vars:
        raise ValueError("p_samples must be [n_samples, n_vars]")
    out = np.empty(p.shape[0], dtype=np.float64)
    view = _PdagView(pd)
    _check(load_library().mcs_independence_logits(C.byref(view.struct), _p(p), p.shape[0], _p(out)))
    return out


def monte_carlo(pd: Pdag, p_samples, words: int, ids, passes: int = 1):
    """Top-event estimate of each parameter sample by plain GPU Monte Carlo
    (sample k uses Philox key (key0 + ids[k], key1)). -> (p_hat, seconds)."""
    with MonteCarloRunner(pd, words) as mc:
        return mc(p_samples, ids, passes)


class MonteCarloRunner:
    """Plain GPU Monte Carlo of successive sets of parameter samples of one
    tree through one uncertainty engine: its device arenas are allocated for
    the first set and reused (``UncertaintyEngine.set_samples``), since a
    large arena allocated again after a free costs seconds. Each call gives
    the results a fresh engine would, bit for bit."""

    def __init__(self, pd: Pdag, words: int):
        self.pd, self.words, self.eng = pd, words, None

    def __call__(self, p_samples, ids, passes: int = 1):
        """-> (top-event estimate per sample, seconds)."""
        t = time.perf_counter()
        if self.eng is None:
            self.eng = UncertaintyEngine(self.pd, self.words, p_samples, sample_ids=ids)
        else:
            self.eng.set_samples(p_samples, ids)
        self.eng.run(passes)
        ones, total = self.eng.collect()
        dt = time.perf_counter() - t
        return (ones[:, 0].astype(np.float64) + 0.5) / (total + 1.0), dt

    def close(self):
        if self.eng is not None:
            self.eng.close()
            self.eng = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


@dataclass
class ScreeningResult:
    estimates: np.ndarray        # top-event probability per outer sample
    routed: np.ndarray           # bool: sent to plain Monte Carlo
    disagreed: np.ndarray        # bool: ensemble disagreement above the bound
    audited: np.ndarray          # bool: random audit
    audit_rel_error: dict        # surrogate vs Monte Carlo on the audit subset
    seconds: dict = field(default_factory=dict)

    def stats(self):
        return statistics(self.estimates)


Z99 = 2.3263478740408408  # standard-normal 99th percentile


def lognormal_input_map(model: LognormalModel):
    """(scale, shift, cap) with min(cap, exp(scale z + shift)) = p / q99 for
    p = min(1, exp(mu + sigma z)) and q99 = exp(mu + Z99 sigma)."""
    scale = model.sigma
    shift = -Z99 * model.sigma
    cap = np.exp(-(model.mu + Z99 * model.sigma))
    return scale, shift, cap


def train_surrogate(pd: Pdag, model: LognormalModel, n_train: int, words: int, seed: int,
                    hidden: int = 512, epochs: int = 30, batch: int = 256, lr: float = 1e-3,
                    id_base: int = 1 << 24, chunk: int = 16384, progress=None,
                    baseline: str | None = "independence"):
    """Monte Carlo training pairs (in runs of ``chunk`` samples, each followed
    by ``progress(done, n_train)`` if given) and a fitted two-member ensemble.
    With ``baseline="independence"`` the members fit the residual of the
    Monte Carlo logit over ``independence_logits`` of the same samples,
    clipped to the logits a run of words x 64 trials can give, +-ln(2 T + 1)
    (a run that sees no failure, or no success, reports T + 1/2 over T + 1):
    where the top event is nearly certain or nearly impossible the Monte
    Carlo targets sit at that bound and so does the clipped baseline.
    -> (ensemble, info dict with timings and the training losses)."""
    if baseline not in (None, "independence"):
        raise ValueError(f"unknown baseline {baseline!r}")
    ids = (id_base + np.arange(n_train)).astype(np.uint32)
    p_hat = np.empty(n_train)
    base = np.zeros(n_train)
    z = np.empty((n_train, pd.n_vars), dtype=np.float32)
    t_mc = t_base = 0.0
    with MonteCarloRunner(pd, words) as mc:
        for a, p, zc in model.draw_chunks(n_train, seed, chunk):
            b = a + p.shape[0]
            z[a:b] = zc
            p_hat[a:b], dt = mc(p, ids[a:b])
            t_mc += dt
            if baseline:
                t = time.perf_counter()
                base[a:b] = independence_logits(pd, p)
                t_base += time.perf_counter() - t
            if progress:
                progress(b, n_train)
    t = time.perf_counter()
    clip = float(np.log(2.0 * words * 64 + 1.0))
    ens = Ensemble(pd.n_vars, hidden=hidden, members=2, seed=seed,
                   input_map=lognormal_input_map(model), baseline=baseline, baseline_clip=clip)
    losses = ens.fit(z, logit(p_hat) - np.clip(base, -clip, clip), epochs=epochs, batch=batch,
                     lr=lr)
    t_fit = time.perf_counter() - t
    return ens, {"pairs": n_train, "mc_seconds": t_mc, "baseline_seconds": t_base,
                 "fit_seconds": t_fit, "baseline": baseline,
                 "final_losses": [float(l[-1]) for l in losses]}


def route(lg: np.ndarray, disagreement: float) -> np.ndarray:
    """Samples whose ensemble members' logits lg [members, n] differ by more
    than ``disagreement`` (absolute, in logit space)."""
    return (lg.max(axis=0) - lg.min(axis=0)) > disagreement


def _quantiles(x):
    return {"mean": float(np.mean(x)), "p05": float(np.quantile(x, 0.05)),
            "p50": float(np.median(x)), "p95": float(np.quantile(x, 0.95))}


def paired_errors(surrogate: np.ndarray, mc: np.ndarray) -> dict:
    """Surrogate predictions against Monte Carlo on the same samples: the
    relative error of the mean, 5th, 50th and 95th percentile, and the
    median absolute logit error."""
    s, m = _quantiles(surrogate), _quantiles(mc)
    out = {k: {"surrogate": s[k], "plain": m[k], "rel_error": abs(s[k] - m[k]) / m[k]}
           for k in s}
    out["samples"] = int(len(mc))
    out["median_abs_logit_error"] = float(np.median(np.abs(logit(surrogate) - logit(mc))))
    return out


def screened_uncertainty(pd: Pdag, model: LognormalModel, ens: Ensemble, n_outer: int,
                         words: int, seed: int, disagreement: float = 0.05, audit: float = 0.05,
                         id_base: int = 0, chunk: int = 16384, mc_chunk: int = 1 << 30,
                         progress=None) -> ScreeningResult:
    """Predict every outer sample (in blocks of ``chunk`` samples), route the
    disagreeing and audited ones to plain Monte Carlo (in runs of
    ``mc_chunk`` samples, each followed by ``progress(done, n_routed)``) and
    keep the Monte Carlo value for them."""
    rng = np.random.Generator(np.random.PCG64(seed ^ 0xA0D17))
    audited = rng.random(n_outer) < audit
    disagreed = np.zeros(n_outer, dtype=bool)
    est = np.empty(n_outer)
    routed_p = []
    t0 = time.perf_counter()
    t_pred = 0.0
    for a, p, z in model.draw_chunks(n_outer, seed, chunk):
        b = a + p.shape[0]
        t = time.perf_counter()
        lg = ens.predict(z)                   # [members, rows]
        if ens.baseline:
            c = ens.baseline_clip
            lg += np.clip(independence_logits(pd, p), -c, c)[None, :]
        t_pred += time.perf_counter() - t
        disagreed[a:b] = route(lg, disagreement)
        est[a:b] = sigmoid(lg.mean(axis=0))
        r = disagreed[a:b] | audited[a:b]
        routed_p.append(p[r])
    routed = disagreed | audited
    idx = np.nonzero(routed)[0]
    t_mc = 0.0
    audit_err = {}
    if idx.size:
        p_r = np.concatenate(routed_p)
        p_hat = np.empty(idx.size)
        with MonteCarloRunner(pd, words) as mc:
            for a in range(0, idx.size, mc_chunk):
                b = min(idx.size, a + mc_chunk)
                p_hat[a:b], dt = mc(p_r[a:b], (id_base + idx[a:b]).astype(np.uint32))
                t_mc += dt
                if progress:
                    progress(b, idx.size)
        surrogate_vals = est[idx].copy()
        est[idx] = p_hat
        a = audited[idx] & ~disagreed[idx]
        if a.any():
            audit_err = paired_errors(surrogate_vals[a], p_hat[a])
    return ScreeningResult(est, routed, disagreed, audited, audit_err,
                           {"predict": t_pred, "monte_carlo": t_mc,
                            "total": time.perf_counter() - t0,
                            "mc_per_sample": t_mc / idx.size if idx.size else None})


def plain_uncertainty(pd: Pdag, model: LognormalModel, n_outer: int, words: int, seed: int,
                      id_base: int = 0, chunk: int = 1 << 30, progress=None):
    """Every outer sample by plain Monte Carlo (same samples, same keys as
    screened_uncertainty with the same seed), in runs of ``chunk`` samples,
    each followed by ``progress(done, n_outer)`` if given.
    -> (estimates, seconds)."""
    p, _ = model.draw(n_outer, seed)
    ids = (id_base + np.arange(n_outer)).astype(np.uint32)
    est = np.empty(n_outer)
    seconds = 0.0
    with MonteCarloRunner(pd, words) as mc:
        for a in range(0, n_outer, chunk):
            b = min(n_outer, a + chunk)
            est[a:b], dt = mc(p[a:b], ids[a:b])
            seconds += dt
            if progress:
                progress(b, n_outer)
    return est, seconds


def compare(screened: ScreeningResult, plain: np.ndarray) -> dict:
    s, m = screened.estimates, plain
    out = {}
    for name, f in (("mean", np.mean), ("p05", lambda x: np.quantile(x, 0.05)),
                    ("p50", np.median), ("p95", lambda x: np.quantile(x, 0.95))):
        out[name] = {"screened": float(f(s)), "plain": float(f(m)),
                     "rel_error": float(abs(f(s) - f(m)) / f(m))}
    out["routed_fraction"] = float(screened.routed.mean())
    out["disagreed_fraction"] = float(screened.disagreed.mean())
    return out
