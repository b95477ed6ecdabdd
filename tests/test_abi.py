"""The C ABI library loads, exports every entry point include/mcscram_amd.h
declares, and its host-only logic (argument checks, pass planning, CI
quantiles) behaves as the reference's. No device is touched here.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd.engine import EngineError, _PdagView, _Shape, load_library
from mcscram_amd.pdag import AND, ATLEAST, OR, Pdag

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(mcs_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = load_library()
    declared = _declared()
    assert len(declared) >= 15
    for name in declared:
        assert hasattr(lib, name), name


def test_library_is_built_for_gfx950():
    data = open(mc.library_path(), "rb").read()
    assert b"gfx950" in data


def test_plan_passes_matches_reference_rounding():
    # trials -> nearest multiple of 64, ties down; passes = ceil(t / (64 N_bp))
    # (src/mc/queue/sample_shaper.h:73-83, 139-146).
    assert mc.plan_passes(100, 1) == (128, 2)
    assert mc.plan_passes(96, 1) == (64, 1)    # tie rounds down
    assert mc.plan_passes(97, 1) == (128, 2)
    assert mc.plan_passes(1 << 20, (16, 1024)) == (1 << 20, 1)
    assert mc.plan_passes(1 << 32, 1 << 18) == (1 << 32, 256)
    assert mc.plan_passes(10, 4) == (0, 0)


def _closest_pow2_shape(trials, nodes, max_alloc, limits):
    """Closed form of the reference's shape choice: the power of two no
    larger than the per-node capacity and the work-item limits that is
    closest to the wanted bitpacks (ties to the smaller)."""
    t = trials or (1 << 26)
    r = t % 64
    rounded = t if r == 0 else (t - r if r <= 32 else t + 64 - r)
    cap_bits = max_alloc * 8 // nodes
    want = -(-min(rounded, cap_bits) // 64)
    hp2 = lambda v: 1 << (v.bit_length() - 1) if v else 0  # noqa: E731
    top = min(cap_bits // 64, hp2(limits[0]) * hp2(limits[1]))
    cands = [1 << k for k in range(0, 64) if (1 << k) <= top] or [1]
    best = min(cands, key=lambda p: (abs(p - want), p))
    passes = 0 if trials == 0 else -(-rounded // (64 * best))
    return (1, best), rounded, passes


def test_sample_shape_matches_the_reference_shaper():
    # The small workload: 2^20 trials of a 320-node tree on a large device
    # give 16,384 words in one pass, as the reference runs it.
    assert mc.sample_shape_for(1 << 20, 320, 288 << 30) == ((1, 16384), 1 << 20, 1)
    # No trial limit: 2^26 trials size the shape, no pass count.
    assert mc.sample_shape_for(0, 320, 288 << 30) == ((1, 1 << 20), 1 << 26, 0)
    # Rounding ties down; the per-node memory cap and the work-item limits bind.
    assert mc.sample_shape_for(96, 3, 1 << 30)[1:] == (64, 1)
    assert mc.sample_shape_for(1 << 32, 65_000, 288 << 30) == ((1, 1 << 19), 1 << 32, 128)
    assert mc.sample_shape_for(1 << 40, 1, 1 << 40, (8, 16))[0] == (1, 128)
    import random
    rng = random.Random(3)
    for _ in range(3000):
        trials = rng.choice([0, rng.randrange(1, 1 << 12), rng.randrange(1, 1 << 40)])
        nodes = rng.randrange(1, 200_000)
        alloc = rng.randrange(1 << 16, 1 << 39)
        limits = (rng.randrange(1, 4096), rng.randrange(1, 4096))
        assert mc.sample_shape_for(trials, nodes, alloc, limits) == \
            _closest_pow2_shape(trials, nodes, alloc, limits), (trials, nodes, alloc, limits)
    with pytest.raises(EngineError):
        mc.sample_shape_for(1 << 20, 0, 1 << 30)


def _nvidia_shape(trials, nodes, max_alloc, overhead, line):
    """Closed form of the shaper's NVIDIA rule (src/mc/queue/sample_shaper.h:
    120-134): the largest allocation in words, split over the nodes with the
    overhead, rounded down to whole cache lines; passes = ceil(bits / bits
    per pass)."""
    t = trials or (1 << 26)
    r = t % 64
    rounded = t if r == 0 else (t - r if r <= 32 else t + 64 - r)
    words = int((max_alloc / 8) // (nodes * (1 + min(max(overhead, 0.0), 1e6))))
    words -= words % (line // 8)
    passes = 0 if trials == 0 else -(-rounded // (64 * words))
    return (1, words), rounded, passes


def test_sample_shape_nvidia_rule():
    # 89 nodes on a device with a 6 GiB largest allocation, 128-byte lines and
    # the default 5 % overhead: floor(805306368 / 93.45) = 8,617,510 words,
    # rounded down to 16-word lines = 8,617,504; the trial count is ignored.
    assert mc.sample_shape_for(1 << 20, 89, 6 << 30, nvidia_gpu=True) == \
        ((1, 8617504), 1 << 20, 1)
    assert mc.sample_shape_for(0, 89, 6 << 30, nvidia_gpu=True) == ((1, 8617504), 1 << 26, 0)
    # Overhead is clamped below at 0 and the alignment follows the line size.
    assert mc.sample_shape_for(64, 1, 1 << 20, nvidia_gpu=True, overhead_ratio=-1.0,
                               cache_line_bytes=64) == ((1, 131072), 64, 1)
    assert mc.sample_shape_for(1 << 30, 3, 3000, nvidia_gpu=True, overhead_ratio=0.0,
                               cache_line_bytes=8)[0] == (1, 125)
    import random
    rng = random.Random(5)
    for _ in range(3000):
        trials = rng.choice([0, rng.randrange(1, 1 << 12), rng.randrange(1, 1 << 40)])
        nodes = rng.randrange(1, 200_000)
        alloc = rng.randrange(1 << 24, 1 << 39)
        overhead = rng.choice([0.0, 0.05, rng.random()])
        line = rng.choice([32, 64, 128, 256])
        want = _nvidia_shape(trials, nodes, alloc, overhead, line)
        if want[0][1] == 0:
            with pytest.raises(EngineError):
                mc.sample_shape_for(trials, nodes, alloc, nvidia_gpu=True,
                                    overhead_ratio=overhead, cache_line_bytes=line)
            continue
        assert mc.sample_shape_for(trials, nodes, alloc, nvidia_gpu=True, overhead_ratio=overhead,
                                   cache_line_bytes=line) == want, (trials, nodes, alloc, overhead)
    # No whole line per node: the reference would divide by zero; here an error.
    with pytest.raises(EngineError):
        mc.sample_shape_for(1 << 20, 1000, 1000, nvidia_gpu=True)
    with pytest.raises(EngineError):
        mc.sample_shape_for(1 << 20, 10, 1 << 30, nvidia_gpu=True, cache_line_bytes=4)


def _estimates(ones, total):
    L = load_library()
    from mcscram_amd.engine import _Tally
    t = _Tally()
    t.num_one_bits, t.total_bits = ones, total
    L.mcs_tally_point_estimates(C.byref(t))
    return t.mean, t.std_err, list(t.ci)


def test_point_estimates_follow_the_reference_closed_form():
    """stats::populate_point_estimates (src/mc/stats/ci_utils.h:167-201):
    mean = ones / total, SE = sqrt(mean (1 - mean) / total), ci =
    {lo95, hi95, lo99, hi99} = mean -/+ z SE clamped to [0, 1], with z the
    two-sided 95 % and 99 % normal quantiles; all zero at total 0."""
    from scipy.stats import norm
    z95, z99 = norm.ppf(0.975), norm.ppf(0.995)
    # No samples yet: everything zero.
    assert _estimates(0, 0) == (0.0, 0.0, [0.0] * 4)
    # Edge counts: no ones and all ones give a zero-width interval at 0 and 1.
    assert _estimates(0, 1 << 20) == (0.0, 0.0, [0.0] * 4)
    assert _estimates(1 << 20, 1 << 20) == (1.0, 0.0, [1.0] * 4)
    # Both clamps: a rare count's lower bounds clamp at 0, a near-certain
    # count's upper bounds at 1; the other side stays open.
    m, se, ci = _estimates(1, 64)
    assert ci[0] == 0.0 and ci[2] == 0.0 and 0 < ci[1] < ci[3] < 1
    m, se, ci = _estimates(63, 64)
    assert ci[1] == 1.0 and ci[3] == 1.0 and 0 < ci[2] < ci[0] < 1
    # Order and widths against the closed form with scipy's quantiles.
    rng = np.random.default_rng(11)
    for _ in range(2000):
        total = int(rng.integers(1, 1 << 40))
        ones = int(rng.integers(0, total + 1))
        m, se, ci = _estimates(ones, total)
        mu = ones / total
        want_se = np.sqrt(mu * (1 - mu) / total)
        assert m == mu
        assert se == pytest.approx(want_se, rel=1e-15, abs=0)
        want = [max(0.0, mu - z95 * want_se), min(1.0, mu + z95 * want_se),
                max(0.0, mu - z99 * want_se), min(1.0, mu + z99 * want_se)]
        assert ci == pytest.approx(want, rel=1e-14, abs=1e-300)
        assert ci[2] <= ci[0] <= m <= ci[1] <= ci[3]


def test_normal_quantile_matches_scipy():
    from scipy.stats import norm
    # Acklam's approximation has relative error below 1.15e-9.
    for c in (0.5, 0.9, 0.95, 0.99, 0.999, 0.999999):
        z = norm.ppf(0.5 + c / 2)
        assert abs(mc.normal_quantile_two_sided(c) - z) <= 1.2e-9 * z


def _bad(pd, shape=(1, 64), tally=None, code=None):
    L = load_library()
    view = _PdagView(pd)
    sh = _Shape(*shape)
    idx = np.array(tally or [pd.root], dtype=np.int32)
    h = C.c_void_p()
    st = L.mcs_engine_create(C.byref(view.struct), C.byref(sh), idx.ctypes.data_as(C.c_void_p),
                             idx.shape[0], C.byref(h))
    assert st != 0
    if code is not None:
        assert st == code, L.mcs_last_error()
    assert h.value is None
    assert L.mcs_last_error()


def test_create_rejects_bad_input_before_touching_a_device():
    base = [(2, 0.1), (3, 0.2)]
    _bad(Pdag.from_lists([(2, 1.5), (3, 0.2)], [(4, OR, 0, [2, 3])], 4), code=2)   # p > 1
    _bad(Pdag.from_lists(base, [(4, OR, 0, [2, 9])], 4), code=2)                  # unknown arg
    _bad(Pdag.from_lists(base, [(4, OR, 0, [2, 5]), (5, AND, 0, [3, 4])], 4), code=2)  # cycle
    _bad(Pdag.from_lists(base, [(4, OR, 0, [2, 3])], 2), code=2)                  # root is a variable
    _bad(Pdag.from_lists(base + [(4, 0.3)], [(4, OR, 0, [2, 3])], 4), code=2)     # duplicate index
    _bad(Pdag.from_lists(base, [(4, OR, 0, [2, 3])], 4), shape=(0, 64), code=1)   # empty shape
    _bad(Pdag.from_lists(base, [(4, OR, 0, [2, 3])], 4), tally=[99], code=2)      # unknown tally
    _bad(Pdag.from_lists(base, [(4, 9, 0, [2, 3])], 4), code=2)                   # bad connective


def test_null_handles():
    L = load_library()
    assert L.mcs_engine_pass(None, 1) == 1
    assert L.mcs_engine_wait(None) == 1
    assert L.mcs_engine_collect(None, None, 0) == 1
    assert L.mcs_engine_destroy(None) == 0
    assert L.mcs_probability_analysis(None, None, None, None, 0, None, 0, None, None) == 1
    assert L.mcs_ua_run(None, 1) == 1
    assert L.mcs_ua_set_samples(None, 1, None, None) == 1
    assert L.mcs_lognormal_samples(None, None, 3, 1, 0, 4, None, None) == 0  # no output: no-op


def test_settings_defaults_are_the_reference_defaults():
    s = mc.Settings()
    assert (s.num_trials, s.burn_in_trials, s.ci_confidence, s.ci_rel_margin) == \
        (0, 1 << 20, 0.99, 0.001)
    L = load_library()
    from mcscram_amd.engine import _Settings
    cs = _Settings()
    L.mcs_settings_default(C.byref(cs))
    assert (cs.num_trials, cs.burn_in_trials, cs.ci_confidence, cs.ci_rel_margin) == \
        (0, 1 << 20, 0.99, 0.001)
    assert (cs.key0, cs.key1) == (382307844, 293830103)
    # Bayes is the reference's default stopping rule (src/settings.h:426).
    assert cs.ci_policy == 0 and s.ci_policy == "bayes"
    # Early stop is on by default (src/settings.h:453).
    assert cs.early_stop == 1 and s.early_stop is True


def test_generators_are_deterministic_and_sized():
    a, b = mc.small_tree(1), mc.small_tree(1)
    assert np.array_equal(a.args, b.args) and np.array_equal(a.var_p, b.var_p)
    assert a.n_vars == 200 and a.n_gates == 120
    assert np.all((a.var_p >= 1e-4) & (a.var_p <= 1e-1))
    a.validate()
    t = mc.pdag.Pdag  # noqa: F841
    counts = np.bincount(a.gate_type, minlength=3)
    assert counts[ATLEAST] > 0 and counts[AND] > 0 and counts[OR] > 0
