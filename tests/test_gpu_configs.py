"""The benchmark workloads at their specified sizes on the GPU.

* Batched trees: all 256 trees of the batched workload (gate count
  log-uniform in [50, 5000], events 0.6 x gates), with the autotuned batch
  configuration: every node of every tree bit-exact against the oracle at
  1,024 words per pass, then one full pass of 2^20 words (2^26 trials per
  tree) checked by properties that hold at any size.
* Uncertainty tree: the 2,000-event / 3,000-gate tree with its lognormal
  model: parameter samples bit-exact against the oracle run with each
  sample's probabilities and key, and the full-width 2000-512-512-512-1
  surrogate's loss and gradient against the FP32 oracle with BF16 rounding.
* At-scale uncertainty tree: the large tree's family (25,000 events, 40,002
  gates) with lognormal uncertainty on every event: parameter samples
  bit-exact against the oracle, the top event at the medians neither
  certain nor negligible, and the 25000-512-512-512-1 surrogate's loss and
  gradient against the FP32 oracle.
"""
import json
import os

import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd import ua
from oracle import binding as orc
from oracle import surrogate_ref as ref

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
KEY = (382307844, 293830103)


def _tuned(workload):
    with open(os.path.join(ROOT, "mcscram_amd", "tuned.json")) as f:
        return dict(json.load(f)[workload]["config"])


def _all_nodes(pd):
    return [int(g) for g in pd.gate_index] + [int(v) for v in pd.var_index]


def test_batched_workload_every_tree_bit_exact(gpu):
    trees = mc.batched_trees(seed=3, n_trees=256)
    tallies = [_all_nodes(t) for t in trees]
    words = 1024
    with mc.Batch(trees, words, tallies, **_tuned("batched")) as b:
        info = b.info()
        b.pass_(1)
        got = [[t.num_one_bits for t in b.collect(i)] for i in range(len(trees))]
    assert info["num_trees"] == 256
    assert info["num_fused_trees"] > 0 and info["num_layered_trees"] > 0
    for i, pd in enumerate(trees):
        want, _ = orc.run(pd, words, tally=tallies[i])
        assert got[i] == [int(w) for w in want], (i, pd.n_gates)


def _check_properties(pd, ones, total):
    """A basic event's count is Binomial(total, p); with c_i the count of
    trials on which input i is true, AND <= min c_i, max c_i <= OR <=
    min(sum c_i, total), k * ATLEAST <= sum c_i (fewer than 256 inputs)."""
    n = float(total)
    for v, p in zip(pd.var_index, pd.var_p):
        mu, sd = n * p, np.sqrt(n * p * (1 - p))
        assert abs(ones[int(v)] - mu) <= 5.5 * sd + 1, (int(v), ones[int(v)], mu)
    for s, g in enumerate(pd.gate_index):
        c = [ones[int(x)] if x > 0 else total - ones[int(-x)] for x in pd.gate_args(s)]
        t, got = pd.gate_type[s], ones[int(g)]
        assert 0 <= got <= total
        if t == mc.pdag.AND:
            assert got <= min(c), int(g)
        elif t == mc.pdag.OR:
            assert max(c) <= got <= min(sum(c), total), int(g)
        elif t == mc.pdag.ATLEAST:
            assert int(pd.gate_k[s]) * got <= sum(c), int(g)


def test_batched_workload_full_pass_properties(gpu):
    trees = mc.batched_trees(seed=3, n_trees=256)
    tallies = [_all_nodes(t) for t in trees]
    words = 1 << 20
    with mc.Batch(trees, words, tallies, **_tuned("batched")) as b:
        b.pass_(1)
        got = [b.collect(i) for i in range(len(trees))]
    for pd, tl in zip(trees, got):
        ones = {t.index: t.num_one_bits for t in tl}
        assert {t.total_bits for t in tl} == {words * 64}
        _check_properties(pd, ones, words * 64)


def test_uncertainty_tree_samples_bit_exact(gpu):
    pd, model = ua.uncertainty_tree()
    assert (pd.n_vars, pd.n_gates) == (2000, 3000)
    p, _ = model.sample(6, seed=11)
    ids = np.array([0, 1, 2, 1000, 65535, 1 << 24], dtype=np.uint32)
    tally = [pd.root] + [int(g) for g in pd.gate_index[::97]] + [int(pd.var_index[5])]
    words = 1024
    with ua.UncertaintyEngine(pd, words, p, sample_ids=ids, tally=tally) as eng:
        eng.run(2)
        ones, total = eng.collect()
    assert total == 2 * words * 64
    for k in range(len(ids)):
        key = ((KEY[0] + int(ids[k])) & 0xFFFFFFFF, KEY[1])
        want, _ = orc.run(pd.with_probabilities(p[k]), words, n_iters=2, tally=tally, key=key)
        assert [int(v) for v in ones[k]] == [int(v) for v in want], k


def test_at_scale_tree_samples_bit_exact(gpu):
    pd, model = ua.at_scale_uncertainty()
    assert (pd.n_vars, pd.n_gates) == (25_000, 40_002)
    p, _ = model.sample(3, seed=13)
    ids = np.array([5, 99_999, 1 << 24], dtype=np.uint32)
    tally = [pd.root] + [int(g) for g in pd.gate_index[::997]] + [int(pd.var_index[7])]
    words = 512
    with ua.UncertaintyEngine(pd, words, p, sample_ids=ids, tally=tally) as eng:
        eng.run(1)
        ones, total = eng.collect()
    assert total == words * 64
    for k in range(len(ids)):
        key = ((KEY[0] + int(ids[k])) & 0xFFFFFFFF, KEY[1])
        want, _ = orc.run(pd.with_probabilities(p[k]), words, tally=tally, key=key)
        assert [int(v) for v in ones[k]] == [int(v) for v in want], k
    # The top event at the medians: 2^24 trials put it well inside (1e-3, 0.5).
    with mc.LayerManager(pd, 1 << 18) as eng:
        eng.pass_(1)
        root = next(t for t in eng.collect_tallies() if t.index == pd.root)
    assert 1e-3 < root.mean < 0.5, root.mean


@pytest.mark.parametrize("tree", ["uncertainty", "at_scale"])
def test_full_width_surrogate_gradient_against_oracle(gpu, tree):
    """n_in-512-512-512-1 (n_in = 2,000 or 25,000 basic events) with the
    lognormal input map of its tree: loss and per-tensor gradient of one
    256-row batch against the FP32 oracle with the same BF16 roundings
    (within 2 % norm-wise: accumulation order differs, so a BF16 rounding of
    a backpropagated value can differ by one unit)."""
    from mcscram_amd.screening import lognormal_input_map
    from mcscram_amd.surrogate import Mlp
    pd, model = ua.uncertainty_tree() if tree == "uncertainty" else ua.at_scale_uncertainty()
    n_in, h = pd.n_vars, 512
    net = Mlp(n_in, h, seed=7)
    scale, shift, cap = lognormal_input_map(model)
    net.set_input_map(scale, shift, cap)
    params = net.params()
    _, z = model.sample(256, seed=5)
    x = z.astype(np.float32)
    y = np.random.default_rng(2).standard_normal(256).astype(np.float32)
    loss, g = net.loss_grad(x, y)
    u = ref.apply_input_map(x, scale.astype(np.float32), shift.astype(np.float32),
                            cap.astype(np.float32))
    rloss, rg = ref.loss_and_grad(params, u, y, n_in, h, bf16_mode=True)
    assert abs(loss - rloss) <= 1e-3 * max(1.0, rloss), (loss, rloss)
    sizes = [n_in * h, h, h * h, h, h * h, h, h, 1]
    o = 0
    for name, n in zip(("W1", "b1", "W2", "b2", "W3", "b3", "w4", "b4"), sizes):
        d = np.linalg.norm(g[o:o + n] - rg[o:o + n])
        assert d <= 2e-2 * np.linalg.norm(rg[o:o + n]) + 1e-7, (name, d, np.linalg.norm(rg[o:o + n]))
        o += n
