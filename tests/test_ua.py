"""Uncertainty analysis: the lognormal parameter model and the statistics
against their closed forms, and the GPU engine's per-sample Monte Carlo runs
bit-exact against the oracle run with the same probabilities and key.
"""
import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd import ua
from oracle import binding as orc
from tests.models import small_random

KEY = (382307844, 293830103)


def test_lognormal_model_median_and_error_factor():
    # ln p ~ N(ln m, (ln EF / z95)^2): the median is m and the 95th
    # percentile is m * EF (src/expression/random_deviate.cc:91-108).
    m = ua.LognormalModel(np.array([1e-3, 1e-4]), np.array([3.0, 10.0]))
    p, z = m.sample(200_000, seed=1)
    med = np.median(p, axis=0)
    q95 = np.quantile(p, 0.95, axis=0)
    assert np.allclose(med, [1e-3, 1e-4], rtol=0.02)
    assert np.allclose(q95 / med, [3.0, 10.0], rtol=0.03)
    assert np.allclose(np.log(p), m.mu + m.sigma * z)


def test_sampled_probabilities_are_clamped():
    m = ua.LognormalModel(np.array([0.5]), np.array([10.0]))
    p, _ = m.sample(10_000, seed=2)
    assert p.max() == 1.0 and p.min() > 0.0


def test_statistics_formulas():
    # mean, sigma with n/(n-1), EF = exp(1.96 sigma), CI of the mean
    # (src/uncertainty_analysis.cc:145-150).
    x = np.array([0.1, 0.2, 0.4, 0.3])
    s = ua.statistics(x, n_quantiles=4)
    assert s["mean"] == pytest.approx(0.25)
    sd = np.std(x, ddof=1)
    assert s["sigma"] == pytest.approx(sd)
    assert s["error_factor"] == pytest.approx(np.exp(1.96 * sd))
    assert s["ci95"][0] == pytest.approx(0.25 - 1.96 * sd / 2)
    assert list(s["quantiles"]) == [0.25, 0.5, 0.75, 1.0]
    assert s["quantiles"][1.0] == pytest.approx(0.4)


def test_uncertainty_tree_shape():
    pd, m = ua.uncertainty_tree()
    assert pd.n_vars == 2000 and pd.n_gates == 3000
    assert np.all((m.median >= 1e-5) & (m.median <= 1e-2))
    assert np.all((m.error_factor >= 3) & (m.error_factor <= 10))


def _oracle_sample(pd, p, n_bp, key, n_iters, tally):
    return orc.run(pd.with_probabilities(p), n_bp, n_iters=n_iters, tally=tally, key=key)[0]


def _closure(pd, top):
    gates = {int(g): i for i, g in enumerate(pd.gate_index)}
    seen, stack = {top}, [top]
    while stack:
        for a in pd.gate_args(gates[stack.pop()]):
            a = abs(int(a))
            if a not in seen:
                seen.add(a)
                if a in gates:
                    stack.append(a)
    return seen


@pytest.mark.parametrize("seed", range(6))
def test_subgraph_is_the_gates_own_pdag(seed):
    pd = mc.random_layered_tree(seed, 40, 50, 5, all_connectives=True, neg_frac=0.2)
    for top in pd.gate_index[::7]:
        top = int(top)
        sub = pd.subgraph(top)
        sub.validate()
        assert sub.root == top
        nodes = _closure(pd, top)
        assert set(sub.var_index.tolist()) | set(sub.gate_index.tolist()) == nodes
        # Order, probabilities and argument lists are kept.
        keep_v = np.isin(pd.var_index, sub.var_index)
        assert np.array_equal(sub.var_index, pd.var_index[keep_v])
        assert np.array_equal(sub.var_p, pd.var_p[keep_v])
        slot = {int(g): i for i, g in enumerate(pd.gate_index)}
        for i, g in enumerate(sub.gate_index):
            assert np.array_equal(sub.gate_args(i), pd.gate_args(slot[int(g)]))
            assert sub.gate_type[i] == pd.gate_type[slot[int(g)]]
            assert sub.gate_k[i] == pd.gate_k[slot[int(g)]]
        # The same event: the exact probability of the gate is unchanged.
        whole = mc.Pdag(pd.var_index, pd.var_p, pd.gate_index, pd.gate_type, pd.gate_k,
                        pd.arg_off, pd.args, top)
        p_whole, _ = orc.bdd_probability(whole)
        p_sub, _ = orc.bdd_probability(sub)
        assert p_whole == pytest.approx(p_sub, rel=1e-12, abs=1e-300)


def test_subgraph_rejects_a_basic_event():
    pd = mc.small_tree(1)
    with pytest.raises(ValueError):
        pd.subgraph(int(pd.var_index[0]))


@pytest.mark.gpu
def test_ua_samples_bit_exact_against_oracle(gpu):
    pd = small_random(31, n_be=40, n_gates=60, n_layers=6)
    rng = np.random.default_rng(3)
    K = 7
    P = np.clip(pd.var_p[None, :] * np.exp(rng.standard_normal((K, pd.n_vars))), 0, 1)
    P[2, 5] = 1.0
    P[3, :] = 0.0
    tally = [pd.root, int(pd.gate_index[3]), int(pd.var_index[0])]
    ids = np.array([10, 11, 12, 500, 7, 8, 9], dtype=np.uint32)
    with ua.UncertaintyEngine(pd, 1500, P, sample_ids=ids, tally=tally) as eng:
        eng.run(1)
        eng.run(1)
        ones, total = eng.collect()
    assert total == 2 * 1500 * 64
    for k in range(K):
        want = _oracle_sample(pd, P[k], 1500, ((KEY[0] + int(ids[k])) & 0xFFFFFFFF, KEY[1]), 2, tally)
        assert [int(v) for v in ones[k]] == [int(v) for v in want], k


@pytest.mark.gpu
@pytest.mark.parametrize("tree", ["random", "all_connectives"])
def test_ua_samples_per_launch_change_no_bit(gpu, tree):
    # Eleven samples in groups of 1, 3, 4, 8 and 11 per launch, on one or
    # two pipelines (groups of uneven size at each pipeline's end), with
    # tasks of 512, 1024 and 2048 words (1,500 words per pass: the last
    # chunk is partial), over two runs: every sample's tallies equal the
    # oracle run with its own key.
    if tree == "random":
        pd = small_random(33, n_be=40, n_gates=60, n_layers=6)
    else:
        pd = mc.random_layered_tree(34, 40, 70, 6, all_connectives=True, neg_frac=0.2)
    rng = np.random.default_rng(5)
    K = 11
    P = np.clip(pd.var_p[None, :] * np.exp(rng.standard_normal((K, pd.n_vars))), 0, 1)
    P[4, :] = 1.0
    P[7, :] = 0.0
    tally = [pd.root] + [int(g) for g in pd.gate_index[::9]] + [int(pd.var_index[1])]
    ids = (1000 + 3 * np.arange(K)).astype(np.uint32)
    want = [_oracle_sample(pd, P[k], 1500, ((KEY[0] + int(ids[k])) & 0xFFFFFFFF, KEY[1]), 3,
                           tally) for k in range(K)]
    for g, pipes, cw in [(1, 1, 512), (1, 2, 1024), (3, 1, 2048), (4, 2, 0), (8, 2, 512),
                         (11, 1, 1024)]:
        with ua.UncertaintyEngine(pd, 1500, P, sample_ids=ids, tally=tally,
                                  samples_per_launch=g, pipelines=pipes, chunk_words=cw) as eng:
            info = eng.info
            assert (info["samples_per_launch"], info["pipelines"], info["chunk_words"]) == \
                (min(g, -(-K // pipes)), pipes, cw or 1024)
            eng.run(1)
            eng.run(2)
            ones, total = eng.collect()
        assert total == 3 * 1500 * 64
        for k in range(K):
            assert [int(v) for v in ones[k]] == [int(v) for v in want[k]], (g, pipes, k)


@pytest.mark.gpu
def test_ua_set_samples_equals_a_fresh_engine(gpu):
    # One engine loads three sets of samples in turn (7, then 11 on other
    # ids, then 3), on two pipelines with several samples per launch, and
    # runs each for two passes: every sample's tallies equal the oracle run
    # with its own key, as a fresh engine's do, and the trial count restarts.
    pd = small_random(35, n_be=40, n_gates=60, n_layers=6)
    rng = np.random.default_rng(9)
    tally = [pd.root, int(pd.gate_index[5]), int(pd.var_index[2])]
    sets = []
    for K, base in ((7, 10), (11, 900), (3, 40)):
        P = np.clip(pd.var_p[None, :] * np.exp(rng.standard_normal((K, pd.n_vars))), 0, 1)
        sets.append((P, (base + 5 * np.arange(K)).astype(np.uint32)))
    P0, ids0 = sets[0]
    with ua.UncertaintyEngine(pd, 1500, P0, sample_ids=ids0, tally=tally, samples_per_launch=2,
                              pipelines=2) as eng:
        for P, ids in sets:
            if P is not P0:
                eng.set_samples(P, ids)
            eng.run(1)
            eng.run(1)
            ones, total = eng.collect()
            assert total == 2 * 1500 * 64
            assert ones.shape == (P.shape[0], len(tally))
            for k in range(P.shape[0]):
                want = _oracle_sample(pd, P[k], 1500, ((KEY[0] + int(ids[k])) & 0xFFFFFFFF, KEY[1]),
                                      2, tally)
                assert [int(v) for v in ones[k]] == [int(v) for v in want], (P.shape[0], k)


@pytest.mark.gpu
def test_ua_common_random_numbers(gpu):
    pd = small_random(32, n_be=30, n_gates=40, n_layers=5)
    P = np.stack([pd.var_p, pd.var_p, np.clip(pd.var_p * 2, 0, 1)])
    with ua.UncertaintyEngine(pd, 777, P, common_random_numbers=True) as eng:
        eng.run(1)
        ones, _ = eng.collect()
    assert ones[0, 0] == ones[1, 0]  # same stream, same probabilities
    for k in range(3):
        assert int(ones[k, 0]) == int(_oracle_sample(pd, P[k], 777, KEY, 1, [pd.root])[0])


@pytest.mark.gpu
def test_ua_single_sample_equals_engine(gpu):
    # One sample with the reference key and the tree's own probabilities is
    # the plain engine run.
    pd = mc.small_tree(1)
    with ua.UncertaintyEngine(pd, 4096, pd.var_p[None, :], key=(KEY[0], KEY[1]),
                              common_random_numbers=True) as eng:
        eng.run(2)
        ones, _ = eng.collect()
    with mc.LayerManager(pd, 4096) as lm:
        lm.pass_(2)
        assert int(ones[0, 0]) == lm.collect_tallies()[0].num_one_bits


@pytest.mark.gpu
def test_surrogate_screening_gains_over_plain_monte_carlo(gpu):
    # A small tree with lognormal uncertainty: the screened analysis (two-
    # member ensemble, disagreement routing, 5 % audit) must match plain
    # Monte Carlo over the same outer samples within 3 % on the mean and the
    # 95th percentile while sending only a small fraction to Monte Carlo.
    from mcscram_amd import screening as scr
    pd = small_random(41, n_be=40, n_gates=60, n_layers=6)
    rng = np.random.default_rng(8)
    model = ua.LognormalModel(np.clip(pd.var_p * 0.2, 1e-4, 0.2), 3.0 + 7.0 * rng.random(pd.n_vars))
    words = 1 << 12
    ens, info = scr.train_surrogate(pd, model, 8192, words, seed=3, hidden=128, epochs=40,
                                    batch=256, lr=2e-3)
    res = scr.screened_uncertainty(pd, model, ens, 4096, words, seed=99)
    plain, t_plain = scr.plain_uncertainty(pd, model, 4096, words, seed=99)
    cmp = scr.compare(res, plain)
    assert cmp["mean"]["rel_error"] < 0.03, cmp
    assert cmp["p95"]["rel_error"] < 0.03, cmp
    assert cmp["routed_fraction"] < 0.3, cmp
    # Routed samples carry exactly the plain Monte Carlo value (same keys).
    assert np.array_equal(res.estimates[res.routed], plain[res.routed])


def test_chunked_sampling_reproduces_one_draw():
    m = ua.LognormalModel(np.array([1e-3, 2e-2, 5e-5]), np.array([3.0, 10.0, 5.0]))
    p, z = m.sample(1000, seed=4)
    parts = list(m.sample_chunks(1000, seed=4, chunk=333))
    assert [a for a, _, _ in parts] == [0, 333, 666, 999]
    assert np.array_equal(np.concatenate([q for _, q, _ in parts]), p)
    assert np.array_equal(np.concatenate([w for _, _, w in parts]), z)


def test_disagreement_bound_is_absolute_in_logit_space():
    from mcscram_amd import screening as scr
    # Members 0.06 apart are routed at any magnitude (a rare event's logit
    # near -4.6 as much as an even one near 0); 0.04 apart never are.
    for centre in (-9.0, -4.6, 0.0, 2.0):
        lg = np.array([[centre, centre, centre], [centre + 0.06, centre - 0.04, centre + 0.03]])
        assert scr.route(lg, 0.05).tolist() == [True, False, False]


def test_paired_errors_closed_form():
    from mcscram_amd import screening as scr
    mc_vals = np.linspace(0.01, 0.02, 101)
    e = scr.paired_errors(mc_vals * 1.01, mc_vals)
    for k in ("mean", "p05", "p50", "p95"):
        assert e[k]["rel_error"] == pytest.approx(0.01, rel=1e-9)
    assert e["samples"] == 101
    assert e["median_abs_logit_error"] == pytest.approx(
        np.median(np.abs(scr.logit(mc_vals * 1.01) - scr.logit(mc_vals))))


def test_independence_probabilities_exact_on_a_tree():
    # No shared inputs: the independence approximation is exact, so it
    # matches exhaustive enumeration.
    from mcscram_amd.pdag import AND, ATLEAST, NAND, NOR, OR, XOR, Pdag
    from tests.exact import exact_probabilities
    vs = [(2 + i, 0.05 + 0.07 * i) for i in range(12)]
    g = [(20, AND, 0, [2, -3]), (21, OR, 0, [4, 5, -6]), (22, ATLEAST, 2, [7, 8, 9]),
         (23, XOR, 0, [10, -11]), (24, NAND, 0, [20, 13]), (25, NOR, 0, [21, 12]),
         (26, OR, 0, [22, 23, 24, 25])]
    pd = Pdag.from_lists(vs, g, 26)
    got = ua.independence_probabilities(pd)
    want = exact_probabilities(pd)
    for idx in (20, 21, 22, 23, 24, 25, 26):
        assert got[idx] == pytest.approx(want[idx], rel=1e-12), idx


def test_at_scale_tree_shape():
    pd, model = ua.at_scale_uncertainty()
    assert (pd.n_vars, pd.n_gates) == (25_000, 40_002)
    pd.validate()
    gi = {int(g): s for s, g in enumerate(pd.gate_index)}
    root = gi[pd.root]
    assert pd.gate_type[root] == mc.pdag.AND and len(pd.gate_args(root)) == 2
    assert pd.meta["low_gates"] + pd.meta["high_gates"] == 1379
    # Every generated gate and basic event is reachable from the top.
    sub = pd.subgraph(pd.root)
    assert (sub.n_vars, sub.n_gates) == (pd.n_vars, pd.n_gates)
    prob = ua.independence_probabilities(pd)
    assert 5e-3 < prob[pd.root] < 5e-2
    assert np.all((model.error_factor >= 3.0) & (model.error_factor < 10.0))
