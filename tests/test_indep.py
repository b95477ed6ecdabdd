"""Independence approximation of the top event: the CPU oracle against exact
enumeration on trees without shared events (where the approximation is
exact), and the GPU kernel (mcs_independence_logits) against the oracle on
trees with shared events, every connective and complemented inputs.
"""
import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd import ua
from mcscram_amd.pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag
from oracle import indep_ref
from tests.exact import exact_probabilities


def _tree_without_sharing(seed: int) -> Pdag:
    """Random tree in which every event and gate has one consumer: gates of
    every connective, complemented inputs, multi-input NOT/NULL, ATLEAST
    with k from 0 to n + 1."""
    rng = np.random.default_rng(seed)
    pool = list(range(1, 15))  # 14 basic events
    variables = [(v, float(rng.uniform(0.02, 0.9))) for v in pool]
    rng.shuffle(pool)
    gates, nxt = [], 100
    while len(pool) > 1:
        n = int(min(len(pool), rng.integers(1, 5)))
        args = [pool.pop() for _ in range(n)]
        args = [a if rng.random() < 0.7 else -a for a in args]
        t = int(rng.choice([AND, OR, ATLEAST, XOR, NOT, NAND, NOR, NULL]))
        k = int(rng.integers(0, n + 2)) if t == ATLEAST else 0
        gates.append((nxt, t, k, args))
        pool.insert(0, nxt)
        nxt += 1
    root = pool[0]
    if root < 100:  # a single event left: wrap it
        gates.append((nxt, OR, 0, [root]))
        root = nxt
    return Pdag.from_lists(variables, gates, root)


@pytest.mark.parametrize("seed", range(12))
def test_oracle_exact_without_shared_events(seed):
    pd = _tree_without_sharing(seed)
    pd.validate()
    want = exact_probabilities(pd)[pd.root]
    got = indep_ref.independence_top(pd, pd.var_p[None, :])[0]
    assert got == pytest.approx(want, rel=1e-12, abs=1e-15)


def test_oracle_equals_the_scalar_restatement():
    # ua.independence_probabilities (one probability vector, per-gate scalar
    # arithmetic) and the oracle (numpy over samples) on the uncertainty tree.
    pd, _ = ua.uncertainty_tree()
    want = ua.independence_probabilities(pd)[pd.root]
    got = indep_ref.independence_top(pd, pd.var_p[None, :])[0]
    assert got == pytest.approx(want, rel=1e-12)


def _samples(pd, n, seed):
    rng = np.random.default_rng(seed)
    P = np.clip(pd.var_p[None, :] * np.exp(1.5 * rng.standard_normal((n, pd.n_vars))), 0, 1)
    P[0, :] = 0.0
    P[1, :] = 1.0
    P[2, ::3] = 1.0
    return P.astype(np.float32)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", range(4))
def test_gpu_matches_oracle_all_connectives(gpu, seed):
    from mcscram_amd.screening import independence_logits
    pd = mc.random_layered_tree(50 + seed, 60, 120, 7, all_connectives=True, neg_frac=0.2)
    P = _samples(pd, 700, seed)
    got = independence_logits(pd, P)
    want = indep_ref.independence_logits(pd, P.astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
def test_gpu_matches_oracle_on_the_trees_without_sharing(gpu):
    from mcscram_amd.screening import independence_logits
    for seed in range(12):
        pd = _tree_without_sharing(seed)
        P = _samples(pd, 300, seed)
        got = independence_logits(pd, P)
        want = indep_ref.independence_logits(pd, P.astype(np.float64))
        np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
def test_gpu_matches_oracle_on_the_at_scale_tree(gpu):
    from mcscram_amd.screening import independence_logits
    pd, model = ua.at_scale_uncertainty()
    p, _ = model.sample(64, seed=9)
    P = p.astype(np.float32)
    got = independence_logits(pd, P)
    want = indep_ref.independence_logits(pd, P.astype(np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)


@pytest.mark.gpu
def test_gpu_rejects_an_atleast_beyond_its_count_states(gpu):
    from mcscram_amd.engine import EngineError
    from mcscram_amd.screening import independence_logits
    variables = [(i + 1, 0.1) for i in range(40)]
    pd = Pdag.from_lists(variables, [(100, ATLEAST, 20, list(range(1, 41)))], 100)
    with pytest.raises(EngineError):
        independence_logits(pd, np.full((4, 40), 0.1, dtype=np.float32))
    # 40 inputs with k = 3 need four count states: accepted.
    pd = Pdag.from_lists(variables, [(100, ATLEAST, 3, list(range(1, 41)))], 100)
    got = independence_logits(pd, np.full((4, 40), 0.1, dtype=np.float32))
    want = indep_ref.independence_logits(pd, np.full((4, 40), np.float32(0.1), dtype=np.float64))
    np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-9)
