"""Batches of independent trees evaluated whole in LDS (one persistent
launch per pass): every tree's tallies must equal the oracle's for that tree
alone, bit for bit, whatever the tile width, work-item size or graph replay.
"""
import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd.engine import EngineError
from oracle import binding as orc
from tests.models import small_random


def _all_nodes(pd):
    return [int(g) for g in pd.gate_index] + [int(v) for v in pd.var_index]


def _mixed_trees():
    trees = [small_random(s, n_be=10 + 7 * s, n_gates=12 + 9 * s, n_layers=3 + s % 5)
             for s in range(1, 7)]
    trees.append(mc.small_tree(1))
    return trees


def test_batch_rejects_tree_too_large_for_lds():
    with pytest.raises(EngineError, match="too large for one LDS tile"):
        mc.Batch([mc.large_tree()], 1024)


def test_batch_rejects_bad_config():
    with pytest.raises(EngineError, match="power of two"):
        mc.Batch([mc.small_tree(1)], 1024, max_tile_words=48)


def _check_batch(trees, shapes, n_pass=1, **cfg):
    tallies = [_all_nodes(t) for t in trees]
    with mc.Batch(trees, shapes, tallies, **cfg) as b:
        b.pass_(n_pass)
        got = [b.collect(i) for i in range(len(trees))]
        widths = [b.tile_words(i) for i in range(len(trees))]
    for i, (pd, sh) in enumerate(zip(trees, shapes)):
        n_bp = sh[0] * sh[1] if isinstance(sh, tuple) else sh
        want, _ = orc.run(pd, n_bp, n_iters=n_pass, tally=tallies[i])
        g = {t.index: t.num_one_bits for t in got[i]}
        bad = [(idx, g[idx], int(w)) for idx, w in zip(tallies[i], want) if g[idx] != int(w)]
        assert not bad, (i, widths[i], bad[:5])
        assert {t.total_bits for t in got[i]} == {n_pass * n_bp * 64}
    return widths


@pytest.mark.gpu
def test_batch_matches_oracle_per_tree(gpu):
    trees = _mixed_trees()
    shapes = [1000, 4096, (3, 777), 64, 5000, 1, (16, 1024)]
    _check_batch(trees, shapes, n_pass=2)


@pytest.mark.gpu
@pytest.mark.parametrize("cfg", [dict(use_graph=0), dict(lds_bytes=8192, item_words=64),
                                 dict(max_tile_words=1), dict(max_tile_words=16, item_words=100000),
                                 dict(workgroup_lanes=64), dict(workgroup_lanes=256, lds_bytes=150 << 10)])
def test_batch_configs_give_identical_tallies(gpu, cfg):
    trees = _mixed_trees()[:5]
    widths = _check_batch(trees, [3000] * 5, n_pass=2, **cfg)
    if cfg.get("max_tile_words") == 1:
        assert set(widths) == {1}


@pytest.mark.gpu
def test_batch_uneven_random_trees(gpu):
    # Trees of the batched workload's family (gate count log-uniform, events
    # 0.6 x gates), at a 2^20-trial prefix each.
    trees = mc.batched_trees(seed=3, n_trees=6, gate_range=(50, 2000))
    _check_batch(trees, [1 << 14] * len(trees))


@pytest.mark.gpu
def test_batch_root_only_default_and_info(gpu):
    trees = _mixed_trees()[:3]
    with mc.Batch(trees, 2048) as b:
        b.pass_(3)
        info = b.info()
        assert info["iteration"] == 3 and info["num_trees"] == 3
        roots = [b.collect(i)[0] for i in range(3)]
    for pd, t in zip(trees, roots):
        want, _ = orc.run(pd, 2048, n_iters=3)
        assert t.index == pd.root and t.num_one_bits == int(want[0])
