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


def test_batch_rejects_bad_config():
    with pytest.raises(EngineError, match="power of two"):
        mc.Batch([mc.small_tree(1)], 1024, max_tile_words=48)
    with pytest.raises(EngineError, match="layered_streams"):
        mc.Batch([mc.small_tree(1)], 1024, layered_streams=9)
    with pytest.raises(TypeError, match="unknown batch config"):
        mc.Batch([mc.small_tree(1)], 1024, layered_stream=2)


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
                                 dict(max_tile_words=1, min_fused_tile_words=1),
                                 dict(max_tile_words=16, item_words=100000),
                                 dict(workgroup_lanes=64), dict(workgroup_lanes=256, lds_bytes=150 << 10),
                                 dict(min_fused_tile_words=128), dict(lds_bytes=4096),
                                 dict(layered_streams=1),
                                 dict(layered_streams=3, min_fused_tile_words=128),
                                 dict(layered_streams=8, min_fused_tile_words=128, use_graph=0)])
def test_batch_configs_give_identical_tallies(gpu, cfg):
    # Fused only, layered only and mixed batches; tile widths from 1 to 64.
    trees = _mixed_trees()[:5]
    widths = _check_batch(trees, [3000] * 5, n_pass=2, **cfg)
    if cfg.get("max_tile_words") == 1:
        assert set(widths) == {1}
    if cfg.get("min_fused_tile_words") == 128:
        assert set(widths) == {0}


@pytest.mark.gpu
def test_batch_uneven_random_trees(gpu):
    # Trees of the batched workload's family (gate count log-uniform, events
    # 0.6 x gates), at a 2^20-trial prefix each; the large ones go through
    # the layered pipeline, the small ones through the fused launch.
    trees = mc.batched_trees(seed=3, n_trees=8, gate_range=(50, 5000))
    with mc.Batch(trees, 1 << 14) as b:
        info = b.info()
    assert info["num_fused_trees"] > 0 and info["num_layered_trees"] > 0
    _check_batch(trees, [1 << 14] * len(trees), n_pass=2)


@pytest.mark.gpu
def test_batch_layered_pipeline_mixed_shapes(gpu):
    # Two pipelines (two sample shapes), passes split across calls.
    trees = _mixed_trees()
    shapes = [1000, 2048, 1000, 2048, 1000, 2048, 1000]
    tallies = [_all_nodes(t) for t in trees]
    with mc.Batch(trees, shapes, tallies, min_fused_tile_words=128) as b:
        b.pass_(1)
        b.wait()
        b.pass_(2)
        got = [b.collect(i) for i in range(len(trees))]
    for i, (pd, n_bp) in enumerate(zip(trees, shapes)):
        want, _ = orc.run(pd, n_bp, n_iters=3, tally=tallies[i])
        assert [t.num_one_bits for t in got[i]] == [int(x) for x in want], i


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


def _edge_tree():
    """NOT/NULL with several inputs (the last one counts), a negated last
    input, ATLEAST with k = 0, k > n and k = 7 of 6, XOR with negations, and
    a gate wider than the inline inputs."""
    from mcscram_amd.pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag
    vs = [(2 + i, 0.05 + 0.1 * (i % 9)) for i in range(12)]
    g = [
        (20, NOT, 0, [2, 3, 4]), (21, NULL, 0, [5, -6]), (22, NOT, 0, [7, -8]),
        (23, ATLEAST, 0, [2, 3, 4]), (24, ATLEAST, 5, [2, 3, 4, 5]),
        (25, ATLEAST, 7, [2, 3, 4, 5, 6, 7]), (26, ATLEAST, 3, [2, -3, 4, -5, 6, 7]),
        (27, XOR, 0, [-2, 3, -4]), (28, NAND, 0, [9, -10, 11]), (29, NOR, 0, [12, -13]),
        (30, AND, 0, [-2, -3, -4, -5, -6, -7, -8]), (31, ATLEAST, 4, list(range(2, 14))),
        (32, OR, 0, [20, 21, 22, 23, 24, 25, 26, 27, 28, 29, 30, 31]),
    ]
    return Pdag.from_lists(vs, g, 32)


@pytest.mark.gpu
def test_batch_gate_normalisation_edges(gpu):
    _check_batch([_edge_tree(), _edge_tree()], [1000, (2, 512)], n_pass=2)


def test_edge_tree_oracle_against_enumeration():
    from tests.exact import exact_probabilities
    pd = _edge_tree()
    exact = exact_probabilities(pd)
    tally = [int(g) for g in pd.gate_index]
    ones, _ = orc.run(pd, 1 << 13, tally=tally)
    for idx, o in zip(tally, ones):
        m = int(o) / (1 << 19)
        se = max(np.sqrt(m * (1 - m) / (1 << 19)), 1 / (1 << 19))
        assert abs(m - exact[idx]) <= 3.5 * se, (idx, m, exact[idx])


@pytest.mark.gpu
def test_batch_of_all_aralia_models_bit_exact(gpu):
    # The reference's own models as one batch: the small ones fused in LDS,
    # the large ones dealt to two layered pipelines.
    import gzip
    import os
    from mcscram_amd import mef
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mef", "aralia")
    names = sorted(f for f in os.listdir(gdir) if f.endswith(".xml.gz"))
    trees = [mef.load_pdag(gzip.open(os.path.join(gdir, f))) for f in names]
    with mc.Batch(trees, 1000, [_all_nodes(t) for t in trees]) as b:
        info = b.info()
    assert info["num_fused_trees"] > 0 and info["num_layered_trees"] > 0
    _check_batch(trees, [1000] * len(trees), n_pass=2)


@pytest.mark.gpu
def test_batch_graph_replays_match_eager_passes(gpu):
    # Whole-pass graphs: the first pass runs eagerly, the second is captured,
    # later ones replay one of two graphs (pipelines of an odd number of
    # trees alternate event areas). Passes are split across calls with
    # collects (eager work) between replays; every collect must equal the
    # eager engine's, and the last one the oracle's.
    trees = _mixed_trees()
    shapes = [1000, 2048, 1000, 2048, 1000, 2048, 1000]
    tallies = [_all_nodes(t) for t in trees]
    runs = {}
    for g in (0, 1):
        seen = []
        with mc.Batch(trees, shapes, tallies, use_graph=g, min_fused_tile_words=32,
                      layered_streams=2) as b:
            for n in (1, 2, 1, 3):
                b.pass_(n)
                seen.append([[t.num_one_bits for t in b.collect(i)] for i in range(len(trees))])
            info = b.info()
        runs[g] = seen
    assert info["iteration"] == 7 and info["num_layered_trees"] > 0
    assert runs[0] == runs[1]
    for i, (pd, n_bp) in enumerate(zip(trees, shapes)):
        want, _ = orc.run(pd, n_bp, n_iters=7, tally=tallies[i])
        assert runs[1][-1][i] == [int(x) for x in want], i
