"""Sharding a run over trials: the ranks' disjoint iteration ranges cover the
run once, and the all-reduced counts equal the single run's.

The CPU test runs two gloo ranks; each rank's share is computed by the
oracle over its iteration range (the CPU stand-in for that rank's GPU), and
the product's all-reduce combines them. The GPU test runs the engine's
shards one after the other on the one GPU and sums them.
"""
import os

import numpy as np
import pytest

import mcscram_amd as mc
from mcscram_amd import shard
from oracle import binding as orc


def test_iteration_ranges_cover_the_run_once():
    for total in (0, 1, 7, 256, 1001):
        for ws in (1, 2, 3, 8, 13):
            rs = shard.iteration_ranges(total, ws)
            assert len(rs) == ws
            its = [f + i for f, n in rs for i in range(n)]
            assert its == list(range(1, total + 1))
            sizes = [n for _, n in rs]
            assert max(sizes) - min(sizes) <= 1
    with pytest.raises(ValueError):
        shard.iteration_ranges(4, 0)


def _rank(rank, world, port, words, total, out_dir):
    import torch.distributed as dist
    os.environ.update(MASTER_ADDR="127.0.0.1", MASTER_PORT=str(port))
    dist.init_process_group("gloo", rank=rank, world_size=world)
    try:
        pd = mc.small_tree(1)
        tally = [int(g) for g in pd.gate_index]
        first, n = shard.iteration_ranges(total, world)[rank]
        ones, _ = orc.run(pd, words, first_iter=first, n_iters=n, tally=tally) if n else \
            (np.zeros(len(tally), dtype=np.uint64), None)
        summed = shard.allreduce_ones(ones)
        np.save(os.path.join(out_dir, f"rank{rank}.npy"), summed)
    finally:
        dist.destroy_process_group()


def test_two_gloo_ranks_sum_to_the_single_run(tmp_path):
    import socket

    import torch.multiprocessing as tmp
    with socket.socket() as s:
        s.bind(("127.0.0.1", 0))
        port = s.getsockname()[1]
    words, total, world = 256, 5, 2
    tmp.spawn(_rank, args=(world, port, words, total, str(tmp_path)), nprocs=world, join=True)
    pd = mc.small_tree(1)
    want, _ = orc.run(pd, words, first_iter=1, n_iters=total,
                      tally=[int(g) for g in pd.gate_index])
    for r in range(world):
        got = np.load(tmp_path / f"rank{r}.npy")
        assert np.array_equal(got, want.astype(np.uint64)), r


@pytest.mark.gpu
def test_engine_shards_sum_to_one_run(gpu):
    pd = mc.small_tree(1)
    tally = [int(g) for g in pd.gate_index] + [int(v) for v in pd.var_index]
    words, total = 4096, 7
    with mc.LayerManager(pd, words, tally=tally) as eng:
        eng.pass_(total)
        whole = np.array([t.num_one_bits for t in eng.collect_tallies()], dtype=np.uint64)
    for world in (2, 3):
        parts = [shard.run_shard(pd, words, total, r, world, tally=tally) for r in range(world)]
        assert np.array_equal(np.sum(parts, axis=0, dtype=np.uint64), whole), world
