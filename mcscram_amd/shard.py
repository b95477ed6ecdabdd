"""Sharding a Monte Carlo run over trials, one process per GPU.

The Philox counter of every draw is a pure function of (event, sample word,
iteration) (src/mc/kernel/basic_event.h:112-122), so the passes of a run
split into disjoint iteration ranges, one per rank: rank r runs iterations
[first_r, first_r + n_r) on its own GPU, and the sum of the ranks' one-bit
counts is the single-GPU run's count, bit for bit. The only exchange is one
sum all-reduce of the counts at collect time (8 bytes per tallied node).

``iteration_ranges`` deals the passes; ``run_shard`` runs one rank's share
through the engine (restoring its iteration counter to the range's start);
``allreduce_ones`` sums the counts over a torch.distributed process group
(RCCL's "nccl" backend on GPUs, gloo on host tensors).
"""
from __future__ import annotations

import numpy as np


def iteration_ranges(total_passes: int, world_size: int) -> list[tuple[int, int]]:
    """(first iteration, passes) of each rank: the 1-based iterations
    1..total_passes in contiguous ranges whose sizes differ by at most one."""
    if world_size < 1 or total_passes < 0:
        raise ValueError("world_size must be positive and total_passes non-negative")
    base, extra = divmod(total_passes, world_size)
    out, first = [], 1
    for r in range(world_size):
        n = base + (1 if r < extra else 0)
        out.append((first, n))
        first += n
    return out


def run_shard(pdag, shape, total_passes: int, rank: int, world_size: int, tally=None,
              device: int = -1) -> np.ndarray:
    """One rank's share of a run: the engine's iteration counter set to the
    range's first iteration - 1 with zero counts, then its passes.
    -> uint64 one-bit counts in tally order (first appearance)."""
    from .engine import LayerManager
    first, n = iteration_ranges(total_passes, world_size)[rank]
    with LayerManager(pdag, shape, tally=tally, device=device) as eng:
        k = len(eng.collect_tallies())
        eng.restore(first - 1, np.zeros(k, dtype=np.uint64))
        if n:
            eng.pass_(n)
        return np.array([t.num_one_bits for t in eng.collect_tallies()], dtype=np.uint64)


def allreduce_ones(ones: np.ndarray, group=None) -> np.ndarray:
    """Sum of every rank's counts (a torch.distributed SUM all-reduce of
    int64: counts stay far below 2^63). Host tensors, as gloo needs; an
    nccl group takes the same call with the tensor moved to the GPU."""
    import torch
    import torch.distributed as dist
    t = torch.from_numpy(np.ascontiguousarray(ones, dtype=np.uint64).view(np.int64).copy())
    if dist.get_backend(group) == "nccl":
        d = t.cuda()
        dist.all_reduce(d, op=dist.ReduceOp.SUM, group=group)
        t = d.cpu()
    else:
        dist.all_reduce(t, op=dist.ReduceOp.SUM, group=group)
    return t.numpy().view(np.uint64).copy()
