"""Host-side graph compilation: node depth, layers and the basic-event order
that fixes every Philox stream (src/preprocessor.cc:130-176,
src/mc/queue/layer_manager.tpp:33-118). Runs on the CPU: these entry points
of the engine library touch no device.
"""
import random

import numpy as np
import pytest

from mcscram_amd import large_tree, pdag_event_order, pdag_gate_depths, small_tree
from mcscram_amd.pdag import OR, Pdag
from oracle import binding as orc
from tests.introsort import reference_event_order, std_sort
from tests.models import small_random


def test_std_sort_restatement_sorts():
    rng = random.Random(4)
    for n in (0, 1, 2, 15, 16, 17, 100, 1000):
        a = [rng.randrange(50) for _ in range(n)]
        b = list(a)
        std_sort(b, lambda x, y: x < y)
        assert b == sorted(a)


def test_std_sort_restatement_heap_fallback():
    # Depth limit 0 forces the heap-sort path; it must still sort.
    from tests import introsort
    rng = random.Random(5)
    a = [rng.randrange(1000) for _ in range(300)]
    b = list(a)
    introsort._introsort_loop(b, 0, len(b), 0, lambda x, y: x < y)
    assert b == sorted(a)


def _depth_closed_form(pd: Pdag) -> dict:
    """Depth by longest path, computed bottom-up by repeated relaxation."""
    d = {int(v): 0 for v in pd.var_index}
    pending = {int(g): s for s, g in enumerate(pd.gate_index)}
    while pending:
        done = []
        for g, s in pending.items():
            args = [abs(int(a)) for a in pd.gate_args(s)]
            if all(a in d for a in args):
                d[g] = 1 + max(d[a] for a in args)
                done.append(g)
        assert done, "cycle"
        for g in done:
            del pending[g]
    return d


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_gate_depths(seed):
    pd = small_random(seed, n_be=30, n_gates=40, n_layers=6)
    got = pdag_gate_depths(pd)
    want = _depth_closed_form(pd)
    for s, g in enumerate(pd.gate_index):
        assert got[s] == want[int(g)]
    _, odepth, _ = orc.layout(pd)
    assert np.array_equal(odepth, got)


@pytest.mark.parametrize("maker", [lambda: small_tree(1), lambda: small_random(7),
                                   lambda: small_random(8, n_be=300, n_gates=200, n_layers=9)])
def test_event_order_matches_libstdcxx_restatement(maker):
    pd = maker()
    want = reference_event_order(pd)
    got = pdag_event_order(pd).tolist()
    assert got == want
    ev, _, _ = orc.layout(pd)
    inv = [None] * len(want)
    for s, e in enumerate(ev):
        if e >= 0:
            inv[e] = int(pd.var_index[s])
    assert inv == want


def test_event_order_large_tree():
    pd = large_tree()
    got = pdag_event_order(pd).tolist()
    assert sorted(got) == sorted(int(v) for v in pd.var_index)
    assert got == reference_event_order(pd)


def test_event_order_is_not_index_order():
    # The unstable sorts do permute: a plain index order would be a wrong
    # counter layout for most events.
    pd = small_tree(1)
    got = pdag_event_order(pd).tolist()
    assert got != sorted(got)


def test_unreachable_nodes_are_left_out():
    pd = Pdag.from_lists([(2, 0.1), (3, 0.2), (4, 0.3)],
                         [(5, OR, 0, [2, 3]), (6, OR, 0, [3, 4])], 5)
    assert sorted(pdag_event_order(pd).tolist()) == [2, 3]
    assert pdag_gate_depths(pd).tolist() == [1, -1]
