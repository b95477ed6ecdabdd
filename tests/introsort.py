"""Pure-Python restatement of libstdc++'s std::sort (introsort) and of the
reference's node gather, to pin the basic-event order independently of any
C++ build.

The reference fixes each basic event's Philox counter word 1 (event_idx + 1)
by its position after two unstable std::sort calls over a DFS gather of the
PDAG (src/mc/queue/layer_manager.tpp:33-118). The permutation an unstable
sort produces is a property of the library's algorithm, so it is restated
here from libstdc++'s published <bits/stl_algo.h> / <bits/stl_heap.h>
algorithm: introsort loop with a 16-element threshold and a depth limit of
2*floor(log2 n), median-of-three pivot moved to the first slot, unguarded
Hoare partition, heap-sort fallback, final insertion sort.
"""
from __future__ import annotations

from mcscram_amd.pdag import Pdag

_THRESHOLD = 16


def _lg(n: int) -> int:
    return n.bit_length() - 1


def _move_median_to_first(a, result, x, y, z, comp):
    if comp(a[x], a[y]):
        if comp(a[y], a[z]):
            a[result], a[y] = a[y], a[result]
        elif comp(a[x], a[z]):
            a[result], a[z] = a[z], a[result]
        else:
            a[result], a[x] = a[x], a[result]
    elif comp(a[x], a[z]):
        a[result], a[x] = a[x], a[result]
    elif comp(a[y], a[z]):
        a[result], a[z] = a[z], a[result]
    else:
        a[result], a[y] = a[y], a[result]


def _unguarded_partition(a, first, last, pivot, comp):
    while True:
        while comp(a[first], a[pivot]):
            first += 1
        last -= 1
        while comp(a[pivot], a[last]):
            last -= 1
        if not first < last:
            return first
        a[first], a[last] = a[last], a[first]
        first += 1


def _push_heap(a, base, hole, top, value, comp):
    parent = (hole - 1) // 2
    while hole > top and comp(a[base + parent], value):
        a[base + hole] = a[base + parent]
        hole = parent
        parent = (hole - 1) // 2
    a[base + hole] = value


def _adjust_heap(a, base, hole, length, value, comp):
    top = hole
    child = hole
    while child < (length - 1) // 2:
        child = 2 * (child + 1)
        if comp(a[base + child], a[base + child - 1]):
            child -= 1
        a[base + hole] = a[base + child]
        hole = child
    if (length & 1) == 0 and child == (length - 2) // 2:
        child = 2 * (child + 1)
        a[base + hole] = a[base + child - 1]
        hole = child - 1
    _push_heap(a, base, hole, top, value, comp)


def _heap_sort(a, first, last, comp):
    length = last - first
    if length >= 2:
        parent = (length - 2) // 2
        while True:
            _adjust_heap(a, first, parent, length, a[first + parent], comp)
            if parent == 0:
                break
            parent -= 1
    while last - first > 1:
        last -= 1
        value = a[last]
        a[last] = a[first]
        _adjust_heap(a, first, 0, last - first, value, comp)


def _introsort_loop(a, first, last, depth, comp):
    while last - first > _THRESHOLD:
        if depth == 0:
            _heap_sort(a, first, last, comp)
            return
        depth -= 1
        mid = first + (last - first) // 2
        _move_median_to_first(a, first, first + 1, mid, last - 1, comp)
        cut = _unguarded_partition(a, first + 1, last, first, comp)
        _introsort_loop(a, cut, last, depth, comp)
        last = cut


def _unguarded_linear_insert(a, i, comp):
    val = a[i]
    j = i - 1
    while comp(val, a[j]):
        a[i] = a[j]
        i = j
        j -= 1
    a[i] = val


def _insertion_sort(a, first, last, comp):
    if first == last:
        return
    for i in range(first + 1, last):
        if comp(a[i], a[first]):
            val = a[i]
            a[first + 1:i + 1] = a[first:i]
            a[first] = val
        else:
            _unguarded_linear_insert(a, i, comp)


def std_sort(a: list, comp) -> None:
    """In place, with libstdc++'s std::sort permutation for `comp`."""
    n = len(a)
    if n == 0:
        return
    _introsort_loop(a, 0, n, 2 * _lg(n), comp)
    if n > _THRESHOLD:
        _insertion_sort(a, 0, _THRESHOLD, comp)
        for i in range(_THRESHOLD, n):
            _unguarded_linear_insert(a, i, comp)
    else:
        _insertion_sort(a, 0, n, comp)


def reference_event_order(pd: Pdag) -> list[int]:
    """PDAG indices of the basic events in event_idx order, following
    layer_manager.tpp:33-118 and preprocessor.cc:130-176 step by step."""
    var_set = set(int(v) for v in pd.var_index)
    slot = {int(g): s for s, g in enumerate(pd.gate_index)}
    depth: dict[int, int] = {}

    def compute_depth(idx: int) -> int:
        if idx in depth:
            return depth[idx]
        if idx in var_set:
            depth[idx] = 0
            return 0
        d = max(compute_depth(abs(int(a))) for a in pd.gate_args(slot[idx])) + 1
        depth[idx] = d
        return d

    import sys
    old = sys.getrecursionlimit()
    sys.setrecursionlimit(max(old, 10 * (pd.n_gates + pd.n_vars) + 100))
    try:
        compute_depth(pd.root)
        nodes: list[tuple] = []  # (depth, is_var, connective, index)
        visited = set()

        def gather(g: int) -> None:
            if g in visited:
                return
            visited.add(g)
            nodes.append((depth[g], False, int(pd.gate_type[slot[g]]), g))
            args = [abs(int(a)) for a in pd.gate_args(slot[g])]
            for a in args:
                if a not in var_set:
                    gather(a)
            for a in args:
                if a in var_set and a not in visited:
                    visited.add(a)
                    nodes.append((0, True, 0, a))

        gather(pd.root)
    finally:
        sys.setrecursionlimit(old)

    std_sort(nodes, lambda x, y: x[0] < y[0])
    layers: list[list] = [[] for _ in range(nodes[-1][0] + 1)]
    for n in nodes:
        layers[n[0]].append(n)

    def by_kind(l, r):
        if l[1] and not r[1]:
            return True
        if not l[1] and r[1]:
            return False
        if l[1] and r[1]:
            return False
        return l[2] < r[2]

    for layer in layers:
        std_sort(layer, by_kind)
    return [n[3] for n in layers[0]]
