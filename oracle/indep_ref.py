"""CPU oracle for the independence approximation of the top event
(test infrastructure only; mcs_independence_logits computes it on the GPU).

Each gate's probability with its inputs treated as independent: AND the
product, OR one minus the product of the complements, XOR by the parity
recursion, ATLEAST the Poisson-binomial tail P(count >= k) by the full
distribution of the count, NOT/NULL the last input in the reference's input
order (positive variables, positive gates, complemented variables,
complemented gates; src/mc/queue/kernel_builder.h:298-426), NOT/NAND/NOR
complemented, a complemented input as 1 - q. numpy over samples, float64.
The approximation has no counterpart in the reference; this oracle is the
mathematics of the approximation itself, pinned in tests/test_indep.py by
exact enumeration on trees without shared events (where it is exact).
"""
from __future__ import annotations

import numpy as np

AND, OR, ATLEAST, XOR, NOT, NAND, NOR, NULL = range(8)


def _post_order(pd):
    gates = {int(g): s for s, g in enumerate(pd.gate_index)}
    order, state, stack = [], {}, [(int(pd.root), False)]
    while stack:
        x, done = stack.pop()
        if done:
            if state.get(x) != 2:
                state[x] = 2
                order.append(x)
            continue
        if state.get(x):
            continue
        state[x] = 1
        stack.append((x, True))
        for a in pd.gate_args(gates[x]):
            a = abs(int(a))
            if a in gates and not state.get(a):
                stack.append((a, False))
    return gates, order


def independence_top(pd, P) -> np.ndarray:
    """P: [n, n_vars] probabilities (order of pd.var_index) -> [n] top-event
    probability under the independence approximation."""
    P = np.asarray(P, dtype=np.float64)
    gates, order = _post_order(pd)
    col = {int(v): i for i, v in enumerate(pd.var_index)}
    val = {}

    def get(a):
        v = val[abs(a)] if abs(a) in val else P[:, col[abs(a)]]
        return v if a > 0 else 1.0 - v

    for g in order:
        s = gates[g]
        ps = [get(int(a)) for a in pd.gate_args(s)]
        t = int(pd.gate_type[s])
        n = P.shape[0]
        if t in (AND, NAND):
            v = np.prod(ps, axis=0) if ps else np.ones(n)
        elif t in (OR, NOR):
            v = 1.0 - (np.prod([1.0 - q for q in ps], axis=0) if ps else np.ones(n))
        elif t == XOR:
            v = np.zeros(n)
            for q in ps:
                v = v * (1.0 - q) + (1.0 - v) * q
        elif t in (NOT, NULL):
            # The last input in the reference's order: positive variables,
            # positive gates, complemented variables, complemented gates.
            args = [int(a) for a in pd.gate_args(s)]
            rank = [(2 if a < 0 else 0) + (0 if abs(a) in col else 1) for a in args]
            v = ps[max(range(len(args)), key=lambda i: (rank[i], i))] if ps else np.zeros(n)
        else:
            k = int(pd.gate_k[s])
            dist = [np.ones(n)] + [np.zeros(n) for _ in ps]  # P(count = j)
            for q in ps:
                for j in range(len(dist) - 1, 0, -1):
                    dist[j] = dist[j] * (1.0 - q) + dist[j - 1] * q
                dist[0] = dist[0] * (1.0 - q)
            v = np.sum(dist[k:], axis=0) if k <= len(ps) else np.zeros(n)
        if t in (NOT, NAND, NOR):
            v = 1.0 - v
        val[g] = v
    return val[int(pd.root)]


def independence_logits(pd, P) -> np.ndarray:
    """logit of independence_top with p clamped to [1e-300, 1 - 2^-52]."""
    p = np.clip(independence_top(pd, P), 1e-300, 1.0 - 2.0 ** -52)
    return np.log(p) - np.log1p(-p)
