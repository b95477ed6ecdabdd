"""Exact gate probabilities by enumeration of every basic-event state.

Independent of the oracle and of the engine: a gate's value on one state is
its Boolean definition (AND = all inputs, OR = any, XOR = parity, ATLEAST =
at least k true, NOT/NULL of one input, NAND/NOR the complements), and its
probability is the sum of the state weights where it is true. Practical up
to about 20 basic events.
"""
from __future__ import annotations

import numpy as np

from mcscram_amd.pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag


def exact_probabilities(pd: Pdag) -> dict[int, float]:
    """PDAG index -> exact probability, for every variable and gate."""
    n = pd.n_vars
    if n > 22:
        raise ValueError("too many basic events to enumerate")
    states = np.arange(1 << n, dtype=np.int64)
    weight = np.ones(1 << n, dtype=np.float64)
    value: dict[int, np.ndarray] = {}
    for i in range(n):
        bit = ((states >> i) & 1).astype(bool)
        p = float(pd.var_p[i])
        weight *= np.where(bit, p, 1.0 - p)
        value[int(pd.var_index[i])] = bit
    slot_of = {int(g): s for s, g in enumerate(pd.gate_index)}

    def gate(idx: int) -> np.ndarray:
        if idx in value:
            return value[idx]
        s = slot_of[idx]
        args = [int(a) for a in pd.gate_args(s)]
        # The reference's input order: positive inputs, then negated ones,
        # variables before gates within each (kernel_builder.h:298-426);
        # NOT and NULL read the last of them (gate.h:321-351).
        is_var = lambda a: abs(a) in value and abs(a) not in slot_of  # noqa: E731
        order = [a for a in args if a > 0 and is_var(a)] + [a for a in args if a > 0 and not is_var(a)] + \
                [a for a in args if a < 0 and is_var(a)] + [a for a in args if a < 0 and not is_var(a)]
        ins = [gate(abs(a)) if a > 0 else ~gate(abs(a)) for a in order]
        t = int(pd.gate_type[s])
        if t in (AND, NAND):
            r = np.logical_and.reduce(ins)
        elif t in (OR, NOR):
            r = np.logical_or.reduce(ins)
        elif t == XOR:
            r = np.logical_xor.reduce(ins)
        elif t == ATLEAST:
            r = np.sum(ins, axis=0) >= int(pd.gate_k[s])
        elif t in (NOT, NULL):
            r = ins[-1] if ins else np.zeros_like(weight, dtype=bool)
        else:
            raise ValueError(f"connective {t}")
        if t in (NAND, NOR, NOT):
            r = ~r
        value[idx] = r
        return r

    out = {}
    for idx in list(value.keys()) + [int(g) for g in pd.gate_index]:
        out[idx] = float(weight[gate(idx)].sum())
    return out
