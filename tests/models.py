"""Small PDAGs for tests: the reference's core known-answer models and
random small trees.

The core models restate tests/input/core/*.xml of the reference as PDAGs
(variables A, B, C = indices 2, 3, 4 with p = 0.1, 0.2, 0.3), with the exact
top-event probabilities its tests assert (tests/bench_core_tests.cc:23-31,
39-47, 55-63, 81-89, 328-336, and nand.xml).
"""
from __future__ import annotations

from mcscram_amd.pdag import AND, ATLEAST, NAND, OR, XOR, Pdag, random_layered_tree

ABC = [(2, 0.1), (3, 0.2), (4, 0.3)]


def core_models() -> dict[str, tuple[Pdag, float]]:
    """name -> (pdag, exact top-event probability asserted by the reference)."""
    m = {}
    m["abc"] = (Pdag.from_lists(ABC, [(5, OR, 0, [2, 3, 4])], 5), 0.496)
    m["ab_bc"] = (Pdag.from_lists(ABC, [(5, AND, 0, [2, 3]), (6, AND, 0, [3, 4]),
                                        (7, OR, 0, [5, 6])], 7), 0.074)
    m["ab_or_not_ac"] = (Pdag.from_lists(ABC, [(5, AND, 0, [2, 3]), (6, AND, 0, [-2, 4]),
                                               (7, OR, 0, [5, 6])], 7), 0.29)
    m["atleast"] = (Pdag.from_lists(ABC, [(5, ATLEAST, 2, [2, 3, 4])], 5), 0.098)
    m["xor"] = (Pdag.from_lists(ABC, [(5, XOR, 0, [3, 4]), (6, XOR, 0, [2, 5])], 6), 0.404)
    m["nand"] = (Pdag.from_lists(ABC[:2], [(4, NAND, 0, [2, 3])], 4), 0.98)
    return m


def small_random(seed: int, n_be: int = 12, n_gates: int = 10, n_layers: int = 4,
                 all_connectives: bool = True, neg_frac: float = 0.15) -> Pdag:
    """A tree small enough for exact enumeration, with every connective."""
    return random_layered_tree(seed, n_be, n_gates, n_layers, p_range=(0.05, 0.6),
                               fan_in=(2, 4), mix=(0.4, 0.4, 0.2), atleast_n=(3, 5),
                               neg_frac=neg_frac, all_connectives=all_connectives)
