"""Flattened PDAG description and synthetic layered fault-tree generators.

The engine consumes a propositional DAG (PDAG) in the shape the reference's
Monte Carlo engine walks it (src/pdag.h:298-397):

* variables (basic events) carry a PDAG index (sequential from 2, as in
  src/pdag.h:804 and src/pdag.cc:540-551) and a probability;
* gates carry a PDAG index, a connective (numbering of src/pdag.h:265-275),
  a min number for ATLEAST, and signed argument indices (negative = the
  complement of the argument) in insertion order. The reference keeps
  variable and gate arguments in two insertion-ordered maps
  (src/pdag.h:311-321); the order of a flat list restricted to either kind
  is that map's order.

Everything is held as numpy arrays so it can be handed to the C ABI without
copies. The generators use their own SplitMix64 stream so a seed gives the
same tree on every platform (no library distribution is involved).
"""
from __future__ import annotations

import bisect
import math
from dataclasses import dataclass, field

import numpy as np

# Connective numbering (src/pdag.h:265-275).
AND, OR, ATLEAST, XOR, NOT, NAND, NOR, NULL = range(8)
CONNECTIVE_NAMES = ("and", "or", "atleast", "xor", "not", "nand", "nor", "null")
VARIABLE_START_INDEX = 2


@dataclass
class Pdag:
    """Flattened PDAG: struct-of-arrays, ready for the C ABI."""

    var_index: np.ndarray  # int32 [n_vars]
    var_p: np.ndarray  # float64 [n_vars]
    gate_index: np.ndarray  # int32 [n_gates]
    gate_type: np.ndarray  # uint8 [n_gates]
    gate_k: np.ndarray  # int32 [n_gates]
    arg_off: np.ndarray  # int64 [n_gates + 1]
    args: np.ndarray  # int32 [n_args], signed pdag indices
    root: int
    meta: dict = field(default_factory=dict)

    @property
    def n_vars(self) -> int:
        return int(self.var_index.shape[0])

    @property
    def n_gates(self) -> int:
        return int(self.gate_index.shape[0])

    @property
    def n_args(self) -> int:
        return int(self.args.shape[0])

    def fan_in(self) -> np.ndarray:
        return np.diff(self.arg_off)

    def gate_args(self, slot: int) -> np.ndarray:
        return self.args[self.arg_off[slot]:self.arg_off[slot + 1]]

    def with_probabilities(self, p: np.ndarray) -> "Pdag":
        p = np.ascontiguousarray(p, dtype=np.float64)
        if p.shape != self.var_p.shape:
            raise ValueError("probability vector has the wrong length")
        return Pdag(self.var_index, p, self.gate_index, self.gate_type, self.gate_k,
                    self.arg_off, self.args, self.root, dict(self.meta))

    def subgraph(self, top: int) -> "Pdag":
        """The PDAG of gate ``top`` alone: the nodes reachable from it, with
        their indices, probabilities and argument lists unchanged, in their
        original order, and ``top`` as the root. This is the graph a fault
        tree whose top event is that gate compiles to."""
        slot = {int(g): s for s, g in enumerate(self.gate_index)}
        if int(top) not in slot:
            raise ValueError("top is not a gate")
        keep_g = np.zeros(self.n_gates, dtype=bool)
        keep_v = set()
        stack = [slot[int(top)]]
        keep_g[stack[0]] = True
        while stack:
            for a in self.gate_args(stack.pop()):
                a = abs(int(a))
                s = slot.get(a)
                if s is None:
                    keep_v.add(a)
                elif not keep_g[s]:
                    keep_g[s] = True
                    stack.append(s)
        vmask = np.isin(self.var_index, np.fromiter(keep_v, dtype=np.int64, count=len(keep_v)))
        gs = np.flatnonzero(keep_g)
        lens = self.arg_off[gs + 1] - self.arg_off[gs]
        off = np.zeros(gs.shape[0] + 1, dtype=np.int64)
        np.cumsum(lens, out=off[1:])
        args = np.concatenate([self.gate_args(s) for s in gs]).astype(np.int32)
        return Pdag(self.var_index[vmask].copy(), self.var_p[vmask].copy(), self.gate_index[gs].copy(),
                    self.gate_type[gs].copy(), self.gate_k[gs].copy(), off, args, int(top),
                    dict(self.meta, top=int(top)))

    def validate(self) -> None:
        idx = np.concatenate([self.var_index, self.gate_index])
        if np.unique(idx).shape[0] != idx.shape[0]:
            raise ValueError("duplicate pdag index")
        if self.arg_off.shape[0] != self.n_gates + 1 or self.arg_off[0] != 0 or \
                self.arg_off[-1] != self.n_args:
            raise ValueError("bad argument offsets")
        if not np.all((self.var_p >= 0.0) & (self.var_p <= 1.0)):
            raise ValueError("probability outside [0, 1]")
        known = set(idx.tolist())
        if not set(np.abs(self.args).tolist()) <= known:
            raise ValueError("unknown argument index")
        if self.root not in set(self.gate_index.tolist()):
            raise ValueError("root is not a gate")

    @staticmethod
    def from_lists(variables, gates, root) -> "Pdag":
        """variables: [(index, p)]; gates: [(index, connective, k, [signed args])]."""
        var_index = np.array([v[0] for v in variables], dtype=np.int32)
        var_p = np.array([v[1] for v in variables], dtype=np.float64)
        gate_index = np.array([g[0] for g in gates], dtype=np.int32)
        gate_type = np.array([g[1] for g in gates], dtype=np.uint8)
        gate_k = np.array([g[2] for g in gates], dtype=np.int32)
        off = [0]
        flat = []
        for g in gates:
            flat.extend(int(a) for a in g[3])
            off.append(len(flat))
        return Pdag(var_index, var_p, gate_index, gate_type, gate_k,
                    np.array(off, dtype=np.int64), np.array(flat, dtype=np.int32), int(root))


class SplitMix64:
    """SplitMix64 (Steele, Lea, Flood 2014): a portable 64-bit stream."""

    MASK = (1 << 64) - 1

    def __init__(self, seed: int):
        self.state = seed & self.MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & self.MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & self.MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & self.MASK
        return z ^ (z >> 31)

    def uniform(self) -> float:
        """Uniform double in [0, 1) from the top 53 bits."""
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def randint(self, lo: int, hi: int) -> int:
        """Uniform integer in [lo, hi] (multiply-shift, bias < 2^-32 here)."""
        span = hi - lo + 1
        return lo + ((self.next_u64() * span) >> 64)

    def log_uniform(self, lo: float, hi: float) -> float:
        return math.exp(math.log(lo) + self.uniform() * (math.log(hi) - math.log(lo)))

    def sample_distinct(self, n_pool: int, k: int) -> list[int]:
        """k distinct draws from range(n_pool), in draw order."""
        k = min(k, n_pool)
        if k * 4 > n_pool:
            pool = list(range(n_pool))
            for i in range(k):  # partial Fisher-Yates
                j = self.randint(i, n_pool - 1)
                pool[i], pool[j] = pool[j], pool[i]
            return pool[:k]
        seen, out = set(), []
        while len(out) < k:
            c = self.randint(0, n_pool - 1)
            if c not in seen:
                seen.add(c)
                out.append(c)
        return out


def random_layered_tree(seed: int, n_be: int, n_gates: int, n_layers: int,
                        p_range=(1e-4, 1e-1), fan_in=(2, 4), mix=(0.45, 0.45, 0.10),
                        atleast_n=(3, 6), neg_frac=0.0, all_connectives=False,
                        top="orphans") -> Pdag:
    """Random layered fault tree with a single OR top gate.

    * Basic events: PDAG indices 2..n_be+1, p log-uniform in ``p_range``.
    * ``n_gates`` gates including the top. The top OR gate is generator layer
      ``n_layers``; the other gates are spread evenly over layers
      1..n_layers-1. A gate in layer l draws its inputs uniformly, without
      repetition, from every node of layers < l (basic events are layer 0).
    * Connective by ``mix`` = (P(AND), P(OR), P(ATLEAST)); AND/OR fan-in
      uniform in ``fan_in``; ATLEAST n uniform in ``atleast_n`` with k uniform
      in 2..n-1. With ``all_connectives`` a fifth of the gates are instead
      drawn from XOR (2 inputs), NOT and NULL (1 input), NAND and NOR.
    * Each input is complemented with probability ``neg_frac``.
    * Basic events no gate drew are appended to layer-1 gates (OR gates
      first) so every basic event is reachable.
    * ``top="orphans"``: every gate no other gate drew feeds the top OR gate.
      ``top="last_layer"``: the top OR gate reads the last generator layer's
      gates, and gates of lower layers that no gate drew become an extra
      input of a random AND gate of a higher layer (any gate if the higher
      layers hold no AND gate). Both keep every gate reachable; the second
      keeps the top gate's fan-in, and so its probability, moderate.

    The reference's layer of a gate is its computed depth, which can be lower
    than the generator layer.
    """
    rng = SplitMix64(seed)
    var_index = np.arange(VARIABLE_START_INDEX, VARIABLE_START_INDEX + n_be, dtype=np.int32)
    var_p = np.array([rng.log_uniform(*p_range) for _ in range(n_be)], dtype=np.float64)

    n_body = n_gates - 1
    n_mid = max(n_layers - 1, 1)
    per_layer = [n_body // n_mid + (1 if l < n_body % n_mid else 0) for l in range(n_mid)]
    first_gate_index = VARIABLE_START_INDEX + n_be
    pool = list(range(VARIABLE_START_INDEX, VARIABLE_START_INDEX + n_be))  # nodes of lower layers
    gates = []  # [index, type, k, args]
    used = set()
    next_index = first_gate_index
    layer_of = []
    for l, count in enumerate(per_layer, start=1):
        new = []
        for _ in range(count):
            u = rng.uniform()
            k = 0
            if all_connectives and rng.uniform() < 0.2:
                t = (XOR, NOT, NULL, NAND, NOR)[rng.randint(0, 4)]
                f = {XOR: 2, NOT: 1, NULL: 1}.get(t, rng.randint(*fan_in))
            elif u < mix[0]:
                t, f = AND, rng.randint(*fan_in)
            elif u < mix[0] + mix[1]:
                t, f = OR, rng.randint(*fan_in)
            else:
                t = ATLEAST
                f = rng.randint(*atleast_n)
                k = rng.randint(2, f - 1)
            picks = rng.sample_distinct(len(pool), f)
            args = []
            for pi in picks:
                node = pool[pi]
                used.add(node)
                args.append(-node if (neg_frac > 0 and rng.uniform() < neg_frac) else node)
            if t == ATLEAST and len(args) <= k:
                k = max(1, len(args) - 1)
            gates.append([next_index, t, k, args])
            layer_of.append(l)
            new.append(next_index)
            next_index += 1
        pool.extend(new)

    # Unused basic events go to layer-1 gates, OR gates first.
    orphans = [v for v in range(VARIABLE_START_INDEX, VARIABLE_START_INDEX + n_be) if v not in used]
    l1 = [i for i, l in enumerate(layer_of) if l == 1 and gates[i][1] not in (NOT, NULL, XOR)]
    l1_or = [i for i in l1 if gates[i][1] == OR] or l1
    for j, v in enumerate(orphans):
        gi = l1_or[j % len(l1_or)]
        gates[gi][3].append(v)
        used.add(v)

    if top == "last_layer":
        last = max(layer_of)
        # Gates are generated layer by layer, so the gates of layers above l
        # are the index range from first_above[l] on; AND gates likewise.
        first_above = {l: bisect.bisect_right(layer_of, l) for l in set(layer_of)}
        ands_below_last = [j for j in range(len(gates))
                           if gates[j][1] == AND and layer_of[j] < last]
        ands_all = [j for j in range(len(gates)) if gates[j][1] == AND]
        for gi, g in enumerate(gates):
            if layer_of[gi] == last or g[0] in used:
                continue
            lo = first_above[layer_of[gi]]
            ands = ands_below_last[bisect.bisect_left(ands_below_last, lo):] or \
                ands_all[bisect.bisect_left(ands_all, lo):] or range(lo, len(gates))
            j = ands[rng.randint(0, len(ands) - 1)]
            gates[j][3].append(g[0])
            used.add(g[0])
        top_args = [g[0] for gi, g in enumerate(gates) if layer_of[gi] == last]
    elif top == "orphans":
        top_args = [g[0] for g in gates if g[0] not in used]
    else:
        raise ValueError(f"top: {top}")
    if len(top_args) < 2:
        extra = [g[0] for g in gates[::-1] if g[0] not in top_args]
        top_args.extend(extra[:2 - len(top_args)])
    root = next_index
    gates.append([root, OR, 0, top_args])

    variables = list(zip(var_index.tolist(), var_p.tolist()))
    pdag = Pdag.from_lists(variables, [tuple(g) for g in gates], root)
    pdag.meta = {"seed": seed, "n_be": n_be, "n_gates": n_gates, "n_layers": n_layers,
                 "generator": "random_layered_tree"}
    return pdag


def small_tree(seed: int = 1) -> Pdag:
    """Plain MC on a small random layered tree: 200 basic events with p
    log-uniform in [1e-4, 1e-1]; 120 gates in 8 layers, 45% AND / 45% OR /
    10% ATLEAST (n 3-6, k 2..n-1), fan-in 2-4, single OR top gate."""
    return random_layered_tree(seed, 200, 120, 8, p_range=(1e-4, 1e-1), fan_in=(2, 4),
                               mix=(0.45, 0.45, 0.10), atleast_n=(3, 6))


def large_tree(seed: int = 2) -> Pdag:
    """HBM-scale tree: 25,000 basic events with p log-uniform in [1e-6, 1e-2];
    40,000 gates in 30 layers, fan-in 2-8, 48% AND / 47% OR / 5% ATLEAST
    (n up to 16), 10% of gate inputs complemented."""
    return random_layered_tree(seed, 25_000, 40_000, 30, p_range=(1e-6, 1e-2), fan_in=(2, 8),
                               mix=(0.48, 0.47, 0.05), atleast_n=(3, 16), neg_frac=0.10)


def batched_trees(seed: int = 3, n_trees: int = 256, gate_range=(50, 5000)) -> list[Pdag]:
    """Independent trees of uneven size: gate count log-uniform in
    ``gate_range``, basic events = 0.6 x gates, mix and p as the small tree."""
    rng = SplitMix64(seed)
    trees = []
    for t in range(n_trees):
        n_g = int(round(rng.log_uniform(*gate_range)))
        n_be = max(2, int(round(0.6 * n_g)))
        n_layers = max(3, min(30, int(round(math.log2(n_g)))))
        tree_seed = rng.next_u64()
        trees.append(random_layered_tree(tree_seed, n_be, n_g, n_layers, p_range=(1e-4, 1e-1),
                                         fan_in=(2, 4), mix=(0.45, 0.45, 0.10), atleast_n=(3, 6)))
    return trees
