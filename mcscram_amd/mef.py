"""Open-PSA MEF fault trees -> flattened PDAG, as the reference builds it with
no preprocessing (compilation level 0, src/mc/core/preprocessor.cc:26-28).

What is restated, and from where:

* Formulas (src/initializer.cc:792-869): the connective is the element name;
  an element that carries a ``name`` attribute is a pass-through (NULL)
  formula over that one event; ``<not>`` around an event reference is a
  complemented argument; ``<event name=... type=...>`` is resolved by type, or
  by name when the type is absent. A formula names each event at most once
  (src/event.cc:60-71).
* Top events (src/fault_tree.cc:63-93): the gates of a fault tree that no
  other gate of that tree uses.
* PDAG construction from a top gate (src/pdag.cc:471-496):
  - the constant node is created first and takes index 1; variables take
    indices 2, 3, ... in the order GatherVariables meets them, a depth-first
    walk of formula arguments in document order that descends into a gate the
    first time it is seen (src/pdag.cc:516-554);
  - gates take the next indices in the order ConstructGate creates them: the
    parent before its arguments, an argument gate when first met
    (src/pdag.cc:614-656);
  - arguments keep formula order; a complemented argument is the negative
    index (src/pdag.cc:575-611);
  - iff and imply become NULL(~XOR(a, b)) and OR(~a, b)
    (src/pdag.cc:658-676).
* Basic-event probability: the value of the event's expression, which the
  Monte Carlo engine reads as ``expression().value()``
  (src/mc/queue/kernel_builder.h:174-184). Supported expressions: float, int,
  bool, parameter references, exponential (1 - exp(-lambda t)), the system
  mission time (default 8760, src/settings.h:445) and the deviates whose value
  is their mean (uniform, normal, lognormal with a mean argument).

Not supported, each with a clear error: house events and constants (the
reference's Monte Carlo engine has no kernel for the constant node, so it needs
preprocessing to remove them), CCF groups, cardinality gates, and substitutions.
"""
from __future__ import annotations

import math
import os
import xml.etree.ElementTree as ET
from dataclasses import dataclass, field

from .pdag import AND, ATLEAST, NAND, NOR, NOT, NULL, OR, XOR, Pdag, VARIABLE_START_INDEX

_CONNECTIVES = {"and": AND, "or": OR, "atleast": ATLEAST, "xor": XOR, "not": NOT,
                "nand": NAND, "nor": NOR, "null": NULL}
_COMPLEX = ("iff", "imply", "cardinality")
_SKIP = ("label", "attributes")


class MefError(ValueError):
    """A model the loader cannot turn into a PDAG."""


@dataclass
class Formula:
    connective: str
    args: list  # [(complement: bool, kind: str, name: str)]
    min_number: int | None = None


@dataclass
class Model:
    """The parts of an MEF model the Monte Carlo path reads."""

    gates: dict = field(default_factory=dict)  # name -> Formula
    gate_tree: dict = field(default_factory=dict)  # name -> fault-tree name
    basic_events: dict = field(default_factory=dict)  # name -> probability
    house_events: set = field(default_factory=set)
    parameters: dict = field(default_factory=dict)  # name -> XML expression element
    fault_trees: list = field(default_factory=list)
    mission_time: float = 8760.0

    def top_events(self, fault_tree: str | None = None) -> list[str]:
        """Gates of a fault tree that no gate of that tree uses, in definition order."""
        names = [g for g, t in self.gate_tree.items() if fault_tree is None or t == fault_tree]
        own = set(names)
        used = set()
        for g in names:
            for _, kind, a in self.gates[g].args:
                if kind in ("gate", "event") and a in own and \
                        self.gate_tree[a] == self.gate_tree[g]:
                    used.add(a)
        return [g for g in names if g not in used]


def _local(tag: str) -> str:
    return tag.rsplit("}", 1)[-1]


def _children(el):
    return [c for c in el if isinstance(c.tag, str) and _local(c.tag) not in _SKIP]


class _Parser:
    def __init__(self, model: Model):
        self.m = model
        self._p_cache: dict = {}

    def feed(self, root) -> None:
        if _local(root.tag) != "opsa-mef":
            raise MefError(f"not an Open-PSA MEF document (root <{_local(root.tag)}>)")
        for el in _children(root):
            tag = _local(el.tag)
            if tag == "define-fault-tree":
                self.m.fault_trees.append(el.get("name"))
                self._container(el, el.get("name"))
            elif tag == "model-data":
                self._container(el, None)
            elif tag in ("define-CCF-group", "define-substitution"):
                raise MefError(f"<{tag}> is not supported")
            elif tag in ("define-event-tree", "define-initiating-event", "define-sequence",
                         "define-rule", "define-alignment", "define-consequence"):
                continue  # event-tree analysis is not on the Monte Carlo fault-tree path
            else:
                self._definition(el, None)

    def _container(self, el, tree) -> None:
        for c in _children(el):
            tag = _local(c.tag)
            if tag == "define-component":
                self._container(c, tree)
            elif tag == "define-CCF-group":
                raise MefError("CCF groups are not supported")
            else:
                self._definition(c, tree)

    def _definition(self, el, tree) -> None:
        tag, name = _local(el.tag), el.get("name")
        if tag == "define-gate":
            if tree is None:
                raise MefError(f"gate {name} outside a fault tree")
            if name in self.m.gates:
                raise MefError(f"gate {name} defined twice")
            body = _children(el)
            if len(body) != 1:
                raise MefError(f"gate {name}: expected one formula")
            self.m.gates[name] = self._formula(body[0], name)
            self.m.gate_tree[name] = tree
        elif tag == "define-basic-event":
            if name in self.m.basic_events:
                raise MefError(f"basic event {name} defined twice")
            body = _children(el)
            self.m.basic_events[name] = body[0] if body else None
        elif tag == "define-house-event":
            self.m.house_events.add(name)
        elif tag == "define-parameter":
            body = _children(el)
            if len(body) != 1:
                raise MefError(f"parameter {name}: expected one expression")
            self.m.parameters[name] = body[0]
        elif tag in ("define-CCF-group", "define-substitution"):
            raise MefError(f"<{tag}> is not supported")

    def _formula(self, el, gate: str) -> Formula:
        tag = _local(el.tag)
        if tag == "constant":
            raise MefError(f"gate {gate}: constants need preprocessing (not supported)")
        if el.get("name") is not None:  # pass-through over one event
            return Formula("null", [self._arg(el, False, gate)])
        if tag in _COMPLEX:
            if tag == "cardinality":
                raise MefError(f"gate {gate}: cardinality is not supported")
            args = [self._formula_arg(c, gate) for c in _children(el)]
            if len(args) != 2:
                raise MefError(f"gate {gate}: {tag} takes two arguments")
            return Formula(tag, args)
        if tag not in _CONNECTIVES:
            raise MefError(f"gate {gate}: unknown connective <{tag}>")
        args = [self._formula_arg(c, gate) for c in _children(el)]
        names = [(k, a) for _, k, a in args]
        if len(set(names)) != len(names):
            raise MefError(f"gate {gate}: an event is named twice")
        k = None
        if tag == "atleast":
            if el.get("min") is None:
                raise MefError(f"gate {gate}: atleast without min")
            k = int(el.get("min"))
        return Formula(tag, args, k)

    def _formula_arg(self, c, gate):
        tag = _local(c.tag)
        if tag == "not":
            inner = _children(c)
            if len(inner) != 1:
                raise MefError(f"gate {gate}: <not> takes one event")
            return self._arg(inner[0], True, gate)
        if tag == "constant":
            raise MefError(f"gate {gate}: constants need preprocessing (not supported)")
        return self._arg(c, False, gate)

    def _arg(self, c, complement, gate):
        tag = _local(c.tag)
        name = c.get("name")
        if name is None:
            raise MefError(f"gate {gate}: nested formula <{tag}> is not an event reference")
        kind = c.get("type") or tag
        if kind not in ("gate", "basic-event", "house-event", "event"):
            raise MefError(f"gate {gate}: unknown event kind <{kind}>")
        return (complement, kind, name)

    # Expressions -------------------------------------------------------
    def value(self, el, where: str) -> float:
        tag = _local(el.tag)
        kids = _children(el)
        if tag == "float":
            return float(el.get("value"))
        if tag == "int":
            return float(int(el.get("value")))
        if tag == "bool":
            return 1.0 if el.get("value") in ("true", "1") else 0.0
        if tag == "parameter":
            name = el.get("name")
            if name not in self.m.parameters:
                raise MefError(f"{where}: undefined parameter {name}")
            if name not in self._p_cache:
                self._p_cache[name] = self.value(self.m.parameters[name], f"parameter {name}")
            return self._p_cache[name]
        if tag == "system-mission-time":
            return self.m.mission_time
        if tag == "exponential":
            lam, t = (self.value(k, where) for k in kids)
            return -math.expm1(-lam * t)
        if tag == "uniform-deviate":
            a, b = (self.value(k, where) for k in kids)
            return 0.5 * (a + b)
        if tag in ("normal-deviate", "lognormal-deviate"):
            return self.value(kids[0], where)  # the mean
        raise MefError(f"{where}: expression <{tag}> is not supported")


def load_model(paths, mission_time: float = 8760.0) -> Model:
    """Parse one or more MEF files (paths or file objects) into one model."""
    if isinstance(paths, (str, os.PathLike)) or hasattr(paths, "read"):
        paths = [paths]
    m = Model(mission_time=mission_time)
    parser = _Parser(m)
    for p in paths:
        parser.feed(ET.parse(p).getroot())
    for name, expr in list(m.basic_events.items()):
        if expr is None:
            raise MefError(f"basic event {name} has no probability expression")
        p = parser.value(expr, f"basic event {name}")
        if not 0.0 <= p <= 1.0:
            raise MefError(f"basic event {name}: probability {p} outside [0, 1]")
        m.basic_events[name] = p
    return m


def _resolve(m: Model, kind: str, name: str) -> str:
    if kind == "event":
        if name in m.gates:
            return "gate"
        if name in m.basic_events:
            return "basic-event"
        if name in m.house_events:
            return "house-event"
        raise MefError(f"undefined event {name}")
    table = {"gate": m.gates, "basic-event": m.basic_events, "house-event": m.house_events}[kind]
    if name not in table:
        raise MefError(f"undefined {kind} {name}")
    return kind


def build_pdag(m: Model, top: str | None = None) -> Pdag:
    """The level-0 PDAG of one top gate (the model's only top event when
    ``top`` is None). ``meta`` maps PDAG indices back to MEF names."""
    if top is None:
        tops = m.top_events()
        if len(tops) != 1:
            raise MefError(f"model has {len(tops)} top events; name one of {tops[:8]}")
        top = tops[0]
    if top not in m.gates:
        raise MefError(f"undefined gate {top}")

    # GatherVariables: depth-first over formula arguments, descending into a
    # gate the first time it is met (src/pdag.cc:516-554).
    var_of: dict[str, int] = {}
    seen_gates = {top}
    stack = [iter(m.gates[top].args)]
    while stack:
        arg = next(stack[-1], None)
        if arg is None:
            stack.pop()
            continue
        _, kind, name = arg
        kind = _resolve(m, kind, name)
        if kind == "house-event":
            raise MefError(f"house event {name}: constants need preprocessing (not supported)")
        if kind == "basic-event":
            if name not in var_of:
                var_of[name] = VARIABLE_START_INDEX + len(var_of)
        elif name not in seen_gates:
            seen_gates.add(name)
            stack.append(iter(m.gates[name].args))

    next_index = [VARIABLE_START_INDEX + len(var_of)]
    gates: list = []  # [index, connective, k, args]
    gate_of: dict[str, int] = {}

    def new_gate(conn, k=0):
        g = [next_index[0], conn, k, []]
        next_index[0] += 1
        gates.append(g)
        return g

    def add_arg(parent, complement, kind, name):
        kind = _resolve(m, kind, name)
        if kind == "basic-event":
            idx = var_of[name]
        else:
            if name not in gate_of:
                gate_of[name] = construct(m.gates[name])
            idx = gate_of[name]
        parent[3].append(-idx if complement else idx)

    # ConstructGate (src/pdag.cc:614-676): the parent takes its index before
    # any argument gate. The recursion limit is raised for deep trees.
    def construct(f: Formula) -> int:
        if f.connective == "iff":
            parent = new_gate(NULL)
            inner = new_gate(XOR)
            for a in f.args:
                add_arg(inner, *a)
            parent[3].append(-inner[0])
            return parent[0]
        if f.connective == "imply":
            parent = new_gate(OR)
            (c0, k0, n0), (c1, k1, n1) = f.args
            add_arg(parent, not c0, k0, n0)
            add_arg(parent, c1, k1, n1)
            return parent[0]
        parent = new_gate(_CONNECTIVES[f.connective], f.min_number or 0)
        for a in f.args:
            add_arg(parent, *a)
        return parent[0]

    import sys
    limit = sys.getrecursionlimit()
    sys.setrecursionlimit(max(limit, 4 * len(m.gates) + 1000))
    try:
        root = construct(m.gates[top])
    finally:
        sys.setrecursionlimit(limit)

    names = {v: k for k, v in var_of.items()}
    names.update({v: k for k, v in gate_of.items()})
    names[root] = top
    variables = [(idx, m.basic_events[name]) for name, idx in var_of.items()]
    pd = Pdag.from_lists(variables, [tuple(g) for g in gates], root)
    pd.meta = {"generator": "mef", "top": top, "names": names}
    pd.validate()
    return pd


def load_pdag(paths, top: str | None = None, mission_time: float = 8760.0) -> Pdag:
    """MEF file(s) -> the level-0 PDAG of one top gate."""
    return build_pdag(load_model(paths, mission_time), top)
