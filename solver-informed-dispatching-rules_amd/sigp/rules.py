"""Rule representation: GP expression trees, their prefix text, the postfix
code the evaluator runs, and the terminal vocabulary.

Restates engine/tree.hpp:14-96 (nodes, prefix parse/print, depth/size),
engine/vocab.hpp:16-125 (slot map, aliases, NS-11 and the 18-terminal menu,
compile_rmod) and engine/factory.cuh:20-40 (PackedTrees / pack_trees).  The
postfix format is the reference's: ops -1 terminal (args = slot, 0),
-2 constant (args = llround(v * 1000), 0), 0..5 = + - * / min max,
6 = rmod (args = left, right child indices).
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Iterable, List, Sequence, Tuple

import numpy as np

__all__ = ["N_TERMS", "MAX_N", "MAX_STACK", "TERM_MAP", "TERM_ALIASES", "SLOT_NAMES",
           "NS_TERMINALS", "EXTRA", "FUNCTIONS", "FUNCS_RMOD", "Node", "parse_prefix",
           "to_prefix", "tree_size", "tree_depth", "term_index", "term_mask_from",
           "compile_tree", "Packed", "pack_trees", "pack_compiled", "stack_depth"]

N_TERMS = 61      # engine/vocab.hpp:16
MAX_N = 512       # engine/vocab.hpp:17
MAX_STACK = 32    # engine/vocab.hpp:19

SLOT_NAMES = [
    "SCORE", "DUR", "TO", "TC", "TA", "TR", "TSV", "TFV", "SL", "RemT", "NS", "REGRET",
    "NSROB", "MAXN", "REACH2", "NSTW", "RAND", "FRAGCNT", "FRAGSCORE", "DOWNSLACK", "PBUST",
    "REACH1", "REACH3", "REACH5", "SAT", "VFRAC", "NFEAS", "VEH", "HARVEST", "BOLSAO",
    "ATRISK", "ZTC", "ZREACH", "ZWLOSS", "ZATRISK", "ZD2D", "C1", "C2", "C3", "C4", "C5",
    "REACH2ROB", "RELCOST", "NETREACH", "SKIPLOSS", "FRAGILITY", "SROB", "REACH3ROB",
    "REACH5ROB", "NSA2", "NSA05", "FEASCORE", "NFEASROB", "NSROB05", "NSROB2", "NNFWD",
    "NNK3", "NNCHAIN3", "SAVINGS", "DETOUR", "SCORERANK"]
TERM_MAP = {name: i for i, name in enumerate(SLOT_NAMES)}      # engine/vocab.hpp:21-55
TERM_ALIASES = {"D2D": 5, "SLACK": 8, "WLOSS": 11}             # engine/vocab.hpp:57-62
assert len(SLOT_NAMES) == N_TERMS

NS_TERMINALS = ["SCORE", "DUR", "TO", "TC", "TA", "TR", "TSV", "TFV", "SL", "RemT", "NS"]
EXTRA = ["REGRET", "NSROB", "MAXN", "REACH2", "NSTW", "FRAGCNT", "FRAGSCORE", "PBUST",
         "REACH1", "REACH3", "REACH5", "SAT", "VFRAC", "NFEAS", "HARVEST", "BOLSAO",
         "ATRISK", "REACH2ROB"]                                  # engine/vocab.hpp:68-71
FUNCTIONS = ["+", "-", "*", "/", "min", "max"]                   # engine/tree.hpp:33
FUNCS_RMOD = ["+", "-", "*", "/", "min", "max", "rmod"]          # engine/tree.hpp:34
_OPCODE = {"+": 0, "-": 1, "*": 2, "/": 3, "min": 4, "max": 5, "rmod": 6}


class Node:
    """Immutable expression-tree node; subtrees may be shared."""

    __slots__ = ("label", "children")

    def __init__(self, label: str, children: Tuple["Node", ...] = ()):
        self.label = label
        self.children = tuple(children)

    @property
    def is_terminal(self) -> bool:
        return not self.children

    def __repr__(self) -> str:
        return "Node(%s)" % to_prefix(self)


def to_prefix(n: Node) -> str:
    """engine/tree.hpp:62-67."""
    if n.is_terminal:
        return n.label
    return "(" + n.label + "".join(" " + to_prefix(c) for c in n.children) + ")"


def parse_prefix(s: str) -> Node:
    """engine/tree.hpp:69-96."""
    toks = s.replace("(", " ( ").replace(")", " ) ").split()
    pos = 0

    def at() -> Node:
        nonlocal pos
        tok = toks[pos]
        if tok == "(":
            pos += 1
            sym = toks[pos]
            pos += 1
            ch = []
            while toks[pos] != ")":
                ch.append(at())
            pos += 1
            return Node(sym, tuple(ch))
        pos += 1
        return Node(tok)

    return at()


def tree_size(n: Node) -> int:
    return 1 + sum(tree_size(c) for c in n.children)


def tree_depth(n: Node) -> int:
    """engine/tree.hpp:55-60: a leaf has depth 0."""
    if n.is_terminal:
        return 0
    return 1 + max(tree_depth(c) for c in n.children)


def term_index(label: str) -> int:
    """engine/vocab.hpp:73-81; -1 for non-terminals (numeric constants)."""
    if label in TERM_MAP:
        return TERM_MAP[label]
    return TERM_ALIASES.get(label, -1)


def term_mask_from(names: Iterable[str]) -> np.ndarray:
    """engine/sim_host.hpp:85-94: 0/1 int64 mask over the 61 slots; ERC ignored."""
    m = np.zeros(N_TERMS, dtype=np.int64)
    for nm in names:
        if nm == "ERC":
            continue
        ti = term_index(nm)
        if ti < 0:
            raise ValueError("unknown terminal " + nm)
        m[ti] = 1
    return m


def _llround(x: float) -> int:
    """C llround: nearest integer, halves away from zero."""
    a = abs(x)
    r = math.floor(a)
    if a - r >= 0.5:
        r += 1
    return int(r) if x >= 0 else -int(r)


def _strtod(s: str) -> float:
    """Leading numeric prefix of s as strtod reads it (0.0 when none)."""
    best = 0.0
    for k in range(len(s), 0, -1):
        try:
            best = float(s[:k])
            return best
        except ValueError:
            continue
    return best


def compile_tree(tree: Node) -> Tuple[List[int], List[Tuple[int, int]]]:
    """engine/vocab.hpp:101-125 (compile_visit / compile_rmod): post-order."""
    ops: List[int] = []
    args: List[Tuple[int, int]] = []

    def visit(nd: Node) -> None:
        if nd.is_terminal:
            ti = term_index(nd.label)
            if ti >= 0:
                ops.append(-1)
                args.append((ti, 0))
            else:
                ops.append(-2)
                args.append((_llround(_strtod(nd.label) * 1000.0), 0))
            return
        visit(nd.children[0])
        left = len(ops) - 1
        visit(nd.children[1])
        right = len(ops) - 1
        if nd.label not in _OPCODE:
            raise ValueError("unknown operator " + nd.label)
        ops.append(_OPCODE[nd.label])
        args.append((left, right))

    visit(tree)
    return ops, args


def stack_depth(ops: Sequence[int]) -> int:
    """Maximum postfix stack depth of compiled code."""
    d = mx = 0
    for o in ops:
        d = d + 1 if o < 0 else d - 1
        mx = max(mx, d)
    return mx


@dataclass
class Packed:
    """engine/factory.cuh:20-23 (PackedTrees): int32 arrays."""

    ops: np.ndarray
    args: np.ndarray
    offsets: np.ndarray
    lengths: np.ndarray

    @property
    def n_rules(self) -> int:
        return int(self.offsets.shape[0])

    @property
    def n_nodes(self) -> int:
        return int(self.ops.shape[0])

    def rule(self, p: int) -> Tuple[np.ndarray, np.ndarray]:
        o, L = int(self.offsets[p]), int(self.lengths[p])
        return self.ops[o:o + L], self.args[2 * o:2 * (o + L)]


def pack_compiled(compiled: Sequence[Tuple[List[int], List[Tuple[int, int]]]]) -> Packed:
    ops: List[int] = []
    args: List[int] = []
    offs: List[int] = []
    lens: List[int] = []
    for o, a in compiled:
        offs.append(len(ops))
        lens.append(len(o))
        ops.extend(o)
        for x, y in a:
            args.append(x)
            args.append(y)
    i32 = lambda v: np.ascontiguousarray(np.array(v, dtype=np.int32))
    return Packed(i32(ops), i32(args), i32(offs), i32(lens))


def pack_trees(trees: Sequence[Node]) -> Packed:
    """engine/factory.cuh:25-40."""
    return pack_compiled([compile_tree(t) for t in trees])
