"""The GP loop on the device (sigp_gp_* of include/sigp.h).

Replaces the host loop of the reference's generation engine: evolve_one
(engine/gp_engine.cu:66-84) for one seed, exhaust_batched
(engine/gp_engine.cu:131-183) for many seeds evolved side by side.  The
populations, their breeding (engine/gp.hpp:67-194, every draw from a
CPython-compatible Mersenne Twister per run, in the reference's order) and
their evaluation stay on the device; `run` executes a whole loop with one
generation captured as a HIP graph.

    gp = DeviceGP(ctx, TreeConfig(vocab), EvolutionConfig(), seeds=[0], T=50,
                  term_mask=term_mask_from(vocab))
    gp.init()
    res = gp.run(all_sampled, generations=50)   # fitness history, best rules, times
    final = gp.population()

With `offspring` = k x population and a trained surrogate network
(`set_surrogate`), each generation breeds k gp_steps (step s on the stream
Mt19937(seed + s * 2**32)), and the network's phenotype scores choose the
next parents among them: the elites, then the best-scored offspring.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from typing import List, Optional, Sequence, Tuple

import numpy as np

from ._lib import GP_MAX_NODES, GP_TOK_CONST, GP_TOK_FUNC, MAP_AUTO, GpConfig, LaunchConfig, check, lib
from ._lib import Situations as _Situations
from .api import FactoryContext, _masks, _ptr
from .gp import EvolutionConfig, TreeConfig
from .rules import _OPCODE, Node, Packed, term_index, to_prefix

__all__ = ["DeviceGP", "RunResult", "tokens_to_tree", "SUB_STREAM"]

SUB_STREAM = 2 ** 32   # sub-step s of a screened run draws from Mt19937(seed + s * SUB_STREAM)


def tokens_to_tree(tokens: Sequence[int], terminals: Sequence[str], functions: Sequence[str]) -> Node:
    """A device tree (prefix-order tokens) as a Node: a terminal token is an
    index into `terminals`, GP_TOK_FUNC | f a function, GP_TOK_CONST | k the
    ERC constant k / 1000 labelled "%.3f" as the reference labels it."""
    pos = 0

    def at() -> Node:
        nonlocal pos
        t = int(tokens[pos])
        pos += 1
        if t & GP_TOK_FUNC:
            a = at()
            b = at()
            return Node(functions[t & 0xFF], (a, b))
        if t & GP_TOK_CONST:
            return Node("%.3f" % ((t & (GP_TOK_CONST - 1)) / 1000.0))
        return Node(terminals[t])

    tree = at()
    if pos != len(tokens):
        raise ValueError("token array is not one tree")
    return tree


@dataclass
class RunResult:
    fitness: np.ndarray          # (generations, runs, pop)
    best: List[List[Node]]       # [generation][run]: the best rule of each evaluation
    generation_ms: np.ndarray    # device time from the start to the end of each generation

    @property
    def ms(self) -> float:
        return float(self.generation_ms[-1])


def _situations(sit):
    """The C struct over `sit`'s arrays, and the arrays it points into."""
    arrs = [np.ascontiguousarray(sit.cur, dtype=np.int32), np.ascontiguousarray(sit.ct, dtype=np.float64),
            np.ascontiguousarray(sit.remaining, dtype=np.int32),
            np.ascontiguousarray(sit.visited, dtype=np.uint8), np.ascontiguousarray(sit.cand, dtype=np.int32)]
    return _Situations(sit.n_sit, sit.n_cand, *[a.ctypes.data for a in arrs]), arrs


class DeviceGP:
    """Populations of `len(seeds)` runs on `ctx`'s instance; run r uses
    Mt19937(seeds[r]) and reads scenario rows r*T .. r*T+T-1 of each
    generation (the reference's grouped evaluation, group_size = population)."""

    def __init__(self, ctx: FactoryContext, tree_cfg: TreeConfig, evo_cfg: EvolutionConfig,
                 seeds: Sequence[int], T: int, term_mask, waves_per_item: int = 0, mapping: int = MAP_AUTO,
                 offspring: int = 0):
        self.ctx = ctx
        self.tree_cfg = tree_cfg
        self.evo_cfg = evo_cfg
        self.runs = len(seeds)
        self.pop = evo_cfg.population_size
        self.T = int(T)
        self.n = ctx.inst.n
        slots = []
        for t in tree_cfg.terminals:
            if t == "ERC":
                slots.append(-1)
            else:
                s = term_index(t)
                if s < 0:
                    raise ValueError("unknown terminal %r" % t)
                slots.append(s)
        self._slots = np.ascontiguousarray(slots, dtype=np.int32)
        self._fops = np.ascontiguousarray([_OPCODE[f] for f in tree_cfg.functions], dtype=np.int32)
        cfg = GpConfig()
        cfg.n_runs = self.runs
        cfg.population = self.pop
        cfg.offspring = int(offspring)
        self.offspring = int(offspring) or self.pop
        cfg.n_terminals = len(slots)
        cfg.terminal_slots = self._slots.ctypes.data
        cfg.n_functions = len(self._fops)
        cfg.function_ops = self._fops.ctypes.data
        cfg.min_depth = tree_cfg.min_depth
        cfg.max_depth = tree_cfg.max_depth
        cfg.depth_limit = evo_cfg.max_depth
        cfg.tournament_size = evo_cfg.tournament_size
        cfg.elitism = evo_cfg.elitism
        cfg.crossover_rate = evo_cfg.crossover_rate
        cfg.mutation_rate = evo_cfg.mutation_rate
        cfg.T = self.T
        cfg.launch = LaunchConfig(waves_per_item, mapping)
        seeds_a = np.ascontiguousarray([int(s) for s in seeds], dtype=np.uint64)
        mask = _masks(term_mask)
        if mask.shape[0] != 1:
            raise ValueError("the GP loop evaluates with one terminal mask")
        self._h = C.c_void_p()
        check(lib().sigp_gp_create(ctx._h, C.byref(cfg), _ptr(seeds_a), _ptr(mask), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().sigp_gp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_surrogate(self, net, sit=None) -> None:
        """Screen offspring with `net` (sigp.surrogate.Mlp, kept alive here)
        on the decision situations `sit` (64 x 8).  net = None detaches the
        screen: the next steps are plain gp_steps (sub-step 0's offspring),
        as they are before a network is first attached."""
        if net is None:
            check(lib().sigp_gp_set_surrogate(self._h, None, None))
            self._net = None
            return
        s, keep = _situations(sit)
        check(lib().sigp_gp_set_surrogate(self._h, net._h, C.byref(s)))
        self._net = net

    def phenotype(self, sit) -> np.ndarray:
        """int32 ranks (runs * pop, n_sit, n_cand) of the current population
        on the decision situations `sit`, as sigp.surrogate.phenotype of the
        same trees."""
        s, keep = _situations(sit)
        out = np.zeros((self.runs * self.pop, sit.n_sit, sit.n_cand), dtype=np.int32)
        check(lib().sigp_gp_phenotype(self._h, C.byref(s), _ptr(out)))
        return out

    def set_sampler(self, run_seeds: Sequence[int], mode: str = "gamma", sigma: float = 0.2, shape: float = 4.0,
                    gen_stride: int = 1, gen_offset: int = 0) -> None:
        """Sample each generation's rows on the device inside the loop:
        row t of run r in generation g draws from
        Mt19937(run_seeds[r] + (gen_offset + g) * gen_stride + t); `run`
        then takes sampled_all = None."""
        from .api import sampler
        smp = sampler(mode, sigma, shape)
        seeds = np.ascontiguousarray([int(v) for v in run_seeds], dtype=np.int64)
        if seeds.size != self.runs:
            raise ValueError("need one seed per run")
        check(lib().sigp_gp_set_sampler(self._h, C.byref(smp), _ptr(seeds), int(gen_stride), int(gen_offset)))

    def init(self, screened: bool = False) -> None:
        """ramped_half_and_half of every run (engine/gp.hpp:67-77); with
        screened=True (a screen attached), k such populations per run, one
        per stream, of which the screen keeps the population
        (sigp_gp_init_screened)."""
        check(lib().sigp_gp_init_screened(self._h) if screened else lib().sigp_gp_init(self._h))

    def _rows(self, sampled, rows: int) -> np.ndarray:
        a = np.ascontiguousarray(sampled, dtype=np.float64).reshape(-1)
        if a.size != rows * self.n:
            raise ValueError("need %d scenario rows of %d durations" % (rows, self.n))
        return a

    def evaluate(self, sampled) -> np.ndarray:
        """Mean reward of every rule, (runs, pop): run r reads rows r*T..r*T+T-1."""
        rows = self.runs * self.T
        a = self._rows(sampled, rows)
        fit = np.zeros(self.runs * self.pop, dtype=np.float64)
        check(lib().sigp_gp_evaluate(self._h, _ptr(a), rows, _ptr(fit)))
        return fit.reshape(self.runs, self.pop)

    def step(self, fitness=None) -> None:
        """gp_step of every run (engine/gp.hpp:170-194) with `fitness`
        (runs x pop), or the last evaluation's."""
        if fitness is None:
            check(lib().sigp_gp_step(self._h, None))
            return
        f = np.ascontiguousarray(fitness, dtype=np.float64).reshape(-1)
        if f.size != self.runs * self.pop:
            raise ValueError("need %d fitness values" % (self.runs * self.pop))
        check(lib().sigp_gp_step(self._h, _ptr(f)))

    def run(self, sampled_all, generations: int, graph: bool = True, step_last: bool = False) -> RunResult:
        """The whole loop on the device: generation g evaluates on rows
        g*runs*T .. and steps, except the last unless `step_last` (a loop
        run in several calls).  Generations are numbered from this call."""
        if sampled_all is None:
            rows, a = 0, None
        else:
            rows = generations * self.runs * self.T
            a = self._rows(sampled_all, rows)
        hist = np.zeros(generations * self.runs * self.pop, dtype=np.float64)
        btok = np.zeros(generations * self.runs * GP_MAX_NODES, dtype=np.int16)
        blen = np.zeros(generations * self.runs, dtype=np.int32)
        gms = np.zeros(generations, dtype=np.float64)
        flags = (1 if graph else 0) | (2 if step_last else 0)
        check(lib().sigp_gp_run(self._h, _ptr(a), rows, generations, flags, _ptr(hist), _ptr(btok),
                                _ptr(blen), _ptr(gms)))
        btok = btok.reshape(generations, self.runs, GP_MAX_NODES)
        blen = blen.reshape(generations, self.runs)
        best = [[tokens_to_tree(btok[g, r, :blen[g, r]], self.tree_cfg.terminals, self.tree_cfg.functions)
                 for r in range(self.runs)] for g in range(generations)]
        return RunResult(hist.reshape(generations, self.runs, self.pop), best, gms)

    def tokens(self) -> Tuple[np.ndarray, np.ndarray]:
        """(tokens (runs, pop, GP_MAX_NODES) int16, lengths (runs, pop))."""
        tok = np.zeros(self.runs * self.pop * GP_MAX_NODES, dtype=np.int16)
        ln = np.zeros(self.runs * self.pop, dtype=np.int32)
        check(lib().sigp_gp_population(self._h, _ptr(tok), _ptr(ln)))
        return tok.reshape(self.runs, self.pop, GP_MAX_NODES), ln.reshape(self.runs, self.pop)

    def population(self, run: Optional[int] = None) -> List:
        """Trees of one run, or a list per run."""
        tok, ln = self.tokens()
        conv = lambda r: [tokens_to_tree(tok[r, i, :ln[r, i]], self.tree_cfg.terminals, self.tree_cfg.functions)
                          for i in range(self.pop)]
        return conv(run) if run is not None else [conv(r) for r in range(self.runs)]

    def prefixes(self, run: int = 0) -> List[str]:
        return [to_prefix(t) for t in self.population(run)]

    def packed(self) -> Packed:
        """The population compiled by the device, as pack_trees lays it out."""
        P = self.runs * self.pop
        ops = np.zeros(P * GP_MAX_NODES, dtype=np.int32)
        args = np.zeros(2 * P * GP_MAX_NODES, dtype=np.int32)
        ln = np.zeros(P, dtype=np.int32)
        check(lib().sigp_gp_packed(self._h, _ptr(ops), _ptr(args), _ptr(ln)))
        o_ops, o_args, offs = [], [], []
        k = 0
        for i in range(P):
            L = int(ln[i])
            offs.append(k)
            o_ops.append(ops[i * GP_MAX_NODES:i * GP_MAX_NODES + L])
            o_args.append(args[2 * i * GP_MAX_NODES:2 * (i * GP_MAX_NODES + L)])
            k += L
        return Packed(np.ascontiguousarray(np.concatenate(o_ops)), np.ascontiguousarray(np.concatenate(o_args)),
                      np.array(offs, dtype=np.int32), np.ascontiguousarray(ln))

    def rng_state(self, run: int = 0, sub: int = 0) -> Tuple[int, ...]:
        """As random.Random.getstate()[1]: 624 words and the position."""
        st = np.zeros(625, dtype=np.uint32)
        check(lib().sigp_gp_rng_state(self._h, run, sub, _ptr(st)))
        return tuple(int(v) for v in st)
