"""Surrogate pre-screening of rules (no counterpart in the reference).

A rule's phenotype is its ranking of candidate customers in a fixed set of
seeded decision situations (64 situations x 8 candidates = 512 integer
ranks, computed on the device by sigp_phenotype).  A 512 -> 256 -> 256 -> 1
MLP (sigp_mlp_*, BF16 matrix cores, FP32 accumulation) learns the mapping
from phenotype to the rule's mean simulated reward.  The screen scores many
candidate rules by the network and sends only the best-scored to exact
simulation; the simulator stays both the fallback and the judge of the
network's error (Spearman rank correlation and mean absolute error against
exact rewards).
"""
from __future__ import annotations

import ctypes as C
import random
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from ._lib import Situations as _Situations
from ._lib import check, lib
from .api import FactoryContext, _masks, _ptr, _rules
from .instance import Instance
from .rules import Packed

N_SIT, N_CAND = 64, 8
N_FEATURES = N_SIT * N_CAND          # 512
N_PARAMS = 197377

__all__ = ["DecisionSituations", "make_situations", "phenotype", "features", "Mlp", "spearman",
           "N_SIT", "N_CAND", "N_FEATURES", "N_PARAMS"]


@dataclass
class DecisionSituations:
    cur: np.ndarray        # int32 [S]   current POI (0 = depot)
    ct: np.ndarray         # float64 [S] current time
    remaining: np.ndarray  # int32 [S]   vehicles after the current one
    visited: np.ndarray    # uint8 [S, n]
    cand: np.ndarray       # int32 [S, C] customers to rank, ascending

    @property
    def n_sit(self) -> int:
        return int(self.cand.shape[0])

    @property
    def n_cand(self) -> int:
        return int(self.cand.shape[1])


def _feasible(inst: Instance, cur: int, ct: float, visited) -> list:
    """pick_best's candidate test (engine/sim_core.hpp:76-82)."""
    out = []
    D = np.asarray(inst.dist).reshape(-1)
    for i in range(inst.n):
        if visited[i]:
            continue
        arr = ct + D[cur * (inst.n + 1) + i + 1]
        if arr > inst.close_t[i]:
            continue
        st = arr if arr > inst.open_t[i] else inst.open_t[i]
        if st + inst.durations[i] + inst.dist_to_depot[i] > inst.tmax:
            continue
        out.append(i)
    return out


def make_situations(inst: Instance, n_sit: int = N_SIT, n_cand: int = N_CAND,
                    seed: int = 20240611, max_steps: int = 12) -> DecisionSituations:
    """Seeded decision situations: a partial route is built on expected
    service times by picking uniformly among the feasible customers for a
    random number of steps (vehicles close when nothing is feasible); the
    state is kept when at least n_cand customers are feasible, and n_cand of
    them, chosen at random and sorted, are the candidates."""
    rng = random.Random(seed)
    n = inst.n
    cur_l, ct_l, rem_l, vis_l, cand_l = [], [], [], [], []
    attempts = 0
    while len(cur_l) < n_sit:
        attempts += 1
        if attempts > 200 * n_sit:
            raise ValueError("instance has too few feasible customers for %d candidates" % n_cand)
        steps = rng.randint(0, max_steps)
        visited = [0] * n
        cur, ct, veh = 0, 0.0, 0
        for _ in range(steps):
            feas = _feasible(inst, cur, ct, visited)
            if not feas:
                veh += 1
                if veh >= inst.m:
                    break
                cur, ct = 0, 0.0
                continue
            i = feas[rng.randrange(len(feas))]
            arr = ct + float(np.asarray(inst.dist).reshape(-1)[cur * (n + 1) + i + 1])
            st = arr if arr > inst.open_t[i] else inst.open_t[i]
            ct = st + inst.durations[i]
            cur = i + 1
            visited[i] = 1
        if veh >= inst.m:
            continue
        feas = _feasible(inst, cur, ct, visited)
        if len(feas) < n_cand:
            continue
        cur_l.append(cur)
        ct_l.append(ct)
        rem_l.append(inst.m - veh - 1)
        vis_l.append(visited)
        cand_l.append(sorted(rng.sample(feas, n_cand)))
    return DecisionSituations(np.array(cur_l, dtype=np.int32), np.array(ct_l, dtype=np.float64),
                              np.array(rem_l, dtype=np.int32), np.array(vis_l, dtype=np.uint8),
                              np.array(cand_l, dtype=np.int32))


def phenotype(ctx: FactoryContext, pk: Packed, masks, sit: DecisionSituations,
              mask_group: Optional[int] = None) -> np.ndarray:
    """sigp_phenotype: int32 [P, n_sit, n_cand] ranks (0 = chosen first)."""
    keep: list = []
    r = _rules(pk, keep)
    m = _masks(masks)
    arrs = [np.ascontiguousarray(sit.cur, dtype=np.int32), np.ascontiguousarray(sit.ct, dtype=np.float64),
            np.ascontiguousarray(sit.remaining, dtype=np.int32),
            np.ascontiguousarray(sit.visited, dtype=np.uint8), np.ascontiguousarray(sit.cand, dtype=np.int32)]
    s = _Situations(sit.n_sit, sit.n_cand, *[a.ctypes.data for a in arrs])
    P = pk.n_rules
    out = np.zeros((P, sit.n_sit, sit.n_cand), dtype=np.int32)
    check(lib().sigp_phenotype(ctx._h, C.byref(r), _ptr(m), m.shape[0], mask_group or max(P, 1),
                               C.byref(s), out.ctypes.data_as(C.c_void_p)))
    return out


def features(ranks: np.ndarray) -> np.ndarray:
    """float32 [P, n_sit * n_cand]: each rank scaled to [0, 1] (0 = chosen first)."""
    R = np.asarray(ranks, dtype=np.float32)
    return np.ascontiguousarray(R.reshape(R.shape[0], -1) / max(R.shape[-1] - 1, 1))


class Mlp:
    """The surrogate network on the device (sigp_mlp_*)."""

    def __init__(self, seed: int = 1):
        self._h = C.c_void_p()
        check(lib().sigp_mlp_create(C.c_uint64(seed), C.byref(self._h)))

    def close(self) -> None:
        if self._h:
            lib().sigp_mlp_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(N_PARAMS, dtype=np.float32)
        check(lib().sigp_mlp_get_params(self._h, _ptr(out)))
        return out

    def set_params(self, p) -> None:
        a = np.ascontiguousarray(p, dtype=np.float32)
        if a.size != N_PARAMS:
            raise ValueError("need %d parameters" % N_PARAMS)
        check(lib().sigp_mlp_set_params(self._h, _ptr(a)))

    def gradients(self, x, t):
        xa = np.ascontiguousarray(x, dtype=np.float32)
        ta = np.ascontiguousarray(t, dtype=np.float32)
        g = np.zeros(N_PARAMS, dtype=np.float32)
        loss = C.c_float(0.0)
        check(lib().sigp_mlp_gradients(self._h, _ptr(xa), _ptr(ta), xa.shape[0], _ptr(g), C.byref(loss)))
        return g, loss.value

    def train(self, x, t, epochs: int = 10, batch: int = 256, lr: float = 1e-3,
              seed: int = 7) -> np.ndarray:
        xa = np.ascontiguousarray(x, dtype=np.float32)
        ta = np.ascontiguousarray(t, dtype=np.float32)
        losses = np.zeros(max(epochs, 1), dtype=np.float32)
        check(lib().sigp_mlp_train(self._h, _ptr(xa), _ptr(ta), xa.shape[0], epochs, batch,
                                   C.c_float(lr), C.c_uint64(seed), _ptr(losses)))
        return losses[:epochs]

    def predict(self, x) -> np.ndarray:
        xa = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros(xa.shape[0], dtype=np.float32)
        check(lib().sigp_mlp_predict(self._h, _ptr(xa), xa.shape[0], _ptr(y)))
        return y


def _ranks(a: np.ndarray) -> np.ndarray:
    """Average ranks (ties share the mean rank)."""
    a = np.asarray(a, dtype=np.float64)
    order = np.argsort(a, kind="mergesort")
    ranks = np.empty(len(a), dtype=np.float64)
    sa = a[order]
    i = 0
    while i < len(a):
        j = i
        while j + 1 < len(a) and sa[j + 1] == sa[i]:
            j += 1
        ranks[order[i:j + 1]] = 0.5 * (i + j)
        i = j + 1
    return ranks


def spearman(a: Sequence[float], b: Sequence[float]) -> float:
    """Spearman rank correlation (Pearson correlation of average ranks)."""
    ra, rb = _ranks(np.asarray(a)), _ranks(np.asarray(b))
    ra -= ra.mean()
    rb -= rb.mean()
    den = np.sqrt((ra * ra).sum() * (rb * rb).sum())
    return float((ra * rb).sum() / den) if den > 0 else 0.0
