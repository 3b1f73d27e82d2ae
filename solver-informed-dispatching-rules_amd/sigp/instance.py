"""STOPTW instances: the reference's readers and SimData staging, plus the
synthetic generators of the throughput workloads.

Readers restate engine/instance.hpp:49-132 (Solomon: score = demand column,
Cordeau: score = 5th column, windows = last two columns, Tmax = depot close).
`Instance` holds what engine/sim_host.hpp:12-40 (SimData) holds: customer
arrays indexed i = 0..n-1 (POI id i+1) and the (n+1)^2 distance matrix with
dist[a][b] = sqrt(dx*dx + dy*dy) (the same IEEE operations, so bit-identical).

The synthetic generators are new (the reference only reads benchmark files).
They draw from CPython's Mersenne Twister (`random.Random`, which the
reference's Mt19937 restates, engine/rng.hpp:13-148) so an instance is a pure
function of its seed.
"""
from __future__ import annotations

import math
import os
import random
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

__all__ = ["POI", "Instance", "parse_solomon", "parse_cordeau", "load_instance",
           "synthetic_uniform", "synthetic_clustered", "synthetic_mixed", "data_dir"]


@dataclass
class POI:
    id: int
    x: float
    y: float
    score: float
    open_time: float
    close_time: float
    duration: float


@dataclass
class Instance:
    """A STOPTW instance staged as the simulator reads it."""

    name: str
    pois: List[POI]
    tmax: float
    m: int = 3
    sigma: float = 0.2
    source: str = ""
    scores: np.ndarray = field(init=False, repr=False)
    durations: np.ndarray = field(init=False, repr=False)
    open_t: np.ndarray = field(init=False, repr=False)
    close_t: np.ndarray = field(init=False, repr=False)
    dist: np.ndarray = field(init=False, repr=False)
    dist_to_depot: np.ndarray = field(init=False, repr=False)

    def __post_init__(self) -> None:
        cust = self.pois[1:]
        f = lambda a: np.ascontiguousarray(np.array(a, dtype=np.float64))
        self.scores = f([p.score for p in cust])
        self.durations = f([p.duration for p in cust])
        self.open_t = f([p.open_time for p in cust])
        self.close_t = f([p.close_time for p in cust])
        cx = f([p.x for p in self.pois])
        cy = f([p.y for p in self.pois])
        dx = cx[:, None] - cx[None, :]
        dy = cy[:, None] - cy[None, :]
        # engine/sim_host.hpp:31-36: sqrt(dx*dx + dy*dy), no fused multiply-add
        self.dist = np.ascontiguousarray(np.sqrt(dx * dx + dy * dy))
        self.dist_to_depot = np.ascontiguousarray(self.dist[1:, 0].copy())

    @property
    def n(self) -> int:
        return len(self.pois) - 1

    @property
    def ld(self) -> int:
        return self.n + 1

    def with_m(self, m: int) -> "Instance":
        return Instance(self.name, self.pois, self.tmax, m, self.sigma, self.source)


def _starts_with_digit(tok: str) -> bool:
    # engine/instance.hpp:34-39: leading '-' characters stripped, rest all digits
    t = tok.lstrip("-")
    return len(t) > 0 and all("0" <= c <= "9" for c in t)


def parse_solomon(path: str, m: int = 3) -> Instance:
    """engine/instance.hpp:49-88."""
    with open(path, "r") as fh:
        lines = [ln.strip(" \t\r\n") for ln in fh]
    lines = [ln for ln in lines if ln]
    name = lines[0].lower()
    cust_idx = 0
    for i, ln in enumerate(lines):
        if "CUST" in ln.upper():
            cust_idx = i
            break
    pois = []
    for ln in lines[cust_idx + 1:]:
        parts = ln.split()
        if len(parts) < 7 or not _starts_with_digit(parts[0]):
            continue
        pid = int(parts[0])
        demand = float(parts[3])
        pois.append(POI(pid, float(parts[1]), float(parts[2]), 0.0 if pid == 0 else demand,
                        float(parts[4]), float(parts[5]), float(parts[6])))
    pois.sort(key=lambda p: p.id)
    return Instance(name, pois, pois[0].close_time, m, 0.2, path)


def parse_cordeau(path: str, m: int = 3) -> Instance:
    """engine/instance.hpp:90-121."""
    stem = os.path.splitext(os.path.basename(path))[0].lower()
    pois = []
    with open(path, "r") as fh:
        for ln in fh:
            parts = ln.split()
            if len(parts) < 9 or not _starts_with_digit(parts[0]):
                continue
            cid = int(parts[0])
            if cid < 0:
                continue
            pois.append(POI(cid, float(parts[1]), float(parts[2]),
                            0.0 if cid == 0 else float(parts[4]),
                            float(parts[-2]), float(parts[-1]), float(parts[3])))
    pois.sort(key=lambda p: p.id)
    return Instance(stem, pois, pois[0].close_time, m, 0.2, path)


def data_dir() -> str:
    """$GITC_DATA (engine/instance.hpp:123-127), else the bundled benchmark files."""
    env = os.environ.get("GITC_DATA")
    if env:
        return env
    here = os.path.dirname(os.path.abspath(__file__))
    return os.path.join(here, "..", "..", "tests", "golden", "instances")


def load_instance(name: str, m: int = 3, root: Optional[str] = None) -> Instance:
    """engine/instance.hpp:129-132: pr* are Cordeau files, the rest Solomon."""
    root = root or data_dir()
    if name.startswith("pr"):
        return parse_cordeau(os.path.join(root, "cordeau", name + ".txt"), m)
    return parse_solomon(os.path.join(root, "solomon", name + ".txt"), m)


# ---------------------------------------------------------------------------
# synthetic instances (new; no reference counterpart)

def _windows(rng: random.Random, cx, cy, tmax, dur, wlo, whi):
    """Customer windows: width ~ U[wlo, whi] centred at c ~ U[d(0,i), tmax - dur - d(i,0)]
    (the latest start that still returns in time; collapsed to d(0,i) when
    that interval is empty), clipped to [0, tmax]."""
    out = []
    for i in range(1, len(cx)):
        d0 = math.sqrt((cx[0] - cx[i]) ** 2 + (cy[0] - cy[i]) ** 2)
        width = wlo + (whi - wlo) * rng.random()
        lo, hi = d0, tmax - dur - d0
        if hi < lo:
            hi = lo
        c = lo + (hi - lo) * rng.random()
        o = max(0.0, c - width / 2.0)
        e = min(tmax, c + width / 2.0)
        out.append((o, e))
    return out


def _assemble(name, cx, cy, wins, scores, tmax, dur, m) -> Instance:
    pois = [POI(0, float(cx[0]), float(cy[0]), 0.0, 0.0, float(tmax), 0.0)]
    for i in range(1, len(cx)):
        o, e = wins[i - 1]
        pois.append(POI(i, float(cx[i]), float(cy[i]), float(scores[i - 1]), o, e, float(dur)))
    return Instance(name, pois, float(tmax), m, 0.2, "synthetic")


def synthetic_uniform(n: int = 100, m: int = 2, seed: int = 1, tmax: float = 230.0,
                      width=(10.0, 60.0), duration: float = 10.0, score_max: int = 50) -> Instance:
    """n customers + depot with integer coordinates U{0..100}^2 (depot drawn
    first), integer scores U{1..score_max}, expected service `duration`."""
    rng = random.Random(seed)
    cx, cy = [], []
    for _ in range(n + 1):
        cx.append(rng.randint(0, 100))
        cy.append(rng.randint(0, 100))
    wins = _windows(rng, cx, cy, tmax, duration, *width)
    scores = [rng.randint(1, score_max) for _ in range(n)]
    return _assemble("syn-u%d-s%d" % (n, seed), cx, cy, wins, scores, tmax, duration, m)


def _cluster_points(rng: random.Random, count: int, n_clusters: int, sigma: float):
    centres = [(100.0 * rng.random(), 100.0 * rng.random()) for _ in range(n_clusters)]
    pts = []
    for _ in range(count):
        cxy = centres[rng.randrange(n_clusters)]
        x = min(100, max(0, int(round(rng.gauss(cxy[0], sigma)))))
        y = min(100, max(0, int(round(rng.gauss(cxy[1], sigma)))))
        pts.append((x, y))
    return pts


def synthetic_clustered(n: int = 100, m: int = 3, seed: int = 1, tmax: float = 230.0,
                        width=(30.0, 120.0), n_clusters: int = 10, sigma: float = 5.0,
                        duration: float = 10.0, score_max: int = 50) -> Instance:
    """n customers in `n_clusters` Gaussian clusters (std `sigma`, centres
    U[0,100]^2, coordinates rounded and clipped to {0..100}); depot uniform."""
    rng = random.Random(seed)
    depot = (rng.randint(0, 100), rng.randint(0, 100))
    pts = _cluster_points(rng, n, n_clusters, sigma)
    cx = [depot[0]] + [p[0] for p in pts]
    cy = [depot[1]] + [p[1] for p in pts]
    wins = _windows(rng, cx, cy, tmax, duration, *width)
    scores = [rng.randint(1, score_max) for _ in range(n)]
    return _assemble("syn-c%d-s%d" % (n, seed), cx, cy, wins, scores, tmax, duration, m)


def synthetic_mixed(n: int = 288, m: int = 4, seed: int = 1, tmax: float = 230.0,
                    width=(20.0, 200.0), n_clusters: int = 10, sigma: float = 5.0,
                    duration: float = 10.0, score_max: int = 50) -> Instance:
    """Half the customers clustered, half uniform (clustered + uniform mix)."""
    rng = random.Random(seed)
    depot = (rng.randint(0, 100), rng.randint(0, 100))
    nc = n // 2
    pts = _cluster_points(rng, nc, n_clusters, sigma)
    pts += [(rng.randint(0, 100), rng.randint(0, 100)) for _ in range(n - nc)]
    cx = [depot[0]] + [p[0] for p in pts]
    cy = [depot[1]] + [p[1] for p in pts]
    wins = _windows(rng, cx, cy, tmax, duration, *width)
    scores = [rng.randint(1, score_max) for _ in range(n)]
    return _assemble("syn-m%d-s%d" % (n, seed), cx, cy, wins, scores, tmax, duration, m)
