"""Readers for the reference's golden fixtures (tests/golden/, copied as data
from the reference's verify/golden/)."""
import gzip
import os
import struct

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def h2d(h: str) -> float:
    return struct.unpack("<d", bytes.fromhex(h)[::-1])[0]


def d2h(x: float) -> str:
    return struct.pack("<d", x)[::-1].hex()


def bits(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64)).view(np.uint64)


def sim_cases():
    """verify/golden/sim.txt: 12 cases with tree, mask, zp, sampled rows, rewards."""
    cases, cur = [], None
    with open(os.path.join(GOLDEN, "sim.txt")) as fh:
        for ln in fh:
            ln = ln.rstrip("\r\n")
            if not ln:
                continue
            if ln.startswith("# case "):
                inst, kind = ln[7:].split()
                cur = dict(inst=inst, kind=kind, sampled=[])
                cases.append(cur)
            elif ln.startswith("tree "):
                cur["tree"] = ln[5:]
            elif ln.startswith("mask "):
                cur["mask"] = np.array([int(v) for v in ln[5:].split()], dtype=np.int64)
            elif ln.startswith("zp "):
                cur["zp"] = np.array([h2d(v) for v in ln[3:].split()])
            elif ln.startswith("sampled "):
                cur["sampled"].append([h2d(v) for v in ln[8:].split()])
            elif ln.startswith("scores "):
                cur["want"] = np.array([h2d(v) for v in ln[7:].split()])
    for c in cases:
        c["sampled"] = np.array(c["sampled"])
    return cases


def gp_cases():
    """verify/golden/gp.txt: per case, 3 generations of (population prefix
    strings, sampled row, fitness hex)."""
    cases, cur = [], None
    with gzip.open(os.path.join(GOLDEN, "gp.txt.gz"), "rt") as fh:
        for ln in fh:
            ln = ln.rstrip("\r\n")
            if not ln:
                continue
            if ln.startswith("# case "):
                cur = dict(kind=ln[ln.rfind(" ") + 1:], gens=[])
                cases.append(cur)
                cur["gens"].append(dict(pop=[]))
            elif ln[0] == "p" and len(ln) > 2 and ln[2] == " ":
                if "fits" in cur["gens"][-1]:
                    cur["gens"].append(dict(pop=[]))
                cur["gens"][-1]["pop"].append(ln[3:])
            elif ln.startswith("sampled "):
                parts = ln[8:].split()
                cur["gens"][-1]["sampled"] = np.array([h2d(v) for v in parts[1:]])
            elif ln.startswith("fits "):
                parts = ln[5:].split()
                cur["gens"][-1]["fits"] = np.array([h2d(v) for v in parts[1:]])
    return cases


def compile_cases():
    out, cur = [], None
    with open(os.path.join(GOLDEN, "compile.txt")) as fh:
        for ln in fh:
            ln = ln.rstrip("\r\n")
            if ln.startswith("tree "):
                cur = dict(tree=ln[5:])
                out.append(cur)
            elif ln.startswith("ops "):
                cur["ops"] = [int(v) for v in ln[4:].split()]
            elif ln.startswith("args "):
                cur["args"] = [int(v) for v in ln[5:].split()]
    return out
