"""Build libsigp.so in-tree with hipcc for gfx950.

Flags: -ffp-contract=off keeps every floating-point operation unfused (the
simulator must reproduce the reference's CPU arithmetic bit for bit; the few
fused multiply-adds it needs are explicit); packed FP32 instructions are
disabled for the device pass.

Each source compiles to its own object (in parallel; kernels do not call
across files), then one link.  `incremental=True` (command line:
--incremental) recompiles only the objects older than their source or any
header; the default rebuilds everything.
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "csrc")
INCLUDE = os.path.join(HERE, "..", "..", "include")
OBJ = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "libsigp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall",
         "-Wno-unused-function"]
SOURCES = ["sigp_sim.hip", "sigp_mlp.hip", "sigp_gp.hip", "sigp_sample.hip"]
NOISE = "packed-fp32-ops' is not a recognized feature"


def _run(cmd):
    res = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(ln for ln in (res.stderr or "").splitlines() if NOISE not in ln)
    if res.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + " ".join(cmd) + "\n" + err)
    return err


def _stale(src, obj, headers):
    if not os.path.exists(obj):
        return True
    t = os.path.getmtime(obj)
    return any(os.path.getmtime(p) > t for p in [src] + headers)


def build(verbose: bool = False, incremental: bool = False) -> str:
    os.makedirs(OBJ, exist_ok=True)
    headers = glob.glob(os.path.join(CSRC, "*.h")) + glob.glob(os.path.join(INCLUDE, "*.h"))
    jobs = []
    for s in SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(OBJ, s.replace(".hip", ".o"))
        if not incremental or _stale(src, obj, headers):
            jobs.append((src, obj))

    def compile_one(job):
        src, obj = job
        tmp = obj + ".tmp%d" % os.getpid()
        err = _run([HIPCC] + FLAGS + ["-c", "-o", tmp, src])
        os.replace(tmp, obj)
        return err

    with ThreadPoolExecutor(max_workers=len(SOURCES)) as ex:
        errs = list(ex.map(compile_one, jobs))
    tmp = OUT + ".tmp%d" % os.getpid()
    errs.append(_run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] +
                     [os.path.join(OBJ, s.replace(".hip", ".o")) for s in SOURCES]))
    if verbose:
        for e in errs:
            if e.strip():
                print(e, file=sys.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv[1:]))
