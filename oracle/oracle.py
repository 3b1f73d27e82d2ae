"""CPU ORACLE -- test infrastructure only.

ctypes access to the C restatement (liboracle.so / liboracle_count.so) and to
the reference's own CPU simulator compiled from its sources
(_ref/libsigpref.so, present only where the reference was available at build
time).  Imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg; never by the product.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
N_TERMS = 61


class OrcInstance(C.Structure):
    _fields_ = [("n", C.c_int), ("m", C.c_int), ("tmax", C.c_double),
                ("score", C.c_void_p), ("dur", C.c_void_p), ("open", C.c_void_p),
                ("close", C.c_void_p), ("dist", C.c_void_p), ("d2d", C.c_void_p)]


def build(reference: bool = True) -> None:
    """Compile the oracle (and, where the reference sources exist, _ref: the
    reference's CPU simulator, and its GPU gate and tools linked against
    libsigp.so)."""
    targets = ["liboracle.so", "liboracle_count.so", "libcrmath.so", "libpowcheck.so"] + (["ref", "gate", "tools", "gpcheck"] if reference else [])
    subprocess.run(["make", "-s", "-C", HERE] + targets, check=True)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class _Staged:
    """Keeps the numpy arrays of an instance alive for an OrcInstance."""

    def __init__(self, inst, m: Optional[int] = None):
        self.arrs = [_f64(inst.scores), _f64(inst.durations), _f64(inst.open_t),
                     _f64(inst.close_t), _f64(inst.dist), _f64(inst.dist_to_depot)]
        self.n = inst.n
        self.m = inst.m if m is None else m
        self.tmax = float(inst.tmax)
        self.st = OrcInstance(self.n, self.m, self.tmax, *[a.ctypes.data for a in self.arrs])


class Oracle:
    """The C restatement (sim_oracle.c)."""

    def __init__(self, counting: bool = False):
        name = "liboracle_count.so" if counting else "liboracle.so"
        path = os.path.join(HERE, name)
        if not os.path.exists(path):
            build(reference=False)
        L = C.CDLL(path)
        vp, i32, dbl = C.c_void_p, C.c_int, C.c_double
        L.orc_sim_one.argtypes = [vp, vp, vp, i32, vp, vp, vp, vp]
        L.orc_sim_one.restype = dbl
        L.orc_zstats.argtypes = [vp, vp]
        L.orc_sim_batch.argtypes = [vp, vp, vp, i32, vp, i32, vp, vp, vp]
        L.orc_evaluate.argtypes = [vp, vp, vp, vp, vp, i32, vp, i32, vp, vp, vp, vp, vp, i32, vp]
        L.orc_pick_best.argtypes = [vp, i32, dbl, i32, vp, vp, i32, vp, vp, vp, vp, vp]
        L.orc_pick_best.restype = i32
        L.orc_terminal.argtypes = [vp, i32, i32, i32, dbl, i32, vp, vp]
        L.orc_terminal.restype = dbl
        L.orc_ops_read.restype = C.c_ulonglong
        L.orc_terms_at.argtypes = [vp, i32, i32, dbl, i32, vp, vp, vp, vp]
        L.orc_terms_at.restype = i32
        L.orc_terms_all.argtypes = [vp, i32, dbl, i32, vp, vp, vp, vp, vp]
        L.orc_terms_all.restype = i32
        L.orc_sim_one_rr.argtypes = [vp, vp, vp, i32, vp, vp, vp]
        L.orc_sim_one_rr.restype = dbl
        self.L = L

    def zstats(self, inst) -> np.ndarray:
        s = _Staged(inst)
        out = np.zeros(10)
        self.L.orc_zstats(C.byref(s.st), _p(out))
        return out

    def sim_batch(self, inst, ops, args, sampled, n_samples, mask, zp) -> np.ndarray:
        s = _Staged(inst)
        o, a = _i32(ops), _i32(args)
        sm, mk, z = _f64(sampled), np.ascontiguousarray(mask, dtype=np.int64), _f64(zp)
        out = np.zeros(n_samples)
        self.L.orc_sim_batch(C.byref(s.st), _p(o), _p(a), len(o), _p(sm), n_samples, _p(mk),
                             _p(z), _p(out))
        return out

    def evaluate(self, inst, pk, sampled, N, masks, zp, group_size=None, mask_group=None,
                 need_rows=None, threads: int = 1) -> np.ndarray:
        """out[p*N+s] with the reference's grouped index maths."""
        s = _Staged(inst)
        P = pk.n_rules
        gs = group_size or max(P, 1)
        mg = mask_group or max(P, 1)
        scen_base = np.ascontiguousarray((np.arange(P) // gs) * N, dtype=np.int64)
        mrow = np.ascontiguousarray(np.arange(P) // mg, dtype=np.int32)
        mk = np.ascontiguousarray(np.atleast_2d(masks), dtype=np.int64)
        nd = None if need_rows is None else np.ascontiguousarray(need_rows, dtype=np.uint8)
        out = np.zeros(P * N)
        arrs = [_i32(pk.ops), _i32(pk.args), _i32(pk.offsets), _i32(pk.lengths), _f64(sampled)]
        z = _f64(zp)
        self.L.orc_evaluate(C.byref(s.st), *[_p(a) for a in arrs[:4]], P, _p(arrs[4]), N,
                            _p(scen_base), _p(mk), _p(mrow), _p(z), _p(nd), threads, _p(out))
        return out

    def pick_best(self, inst, current, ctime, remaining, ops, args, visited, mask, zp):
        s = _Staged(inst)
        o, a = _i32(ops), _i32(args)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        mk = np.ascontiguousarray(mask, dtype=np.int64)
        z = _f64(zp)
        pri = np.zeros(inst.n)
        b = self.L.orc_pick_best(C.byref(s.st), current, ctime, remaining, _p(o), _p(a), len(o),
                                 _p(vis), _p(mk), _p(z), None, _p(pri))
        return b, pri

    def terminal(self, inst, slot, cand, current, ctime, remaining, visited, zp) -> float:
        s = _Staged(inst)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        z = _f64(zp)
        return self.L.orc_terminal(C.byref(s.st), slot, cand, current, ctime, remaining, _p(vis), _p(z))

    def terms_at(self, inst, cand, current, ctime, remaining, visited, mask, zp):
        """The 61 slot values of candidate `cand` (None if it is visited or
        infeasible)."""
        s = _Staged(inst)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        mk = np.ascontiguousarray(mask, dtype=np.int64)
        z = _f64(zp)
        out = np.zeros(N_TERMS)
        hit = self.L.orc_terms_at(C.byref(s.st), cand, current, ctime, remaining, _p(vis), _p(mk),
                                  _p(z), _p(out))
        return out if hit else None

    def terms_all(self, inst, current, ctime, remaining, visited, mask, zp):
        """(n x 61 slot values, feasible flags) of every candidate of one
        decision; rows of infeasible or visited candidates are zero."""
        s = _Staged(inst)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        mk = np.ascontiguousarray(mask, dtype=np.int64)
        z = _f64(zp)
        out = np.zeros((inst.n, N_TERMS))
        pri = np.zeros(inst.n)
        self.L.orc_terms_all(C.byref(s.st), current, ctime, remaining, _p(vis), _p(mk), _p(z), _p(out),
                             _p(pri))
        return out, np.isfinite(pri)

    def sim_one_rr(self, inst, ops, args, sampled_row, mask, zp) -> float:
        s = _Staged(inst)
        o, a = _i32(ops), _i32(args)
        sm, mk, z = _f64(sampled_row), np.ascontiguousarray(mask, dtype=np.int64), _f64(zp)
        return self.L.orc_sim_one_rr(C.byref(s.st), _p(o), _p(a), len(o), _p(sm), _p(mk), _p(z))

    def evaluate_rr(self, inst, pk, sampled, N, masks, zp, group_size=None, mask_group=None,
                    threads: int = 1) -> np.ndarray:
        """out[p*N+s] = orc_sim_one_rr of rule p on scenario row
        (p / group_size) * N + s with mask row p / mask_group: the grouped
        index maths of evaluate, under the round-robin protocol.  Rules run
        on `threads` host threads (ctypes releases the GIL)."""
        from concurrent.futures import ThreadPoolExecutor
        st = _Staged(inst)
        P = pk.n_rules
        gs = group_size or max(P, 1)
        mg = mask_group or max(P, 1)
        S = _f64(sampled).reshape(-1, max(inst.n, 1))
        mk = np.ascontiguousarray(np.atleast_2d(masks), dtype=np.int64)
        z = _f64(zp)
        ops, args = _i32(pk.ops), _i32(pk.args)
        out = np.zeros(P * N)

        def rule(p):
            o, L = int(pk.offsets[p]), int(pk.lengths[p])
            op, ar = ops[o:o + L], args[2 * o:2 * (o + L)]
            row0 = (p // gs) * N
            for s in range(N):
                out[p * N + s] = self.L.orc_sim_one_rr(C.byref(st.st), _p(op), _p(ar), L, _p(S[row0 + s]),
                                                       _p(mk[p // mg]), _p(z))

        with ThreadPoolExecutor(max_workers=max(1, threads)) as ex:
            list(ex.map(rule, range(P)))
        return out

    def ops(self) -> int:
        return int(self.L.orc_ops_read())

    def reset_ops(self) -> None:
        self.L.orc_ops_reset()


class Reference:
    """The reference's own CPU simulator (oracle/_ref), when it was built."""

    path = os.path.join(HERE, "_ref", "libsigpref.so")

    @classmethod
    def available(cls) -> bool:
        return os.path.exists(cls.path)

    def __init__(self):
        L = C.CDLL(self.path)
        vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_longlong, C.c_double
        L.ref_zp_of.argtypes = [i32, i32, dbl] + [vp] * 7
        L.ref_evaluate.argtypes = [i32, i32, dbl] + [vp] * 6 + [vp, vp, vp, vp, i32, vp, i32, vp, vp,
                                                                vp, vp, i32, vp]
        L.ref_gen_durations.argtypes = [i32, vp, dbl, i32, i64, vp]
        L.ref_compile.argtypes = [C.c_char_p, vp, vp, i32]
        L.ref_compile.restype = i32
        L.ref_pick_best_terms.argtypes = [i32, i32, dbl] + [vp] * 6 + [i32, dbl, i32, vp, vp, vp, vp]
        L.ref_pick_best_terms.restype = i32
        L.ref_scan_terminal.argtypes = [i32, i32, dbl] + [vp] * 6 + [i32, i32, dbl, vp, vp, i32, vp]
        L.ref_scan_terminal.restype = i32
        L.ref_sim_one_rr.argtypes = [i32, i32, dbl] + [vp] * 6 + [vp, vp, i32, vp, vp, vp]
        L.ref_sim_one_rr.restype = dbl
        self.L = L

    def _inst(self, inst):
        arrs = [_f64(inst.scores), _f64(inst.durations), _f64(inst.open_t), _f64(inst.close_t),
                _f64(inst.dist), _f64(inst.dist_to_depot)]
        return arrs

    def zp_of(self, inst) -> np.ndarray:
        arrs = self._inst(inst)
        out = np.zeros(10)
        self.L.ref_zp_of(inst.n, inst.m, float(inst.tmax), *[_p(a) for a in arrs], _p(out))
        return out

    def evaluate(self, inst, pk, sampled, N, masks, zp, group_size=None, mask_group=None,
                 threads: int = 1) -> np.ndarray:
        arrs = self._inst(inst)
        P = pk.n_rules
        gs = group_size or max(P, 1)
        mg = mask_group or max(P, 1)
        scen_base = np.ascontiguousarray((np.arange(P) // gs) * N, dtype=np.int64)
        mrow = np.ascontiguousarray(np.arange(P) // mg, dtype=np.int32)
        mk = np.ascontiguousarray(np.atleast_2d(masks), dtype=np.int64)
        rr = [_i32(pk.ops), _i32(pk.args), _i32(pk.offsets), _i32(pk.lengths)]
        sm, z = _f64(sampled), _f64(zp)
        out = np.zeros(P * N)
        self.L.ref_evaluate(inst.n, inst.m, float(inst.tmax), *[_p(a) for a in arrs],
                            *[_p(a) for a in rr], P, _p(sm), N, _p(scen_base), _p(mk), _p(mrow),
                            _p(z), threads, _p(out))
        return out

    def pick_best_terms(self, inst, current, ctime, remaining, visited, mask, zp):
        """The reference's pick_best with a constant rule: (choice, the 61
        slot values of the last feasible candidate it evaluated)."""
        arrs = self._inst(inst)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        mk = np.ascontiguousarray(mask, dtype=np.int64)
        z = _f64(zp)
        out = np.full(N_TERMS, np.nan)
        b = self.L.ref_pick_best_terms(inst.n, inst.m, float(inst.tmax), *[_p(a) for a in arrs],
                                       current, ctime, remaining, _p(vis), _p(mk), _p(z), _p(out))
        return b, out

    def scan_terminal(self, inst, slot, current, ctime, visited, cands):
        """The reference's terminal function of `slot`, called directly for
        each candidate of `cands`."""
        arrs = self._inst(inst)
        vis = np.ascontiguousarray(visited, dtype=np.uint8)
        cd = np.ascontiguousarray(cands, dtype=np.int32)
        out = np.zeros(len(cd))
        rc = self.L.ref_scan_terminal(inst.n, inst.m, float(inst.tmax), *[_p(a) for a in arrs], slot,
                                      current, ctime, _p(vis), _p(cd), len(cd), _p(out))
        if rc != 0:
            raise ValueError("slot %d has no terminal function of its own" % slot)
        return out

    def sim_one_rr(self, inst, ops, args, sampled_row, mask, zp) -> float:
        arrs = self._inst(inst)
        o, a = _i32(ops), _i32(args)
        sm, mk, z = _f64(sampled_row), np.ascontiguousarray(mask, dtype=np.int64), _f64(zp)
        return self.L.ref_sim_one_rr(inst.n, inst.m, float(inst.tmax), *[_p(x) for x in arrs], _p(o),
                                     _p(a), len(o), _p(sm), _p(mk), _p(z))

    def gen_durations(self, expected, n_samples, base_seed, sigma=0.2) -> np.ndarray:
        e = _f64(expected)
        out = np.zeros(n_samples * e.size)
        self.L.ref_gen_durations(e.size, _p(e), sigma, n_samples, base_seed, _p(out))
        return out.reshape(n_samples, e.size)

    def compile(self, prefix: str):
        ops = np.zeros(4096, dtype=np.int32)
        args = np.zeros(8192, dtype=np.int32)
        k = self.L.ref_compile(prefix.encode(), _p(ops), _p(args), 4096)
        return ops[:k].copy(), args[:2 * k].copy()


def _live(s: int, mk) -> bool:
    """A slot's value is live (engine/sim_core.hpp:88-197): slots 0-9 always,
    REGRET/REACH2/ATRISK also through their Z partners, others by mask."""
    if s <= 9:
        return True
    if s == 11:
        return bool(mk[11] or mk[33])
    if s == 14:
        return bool(mk[14] or mk[32])
    if s == 30:
        return bool(mk[30] or mk[34])
    return bool(mk[s])


def need_rows_for(pk, masks, mask_group=None) -> np.ndarray:
    """P x 61 uint8: slots a rule reads that are live, closed under the
    dependencies (Z slots on their source, HARVEST/BOLSAO on NS and the
    structural family) -- the work the GPU evaluator performs."""
    masks = np.atleast_2d(masks)
    P = pk.n_rules
    mg = mask_group or max(P, 1)
    out = np.zeros((P, N_TERMS), dtype=np.uint8)
    for p in range(P):
        mk = masks[p // mg]
        o, L = int(pk.offsets[p]), int(pk.lengths[p])
        for k in range(o, o + L):
            if pk.ops[k] == -1:
                s = int(pk.args[2 * k])
                if _live(s, mk):
                    out[p, s] = 1
        if out[p, 32]:
            out[p, 14] = 1
        if out[p, 33]:
            out[p, 11] = 1
        if out[p, 34]:
            out[p, 30] = 1
    return out


class CrMath:
    """The device sampler's correctly rounded log and sincos on the host
    (libcrmath.so): the functions injected into the host sampler to restate
    the device's rows."""

    def __init__(self):
        path = os.path.join(HERE, "libcrmath.so")
        if not os.path.exists(path):
            build(reference=False)
        L = C.CDLL(path)
        L.orc_cr_log.argtypes = [C.c_double]
        L.orc_cr_log.restype = C.c_double
        L.orc_cr_sincos.argtypes = [C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_cr_sincos.restype = None
        self.L = L
        self._s = C.c_double()
        self._c = C.c_double()

    def log(self, x: float) -> float:
        return self.L.orc_cr_log(x)

    def sincos(self, x: float):
        self.L.orc_cr_sincos(x, C.byref(self._s), C.byref(self._c))
        return self._s.value, self._c.value


class PowCheck:
    """The device's glibc-compatible pow compiled for the host
    (libpowcheck.so), beside libm's pow."""

    def __init__(self):
        path = os.path.join(HERE, "libpowcheck.so")
        if not os.path.exists(path):
            build(reference=False)
        L = C.CDLL(path)
        L.orc_pow_sweep.argtypes = [C.c_void_p, C.c_void_p, C.c_long, C.c_void_p, C.c_void_p]
        L.orc_pow_sweep.restype = C.c_long
        L.orc_pow_restated.argtypes = [C.c_double, C.c_double]
        L.orc_pow_restated.restype = C.c_double
        self.L = L

    def sweep(self, x, y):
        """(restated, libm, number of bitwise differences)."""
        x, y = _f64(x), _f64(y)
        mine, ref = np.zeros_like(x), np.zeros_like(x)
        bad = self.L.orc_pow_sweep(_p(x), _p(y), x.size, _p(mine), _p(ref))
        return mine, ref, int(bad)
