"""Python face of the evaluator, shaped like the reference's FactoryContext
(engine/factory.cuh:124-217) so that callers and tests read the same.

All work runs in libsigp.so on the GPU; nothing here computes a reward.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from ._lib import (MAP_AUTO, MAP_CANDIDATES, MAP_CANDIDATES_2, MAP_SCENARIOS, PROTOCOL_ROUND_ROBIN,
                   PROTOCOL_VEHICLES, RR_MAX_M, InstanceDesc, LaunchConfig, Rules, check, lib)
from .instance import Instance
from .rules import N_TERMS, Packed

__all__ = ["FactoryContext", "Batch", "zp_of", "device_math", "device_info", "evaluate_multi",
           "batch_multi", "sample_durations", "sampler", "MAP_AUTO", "MAP_SCENARIOS", "MAP_CANDIDATES",
           "MAP_CANDIDATES_2", "RR_MAX_M"]


def _ptr(a: Optional[np.ndarray]):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64))


def _desc(inst: Instance, zp: Optional[np.ndarray], keep: list) -> InstanceDesc:
    arrs = [_f64(inst.scores), _f64(inst.durations), _f64(inst.open_t), _f64(inst.close_t),
            _f64(inst.dist), _f64(inst.dist_to_depot)]
    zpa = None if zp is None else _f64(zp)
    keep.extend(arrs + [zpa])
    return InstanceDesc(inst.n, inst.m, float(inst.tmax), *[a.ctypes.data for a in arrs],
                        None if zpa is None else zpa.ctypes.data)


def _rules(pk: Packed, keep: list) -> Rules:
    arrs = [np.ascontiguousarray(x, dtype=np.int32) for x in (pk.ops, pk.args, pk.offsets, pk.lengths)]
    keep.extend(arrs)
    return Rules(*[a.ctypes.data for a in arrs], pk.n_rules, pk.n_nodes)


def _masks(masks) -> np.ndarray:
    m = np.ascontiguousarray(np.asarray(masks, dtype=np.int64))
    if m.ndim == 1:
        m = m.reshape(1, -1)
    if m.shape[1] != N_TERMS:
        raise ValueError("mask rows must have %d entries" % N_TERMS)
    return m


def zp_of(inst: Instance) -> np.ndarray:
    """engine/sim_host.hpp:96-105, computed on the device."""
    keep: list = []
    d = _desc(inst, None, keep)
    out = np.zeros(10, dtype=np.float64)
    check(lib().sigp_zstats(C.byref(d), _ptr(out)))
    return out


def device_info():
    buf = C.create_string_buffer(256)
    ncu = C.c_int32(0)
    clk = C.c_double(0.0)
    check(lib().sigp_device_info(buf, 256, C.byref(ncu), C.byref(clk)))
    return buf.value.decode(), ncu.value, clk.value


def device_math(which: str, x, y=None) -> np.ndarray:
    """Device erf / erfc / exp / sqrt / pow (the simulator's own routines)."""
    code = {"erf": 0, "erfc": 1, "exp": 2, "sqrt": 3, "pow": 4}[which]
    xs = _f64(x).ravel()
    ys = None if y is None else _f64(y).ravel()
    out = np.zeros_like(xs)
    check(lib().sigp_device_math(code, _ptr(xs), _ptr(ys), xs.size, _ptr(out)))
    return out


class Batch:
    """Rules and masks resident on the device (sigp_batch_*), over one
    instance or, from batch_multi, several."""

    def __init__(self, ctx: "FactoryContext", pk: Packed, T: int, group_size: int, masks,
                 mask_group: int, _multi: Optional[Sequence["FactoryContext"]] = None):
        keep: list = []
        r = _rules(pk, keep)
        m = _masks(masks)
        self._h = C.c_void_p()
        if _multi is None:
            check(lib().sigp_batch_create(ctx._h, C.byref(r), T, group_size, _ptr(m), m.shape[0],
                                          mask_group, C.byref(self._h)))
            self.ctxs = [ctx]
        else:
            hs = (C.c_void_p * len(_multi))(*[c._h for c in _multi])
            check(lib().sigp_batch_create_multi(hs, len(_multi), C.byref(r), T, _ptr(m),
                                                m.shape[0], mask_group, C.byref(self._h)))
            self.ctxs = list(_multi)
        self.ctx = ctx
        self.P = pk.n_rules
        self.T = T
        self.group_size = group_size

    def run(self, sampled_ptr: int, n_rows: int, out_pairs_ptr: Optional[int],
            out_mean_ptr: Optional[int] = None, stream: Optional[int] = None) -> None:
        """Launch on `stream` with device pointers (e.g. torch tensors' data_ptr())."""
        check(lib().sigp_batch_run(self._h, C.c_void_p(sampled_ptr), n_rows,
                                   None if out_pairs_ptr is None else C.c_void_p(out_pairs_ptr),
                                   None if out_mean_ptr is None else C.c_void_p(out_mean_ptr),
                                   None if stream is None else C.c_void_p(stream)))

    def configure(self, waves_per_item: int, mapping: int = MAP_AUTO) -> None:
        """Launch mapping: MAP_SCENARIOS / MAP_CANDIDATES / MAP_CANDIDATES_2 /
        MAP_AUTO and the wavefronts per workgroup (1, 2, 4; 0 = the mapping's
        default; MAP_CANDIDATES_2 uses one)."""
        cfg = LaunchConfig(waves_per_item, mapping)
        check(lib().sigp_batch_configure(self._h, C.byref(cfg)))

    def protocol(self, round_robin: bool) -> None:
        """Switch this batch between sim_one (False) and sim_one_rr (True)
        (sigp_batch_protocol)."""
        check(lib().sigp_batch_protocol(self._h, PROTOCOL_ROUND_ROBIN if round_robin else PROTOCOL_VEHICLES))

    def order_scenarios(self, enable: bool) -> bool:
        """Scenario order of scenario lanes (sigp_batch_order_scenarios); returns
        whether it is on (it needs 64 <= T <= 4096)."""
        on = C.c_int32(0)
        check(lib().sigp_batch_order_scenarios(self._h, 1 if enable else 0, C.byref(on)))
        return bool(on.value)

    def autotune(self, sampled_ptr: int, n_rows: int, stream: Optional[int] = None):
        """Time every mapping, keep the fastest; returns (waves_per_item, mapping)."""
        cfg = LaunchConfig(0, 0)
        check(lib().sigp_batch_autotune(self._h, C.c_void_p(sampled_ptr), n_rows,
                                        None if stream is None else C.c_void_p(stream),
                                        C.byref(cfg)))
        return cfg.waves_per_item, cfg.mapping

    def close(self) -> None:
        if self._h:
            lib().sigp_batch_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class FactoryContext:
    """Instance uploaded once; evaluate* return out[p*N + s] like the reference."""

    def __init__(self, inst: Instance, zp: Optional[Sequence[float]] = None):
        self.inst = inst
        self.zp = zp_of(inst) if zp is None else _f64(zp)
        keep: list = []
        d = _desc(inst, self.zp, keep)
        self._h = C.c_void_p()
        check(lib().sigp_instance_create(C.byref(d), C.byref(self._h)))

    def skip_repeats(self, enable: bool) -> None:
        """Later batches and evaluate* calls simulate each distinct rule of a
        scenario group once and copy its rewards to the repeats
        (sigp_instance_skip_repeats)."""
        check(lib().sigp_instance_skip_repeats(self._h, 1 if enable else 0))

    def protocol(self, round_robin: bool) -> None:
        """Simulation protocol of later batches and evaluate* calls
        (sigp_instance_protocol): False = sim_one, vehicle by vehicle
        (engine/sim_core.hpp:207-238, the default); True = sim_one_rr, the
        open vehicles deciding in turn (engine/sim_core.hpp:241-274;
        m <= RR_MAX_M)."""
        check(lib().sigp_instance_protocol(self._h, PROTOCOL_ROUND_ROBIN if round_robin else PROTOCOL_VEHICLES))

    def one_launch(self, enable: bool) -> None:
        """Multi-instance batches created later over instances that all have
        this enabled run every size class in one launch
        (sigp_instance_one_launch)."""
        check(lib().sigp_instance_one_launch(self._h, 1 if enable else 0))

    def configure(self, waves_per_item: int, mapping: int = MAP_AUTO) -> None:
        """Default launch mapping of later batches and evaluate* calls."""
        cfg = LaunchConfig(waves_per_item, mapping)
        check(lib().sigp_instance_configure(self._h, C.byref(cfg)))

    def close(self) -> None:
        if self._h:
            lib().sigp_instance_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _run(self, pk: Packed, sampled, T: int, group_size: int, masks, mask_group: int,
             want_mean: bool = False):
        keep: list = []
        r = _rules(pk, keep)
        s = _f64(sampled)
        n = self.inst.n
        n_rows = s.size // n if n > 0 else 0
        m = _masks(masks)
        P = pk.n_rules
        out = np.zeros(P * T, dtype=np.float64)
        mean = np.zeros(P, dtype=np.float64) if want_mean else None
        check(lib().sigp_evaluate(self._h, C.byref(r), _ptr(s), n_rows, T, group_size, _ptr(m),
                                  m.shape[0], mask_group, _ptr(out), _ptr(mean)))
        return (out, mean) if want_mean else out

    # engine/factory.cuh:144-164
    def evaluate(self, pk: Packed, all_sampled, N: int, term_mask) -> np.ndarray:
        P = max(pk.n_rules, 1)
        return self._run(pk, all_sampled, N, P, term_mask, P)

    # engine/factory.cuh:167-189
    def evaluate_grouped(self, pk: Packed, grouped_sampled, G: int, T: int, group_size: int,
                         term_mask) -> np.ndarray:
        return self._run(pk, grouped_sampled, T, group_size, term_mask, max(pk.n_rules, 1))

    # engine/factory.cuh:193-216
    def evaluate_grouped_masks(self, pk: Packed, grouped_sampled, T: int, group_size: int,
                               mask_group: int, masks) -> np.ndarray:
        return self._run(pk, grouped_sampled, T, group_size, masks, mask_group)

    def evaluate_means(self, pk: Packed, all_sampled, N: int, term_mask):
        """(pairs, per-rule means) in one call."""
        P = max(pk.n_rules, 1)
        return self._run(pk, all_sampled, N, P, term_mask, P, want_mean=True)

    def batch(self, pk: Packed, T: int, masks, group_size: Optional[int] = None,
              mask_group: Optional[int] = None) -> Batch:
        P = max(pk.n_rules, 1)
        return Batch(self, pk, T, group_size or P, masks, mask_group or P)


def batch_multi(ctxs: Sequence[FactoryContext], pk: Packed, T: int, masks,
                mask_group: Optional[int] = None) -> Batch:
    """Every rule of pk on every instance of ctxs (sigp_batch_create_multi);
    run with the instances' scenario blocks back to back, n_rows = len(ctxs) * T,
    results out[(k * P + p) * T + s]."""
    P = max(pk.n_rules, 1)
    return Batch(ctxs[0], pk, T, P, masks, mask_group or P, _multi=list(ctxs))


def evaluate_multi(ctxs: Sequence[FactoryContext], pk: Packed, blocks: Sequence[np.ndarray], T: int,
                   masks, mask_group: Optional[int] = None, want_mean: bool = False):
    """sigp_evaluate_multi: blocks[k] holds instance k's T scenario rows;
    returns out[(k * P + p) * T + s] (and the n_inst * P means)."""
    keep: list = []
    r = _rules(pk, keep)
    m = _masks(masks)
    K, P = len(ctxs), pk.n_rules
    for k, (c, b) in enumerate(zip(ctxs, blocks)):
        if np.asarray(b).size != T * c.inst.n:
            raise ValueError("block %d must hold T x n = %d x %d durations" % (k, T, c.inst.n))
    flat = _f64(np.concatenate([np.asarray(b, dtype=np.float64).ravel() for b in blocks]))
    hs = (C.c_void_p * K)(*[c._h for c in ctxs])
    out = np.zeros(K * P * T, dtype=np.float64)
    mean = np.zeros(K * P, dtype=np.float64) if want_mean else None
    check(lib().sigp_evaluate_multi(hs, K, C.byref(r), _ptr(flat), T, _ptr(m), m.shape[0],
                                    mask_group or max(P, 1), _ptr(out), _ptr(mean)))
    return (out, mean) if want_mean else out


def sampler(mode: str = "gauss", sigma: float = 0.2, shape: float = 4.0):
    """The sampler description of include/sigp.h (sigp_sampler)."""
    from ._lib import SAMPLE_GAMMA, SAMPLE_GAUSS, Sampler
    if mode not in ("gauss", "gamma"):
        raise ValueError("mode must be 'gauss' or 'gamma'")
    return Sampler(SAMPLE_GAUSS if mode == "gauss" else SAMPLE_GAMMA, float(sigma), float(shape))


def sample_durations(ctx: "FactoryContext", n_rows: int, base_seed: int, mode: str = "gauss",
                     sigma: float = 0.2, shape: float = 4.0) -> np.ndarray:
    """Scenario rows sampled on the device (sigp_sample_durations): row r
    draws from Mt19937(base_seed + r), as gen_durations
    (engine/sim_host.hpp:50-58) with correctly rounded log/sin/cos."""
    out = np.zeros((max(n_rows, 0), ctx.inst.n), dtype=np.float64)
    smp = sampler(mode, sigma, shape)
    check(lib().sigp_sample_durations(ctx._h, C.byref(smp), int(base_seed), int(n_rows), _ptr(out)))
    return out
