"""ctypes binding of include/sigp.h (libsigp.so, built in-tree by sigp.build).

There is no fallback: if the library is missing every entry point raises.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional

__all__ = ["lib", "lib_path", "SigpError", "check", "InstanceDesc", "Rules", "LaunchConfig", "Situations",
           "GpConfig", "Sampler", "EXPORTED", "GP_MAX_NODES", "GP_TOK_FUNC", "GP_TOK_CONST",
           "SAMPLE_GAUSS", "SAMPLE_GAMMA",
           "MAP_AUTO", "MAP_SCENARIOS", "MAP_CANDIDATES",
           "MAP_CANDIDATES_2", "PROTOCOL_VEHICLES", "PROTOCOL_ROUND_ROBIN", "RR_MAX_M"]

_HERE = os.path.dirname(os.path.abspath(__file__))
lib_path = os.path.join(_HERE, "libsigp.so")

# Every symbol include/sigp.h declares.
EXPORTED = ["sigp_last_error", "sigp_device_info", "sigp_instance_create", "sigp_instance_destroy",
            "sigp_zstats", "sigp_evaluate", "sigp_evaluate_noalloc", "sigp_batch_create", "sigp_batch_run",
            "sigp_batch_destroy", "sigp_batch_shape", "sigp_device_math", "sigp_batch_configure",
            "sigp_batch_order_scenarios", "sigp_instance_skip_repeats",
            "sigp_instance_protocol", "sigp_batch_protocol", "sigp_instance_one_launch", "sigp_batch_autotune", "sigp_instance_configure", "sigp_batch_create_multi",
            "sigp_evaluate_multi", "sigp_phenotype", "sigp_mlp_create", "sigp_mlp_destroy",
            "sigp_mlp_get_params", "sigp_mlp_set_params", "sigp_mlp_gradients", "sigp_mlp_train",
            "sigp_mlp_predict", "sigp_gp_create", "sigp_gp_destroy", "sigp_gp_init", "sigp_gp_evaluate",
            "sigp_gp_step", "sigp_gp_run", "sigp_gp_population", "sigp_gp_packed", "sigp_gp_rng_state",
            "sigp_gp_set_surrogate", "sigp_gp_init_screened", "sigp_gp_phenotype", "sigp_sample_durations",
            "sigp_sample_durations_async",
            "sigp_gp_set_sampler"]

_STATUS = {0: "SIGP_OK", 1: "SIGP_INVALID_ARG", 2: "SIGP_DEVICE_ERROR", 3: "SIGP_OOM",
           4: "SIGP_NO_DEVICE"}


class SigpError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__("%s: %s" % (_STATUS.get(status, str(status)), msg))
        self.status = status


class InstanceDesc(C.Structure):
    _fields_ = [("n", C.c_int32), ("m", C.c_int32), ("tmax", C.c_double),
                ("scores", C.c_void_p), ("durations", C.c_void_p), ("open_t", C.c_void_p),
                ("close_t", C.c_void_p), ("dist", C.c_void_p), ("dist_to_depot", C.c_void_p),
                ("zp", C.c_void_p)]


MAP_AUTO, MAP_SCENARIOS, MAP_CANDIDATES, MAP_CANDIDATES_2 = 0, 1, 2, 3
PROTOCOL_VEHICLES, PROTOCOL_ROUND_ROBIN, RR_MAX_M = 0, 1, 8


class LaunchConfig(C.Structure):
    _fields_ = [("waves_per_item", C.c_int32), ("mapping", C.c_int32)]


class Situations(C.Structure):
    _fields_ = [("n_sit", C.c_int32), ("n_cand", C.c_int32), ("cur", C.c_void_p), ("ct", C.c_void_p),
                ("remaining", C.c_void_p), ("visited", C.c_void_p), ("cand", C.c_void_p)]


GP_MAX_NODES, GP_TOK_FUNC, GP_TOK_CONST = 512, 0x4000, 0x2000


class GpConfig(C.Structure):
    _fields_ = [("n_runs", C.c_int32), ("population", C.c_int32), ("offspring", C.c_int32),
                ("n_terminals", C.c_int32), ("terminal_slots", C.c_void_p), ("n_functions", C.c_int32),
                ("function_ops", C.c_void_p), ("min_depth", C.c_int32), ("max_depth", C.c_int32),
                ("depth_limit", C.c_int32), ("tournament_size", C.c_int32), ("elitism", C.c_int32),
                ("crossover_rate", C.c_double), ("mutation_rate", C.c_double), ("T", C.c_int32),
                ("launch", LaunchConfig)]


SAMPLE_GAUSS, SAMPLE_GAMMA = 0, 1


class Sampler(C.Structure):
    _fields_ = [("mode", C.c_int32), ("sigma", C.c_double), ("shape", C.c_double)]


class Rules(C.Structure):
    _fields_ = [("ops", C.c_void_p), ("args", C.c_void_p), ("offsets", C.c_void_p),
                ("lengths", C.c_void_p), ("n_rules", C.c_int32), ("n_nodes", C.c_int64)]


_lib: Optional[C.CDLL] = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(lib_path):
        raise RuntimeError("libsigp.so not found at %s: run `python -c \"import __graft_entry__ as g; "
                           "g.build()\"` (hipcc, gfx950) first" % lib_path)
    # A process holds one HIP runtime (soname libamdhip64.so.7).  PyTorch ships
    # its own build of it and fails to find the GPU on another one, so when
    # torch is installed it is loaded first and libsigp binds to the same runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(lib_path)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    sig = {
        "sigp_last_error": ([], C.c_char_p),
        "sigp_device_info": ([C.c_char_p, C.c_size_t, vp, vp], C.c_int),
        "sigp_instance_create": ([vp, vp], C.c_int),
        "sigp_instance_destroy": ([vp], C.c_int),
        "sigp_zstats": ([vp, vp], C.c_int),
        "sigp_evaluate": ([vp, vp, vp, i64, i32, i32, vp, i32, i32, vp, vp], C.c_int),
        "sigp_evaluate_noalloc": ([vp, vp, vp, i64, i32, i32, vp, i32, i32, vp, vp, vp], C.c_int),
        "sigp_batch_create": ([vp, vp, i32, i32, vp, i32, i32, vp], C.c_int),
        "sigp_batch_run": ([vp, vp, i64, vp, vp, vp], C.c_int),
        "sigp_batch_destroy": ([vp], C.c_int),
        "sigp_batch_shape": ([vp, vp, vp], C.c_int),
        "sigp_device_math": ([i32, vp, vp, i64, vp], C.c_int),
        "sigp_batch_configure": ([vp, vp], C.c_int),
        "sigp_batch_order_scenarios": ([vp, C.c_int32, vp], C.c_int),
        "sigp_instance_skip_repeats": ([vp, C.c_int32], C.c_int),
        "sigp_instance_protocol": ([vp, C.c_int32], C.c_int),
        "sigp_batch_protocol": ([vp, C.c_int32], C.c_int),
        "sigp_instance_one_launch": ([vp, C.c_int32], C.c_int),
        "sigp_batch_autotune": ([vp, vp, i64, vp, vp], C.c_int),
        "sigp_instance_configure": ([vp, vp], C.c_int),
        "sigp_batch_create_multi": ([vp, i32, vp, i32, vp, i32, i32, vp], C.c_int),
        "sigp_evaluate_multi": ([vp, i32, vp, vp, i32, vp, i32, i32, vp, vp], C.c_int),
        "sigp_phenotype": ([vp, vp, vp, i32, i32, vp, vp], C.c_int),
        "sigp_mlp_create": ([C.c_uint64, vp], C.c_int),
        "sigp_mlp_destroy": ([vp], C.c_int),
        "sigp_mlp_get_params": ([vp, vp], C.c_int),
        "sigp_mlp_set_params": ([vp, vp], C.c_int),
        "sigp_mlp_gradients": ([vp, vp, vp, i32, vp, vp], C.c_int),
        "sigp_mlp_train": ([vp, vp, vp, i64, i32, i32, C.c_float, C.c_uint64, vp], C.c_int),
        "sigp_mlp_predict": ([vp, vp, i64, vp], C.c_int),
        "sigp_gp_create": ([vp, vp, vp, vp, vp], C.c_int),
        "sigp_gp_destroy": ([vp], C.c_int),
        "sigp_gp_init": ([vp], C.c_int),
        "sigp_gp_evaluate": ([vp, vp, i64, vp], C.c_int),
        "sigp_gp_step": ([vp, vp], C.c_int),
        "sigp_gp_run": ([vp, vp, i64, i32, i32, vp, vp, vp, vp], C.c_int),
        "sigp_gp_set_surrogate": ([vp, vp, vp], C.c_int),
        "sigp_gp_init_screened": ([vp], C.c_int),
        "sigp_gp_phenotype": ([vp, vp, vp], C.c_int),
        "sigp_gp_set_sampler": ([vp, vp, vp, i64, i64], C.c_int),
        "sigp_sample_durations": ([vp, vp, i64, i64, vp], C.c_int),
        "sigp_sample_durations_async": ([vp, vp, i64, i64, vp, vp], C.c_int),
        "sigp_gp_population": ([vp, vp, vp], C.c_int),
        "sigp_gp_packed": ([vp, vp, vp, vp], C.c_int),
        "sigp_gp_rng_state": ([vp, i32, i32, vp], C.c_int),
    }
    for name, (argt, rest) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = argt
        fn.restype = rest
    _lib = L
    return L


def check(status: int) -> None:
    if status != 0:
        raise SigpError(status, lib().sigp_last_error().decode())
