"""sigp: MI355X (gfx950) evaluator of SI-GP dispatching rules.

Drop-in for the reference's GPU boundary (engine/factory.cuh FactoryContext)
through the C ABI in include/sigp.h, plus the host-side mirror of the
reference's instance readers, rule encoding, random streams and GP operators.
"""
from .instance import (Instance, POI, load_instance, parse_cordeau, parse_solomon,
                       synthetic_clustered, synthetic_mixed, synthetic_uniform)
from .rules import (EXTRA, FUNCS_RMOD, FUNCTIONS, N_TERMS, NS_TERMINALS, Node, Packed,
                    compile_tree, pack_compiled, pack_trees, parse_prefix, term_mask_from,
                    to_prefix)
from .scenarios import gen_durations

__all__ = ["Instance", "POI", "load_instance", "parse_cordeau", "parse_solomon",
           "synthetic_clustered", "synthetic_mixed", "synthetic_uniform", "EXTRA", "FUNCS_RMOD",
           "FUNCTIONS", "N_TERMS", "NS_TERMINALS", "Node", "Packed", "compile_tree",
           "pack_compiled", "pack_trees", "parse_prefix", "term_mask_from", "to_prefix",
           "gen_durations", "FactoryContext", "zp_of", "device_math", "device_info"]


def __getattr__(name):
    # the GPU API loads libsigp.so lazily, so CPU-only users of the host mirror
    # do not need the library
    if name in ("FactoryContext", "zp_of", "device_math", "device_info", "Batch"):
        from . import api
        return getattr(api, name)
    raise AttributeError(name)
