"""lslab_amd: MI355X-native LSlab slab-list hash map and set.

The product is liblslab.so (hand-written gfx950 HIP kernels behind the C ABI
in include/lslab_c.h) and the header-only C++ API in include/lslab/.  This
Python package binds the same C ABI for tests and benchmarks.
"""
from . import _lib  # noqa: F401
from ._lib import (COMBINE_BLOCK, COMBINE_WAVE, ENGINE_PLAIN, ENGINE_TWO_PHASE,  # noqa: F401
                   HASH_IDENTITY, HASH_MIX, OP_GET, OP_PUT, OP_REMOVE, RECLAIM_OFF,
                   RECLAIM_ON, RUNS_OFF, RUNS_ON, SLOT_HEADS_OFF, SLOT_HEADS_ON, LslabError)

__all__ = ["Map", "Set", "Kvcg", "Wide", "tuning", "slab_bytes", "workload", "ENGINE_PLAIN", "ENGINE_TWO_PHASE",
           "COMBINE_WAVE", "COMBINE_BLOCK", "HASH_IDENTITY", "HASH_MIX", "RECLAIM_OFF",
           "RECLAIM_ON", "SLOT_HEADS_ON", "SLOT_HEADS_OFF", "RUNS_ON", "RUNS_OFF", "OP_GET",
           "OP_PUT", "OP_REMOVE",
           "LslabError"]


def __getattr__(name):
    # torch-dependent parts load lazily so that the C ABI can be inspected
    # without importing torch.
    import importlib
    if name in ("Map", "Set", "Kvcg", "Wide", "tuning", "slab_bytes"):
        return getattr(importlib.import_module(__name__ + ".table"), name)
    if name in ("workload", "table", "autotune", "build"):
        return importlib.import_module(__name__ + "." + name)
    raise AttributeError(name)
