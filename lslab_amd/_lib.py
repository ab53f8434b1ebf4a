"""ctypes binding of liblslab.so (the C ABI in include/lslab_c.h).

The product path: every call goes to the HIP library.  There is no CPU
fallback; loading fails loudly when the library has not been built.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblslab.so")

OK = 0
ERR_INVALID = 1
ERR_HIP = 2
ERR_OOM = 3
ERR_POOL_EXHAUSTED = 4
ERR_TIMEOUT = 5

OP_GET = 0
OP_PUT = 1
OP_REMOVE = 2

ENGINE_PLAIN = 0
ENGINE_TWO_PHASE = 1

COMBINE_WAVE = 1
COMBINE_BLOCK = 2

HASH_IDENTITY = 1
HASH_MIX = 2

RECLAIM_OFF = 1
RECLAIM_ON = 2
SLOT_HEADS_ON = 1
SLOT_HEADS_OFF = 2
RUNS_ON = 1
RUNS_OFF = 2

CHAIN_BINS = 16

# Every symbol include/lslab_c.h declares.
EXPORTS = (
    "lslab_map_create", "lslab_set_create", "lslab_destroy", "lslab_set_tuning",
    "lslab_get_tuning", "lslab_map_put_pairs", "lslab_map_put", "lslab_map_get",
    "lslab_map_remove", "lslab_map_mixed", "lslab_set_insert", "lslab_set_contains",
    "lslab_set_remove", "lslab_set_mixed", "lslab_exec_batches", "lslab_status",
    "lslab_stats_get", "lslab_dump", "lslab_error_string", "lslab_serve", "lslab_serve_result",
    "lslab_chain_histogram", "lslab_slab_bytes", "lslab_map_create_at", "lslab_set_create_at",
    "lslab_table_info", "lslab_kvcg_create", "lslab_kvcg_destroy", "lslab_kvcg_set_hot",
    "lslab_kvcg_sample_hot", "lslab_kvcg_mixed", "lslab_kvcg_flush", "lslab_kvcg_stats_get",
    "lslab_wide_create", "lslab_wide_destroy", "lslab_wide_put", "lslab_wide_get",
    "lslab_wide_remove", "lslab_wide_mixed", "lslab_wide_status", "lslab_wide_stats",
    "lslab_wide_dump", "lslab_wide_slab_count", "lslab_exec_batches_noalloc",
)

NOALLOC_MAX_SUBBATCH = 65536
NOALLOC_MAX_BLOCKS = 256

SERVE_REJECTED = 0xFFFFFFFF


class Tuning(ctypes.Structure):
    _fields_ = [
        ("lanes_per_op", ctypes.c_int),
        ("engine", ctypes.c_int),
        ("block_threads", ctypes.c_int),
        ("max_blocks", ctypes.c_int),
        ("max_subbatch", ctypes.c_uint64),
        ("combine", ctypes.c_int),
        ("hash", ctypes.c_int),
        ("reclaim", ctypes.c_int),
        ("spin_ms", ctypes.c_int),
        ("slot_heads", ctypes.c_int),
        ("runs", ctypes.c_int),
    ]


class Stats(ctypes.Structure):
    _fields_ = [
        ("buckets", ctypes.c_uint64),
        ("pool_slabs", ctypes.c_uint64),
        ("slabs_used", ctypes.c_uint64),
        ("present", ctypes.c_uint64),
        ("max_chain", ctypes.c_uint64),
        ("slab_bytes", ctypes.c_uint64),
        ("device_bytes", ctypes.c_uint64),
    ]


class ServeParams(ctypes.Structure):
    _fields_ = [
        ("nreq", ctypes.c_uint64),
        ("req_offsets", ctypes.c_void_p),
        ("arrival_ticks", ctypes.c_void_p),
        ("req_reserve", ctypes.c_void_p),
        ("complete_ticks", ctypes.c_void_p),
        ("tick_of", ctypes.c_void_p),
        ("max_tick_ops", ctypes.c_uint64),
        ("max_tick_reqs", ctypes.c_uint32),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("ticks", ctypes.c_uint64),
        ("rejected", ctypes.c_uint64),
        ("served_ops", ctypes.c_uint64),
        ("elapsed_ticks", ctypes.c_uint64),
    ]


class KvcgStats(ctypes.Structure):
    _fields_ = [
        ("hot_keys", ctypes.c_uint64),
        ("hot_capacity", ctypes.c_uint64),
        ("hot_ops", ctypes.c_uint64),
        ("cold_ops", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
    ]


class LslabError(RuntimeError):
    def __init__(self, code, what=""):
        self.code = code
        super().__init__(f"lslab error {code} ({error_string(code)}) {what}".strip())


_LIB = None


def load():
    """Loads liblslab.so; raises if it is missing (no fallback path)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is not built; run python -m lslab_amd.build")
    lib = ctypes.CDLL(LIB_PATH)
    vp, sz, u64, i32 = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "lslab_map_create": [pp, i32, i32, ctypes.c_uint, u64, ctypes.POINTER(Tuning)],
        "lslab_set_create": [pp, i32, ctypes.c_uint, u64, ctypes.POINTER(Tuning)],
        "lslab_destroy": [vp],
        "lslab_set_tuning": [vp, ctypes.POINTER(Tuning)],
        "lslab_get_tuning": [vp, ctypes.POINTER(Tuning)],
        "lslab_map_put_pairs": [vp, vp, vp, sz, vp],
        "lslab_map_put": [vp, vp, vp, vp, sz, vp],
        "lslab_map_get": [vp, vp, vp, sz, vp],
        "lslab_map_remove": [vp, vp, vp, sz, vp],
        "lslab_map_mixed": [vp, vp, vp, vp, vp, sz, vp],
        "lslab_set_insert": [vp, vp, vp, sz, vp],
        "lslab_set_contains": [vp, vp, vp, sz, vp],
        "lslab_set_remove": [vp, vp, vp, sz, vp],
        "lslab_set_mixed": [vp, vp, vp, vp, sz, vp],
        "lslab_exec_batches": [vp, sz, vp, vp, vp, vp, vp, vp],
        "lslab_exec_batches_noalloc": [vp, sz, vp, vp, vp, vp, vp, vp],
        "lslab_status": [vp, vp],
        "lslab_stats_get": [vp, ctypes.POINTER(Stats), vp],
        "lslab_dump": [vp, vp, vp, u64, ctypes.POINTER(ctypes.c_uint64), vp],
        "lslab_serve": [vp, ctypes.POINTER(ServeParams), vp, vp, vp, vp, vp],
        "lslab_serve_result": [vp, ctypes.POINTER(ServeStats), vp],
        "lslab_chain_histogram": [vp, ctypes.POINTER(ctypes.c_uint64), ctypes.c_uint, vp],
        "lslab_map_create_at": [pp, i32, i32, ctypes.c_uint, u64, vp, u64, ctypes.POINTER(Tuning)],
        "lslab_set_create_at": [pp, i32, ctypes.c_uint, u64, vp, u64, ctypes.POINTER(Tuning)],
        "lslab_table_info": [vp] + [ctypes.POINTER(ctypes.c_int)] * 4,
        "lslab_kvcg_create": [pp, vp, u64, u64, i32],
        "lslab_kvcg_destroy": [vp],
        "lslab_kvcg_set_hot": [vp, vp, u64, vp],
        "lslab_kvcg_sample_hot": [vp, vp, u64, u64, vp],
        "lslab_kvcg_mixed": [vp, vp, vp, vp, vp, u64, vp],
        "lslab_kvcg_flush": [vp, vp],
        "lslab_kvcg_stats_get": [vp, ctypes.POINTER(KvcgStats)],
        "lslab_wide_create": [pp, i32, i32, ctypes.c_uint, u64, i32],
        "lslab_wide_destroy": [vp],
        "lslab_wide_put": [vp, vp, vp, vp, vp, sz, vp],
        "lslab_wide_get": [vp, vp, vp, vp, sz, vp],
        "lslab_wide_remove": [vp, vp, vp, vp, sz, vp],
        "lslab_wide_mixed": [vp, vp, vp, vp, vp, vp, sz, vp],
        "lslab_wide_status": [vp, vp],
        "lslab_wide_stats": [vp, ctypes.POINTER(Stats), vp],
        "lslab_wide_dump": [vp, vp, vp, u64, ctypes.POINTER(ctypes.c_uint64), vp],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    lib.lslab_slab_bytes.argtypes = [i32, i32, i32, ctypes.c_uint, u64]
    lib.lslab_slab_bytes.restype = u64
    lib.lslab_wide_slab_count.argtypes = [ctypes.c_uint, u64]
    lib.lslab_wide_slab_count.restype = u64
    lib.lslab_error_string.argtypes = [ctypes.c_int]
    lib.lslab_error_string.restype = ctypes.c_char_p
    _LIB = lib
    return lib


def error_string(code):
    try:
        return load().lslab_error_string(int(code)).decode()
    except Exception:  # pragma: no cover - only for messages
        return "?"


def check(code, what=""):
    if code != OK:
        raise LslabError(code, what)
