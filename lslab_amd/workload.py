"""Seeded synthetic workloads of the benchmark configurations, generated on
the device by the kernels behind include/lslab_workload.h."""
import ctypes

import torch

from . import _lib

_DT = {4: torch.int32, 8: torch.int64}


class GenParams(ctypes.Structure):
    _fields_ = [
        ("seed", ctypes.c_uint64),
        ("first_op", ctypes.c_uint64),
        ("theta", ctypes.c_double),
        ("space_log2", ctypes.c_uint),
        ("p_get", ctypes.c_double),
        ("p_put", ctypes.c_double),
        ("key_bytes", ctypes.c_int),
        ("val_bytes", ctypes.c_int),
    ]


def _lib_gen():
    lib = _lib.load()
    if not getattr(lib, "_gen_ready", False):
        vp, u64, i32 = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
        lib.lslab_gen_keys.argtypes = [vp, i32, u64, u64, u64, vp]
        lib.lslab_gen_values.argtypes = [vp, i32, u64, u64, u64, vp]
        lib.lslab_gen_ops.argtypes = [ctypes.POINTER(GenParams), vp, vp, vp, u64, vp]
        lib.lslab_gen_zipf_ranks.argtypes = [vp, u64, ctypes.c_double, u64, u64, u64, vp]
        for f in ("lslab_gen_keys", "lslab_gen_values", "lslab_gen_ops", "lslab_gen_zipf_ranks"):
            getattr(lib, f).restype = ctypes.c_int
        lib._gen_ready = True
    return lib


def _s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def keys(n, key_bytes=8, first=0, seed=0):
    """n distinct keys: bijective mix of first..first+n-1."""
    out = torch.empty(n, dtype=_DT[key_bytes], device="cuda")
    _lib.check(_lib_gen().lslab_gen_keys(ctypes.c_void_p(out.data_ptr()), key_bytes, first, n,
                                         seed, _s()), "gen_keys")
    return out


def values(n, val_bytes=8, first=0, seed=0):
    out = torch.empty(n, dtype=_DT[val_bytes], device="cuda")
    _lib.check(_lib_gen().lslab_gen_values(ctypes.c_void_p(out.data_ptr()), val_bytes, first, n,
                                           seed, _s()), "gen_values")
    return out


def ops(n, theta, space_log2, p_get, p_put, key_bytes=8, val_bytes=8, seed=0, first_op=0,
        with_values=True):
    """Mixed op stream -> (ops u8, keys, values or None)."""
    p = GenParams(seed, first_op, theta, space_log2, p_get, p_put, key_bytes, val_bytes)
    o = torch.empty(n, dtype=torch.uint8, device="cuda")
    k = torch.empty(n, dtype=_DT[key_bytes], device="cuda")
    v = torch.empty(n, dtype=_DT[val_bytes], device="cuda") if with_values else None
    _lib.check(_lib_gen().lslab_gen_ops(ctypes.byref(p), ctypes.c_void_p(o.data_ptr()),
                                        ctypes.c_void_p(k.data_ptr()),
                                        ctypes.c_void_p(v.data_ptr()) if v is not None else None,
                                        n, _s()), "gen_ops")
    return o, k, v


def zipf_ranks(n, theta, space, seed=0, first=0):
    out = torch.empty(n, dtype=torch.int64, device="cuda")
    _lib.check(_lib_gen().lslab_gen_zipf_ranks(ctypes.c_void_p(out.data_ptr()), n, theta, space,
                                               seed, first, _s()), "gen_zipf")
    return out
