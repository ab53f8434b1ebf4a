"""ctypes front end of the lslab oracle (test infrastructure only).

The oracle is a sequential host replay of the LSlab slab-list map and set
(see lslab_oracle.c for the reference citations).  Only tests/, the smoke
check and the benchmark's CPU-baseline leg import this module, and only to
check or to time the host baseline; the product never does.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

OP_GET = 0
OP_PUT = 1
OP_REMOVE = 2


def build():
    """Compile the oracle's C sources with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "all"], check=True)


def _lib():
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "liblslab_oracle.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp, u64, i64, u8p, u64p = (ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int64,
                               ctypes.c_void_p, ctypes.c_void_p)
    lib.orc_create.restype = vp
    lib.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_uint, u64,
                               ctypes.c_int]
    lib.orc_destroy.argtypes = [vp]
    lib.orc_apply.restype = i64
    lib.orc_apply.argtypes = [vp, u64, u8p, u64p, u64p, u8p, u64p]
    lib.orc_lookup.restype = ctypes.c_int
    lib.orc_lookup.argtypes = [vp, u64, ctypes.POINTER(ctypes.c_uint64)]
    for f in ("orc_present", "orc_slabs_used", "orc_max_chain"):
        getattr(lib, f).restype = u64
        getattr(lib, f).argtypes = [vp]
    lib.orc_set_hash.restype = None
    lib.orc_set_hash.argtypes = [vp, ctypes.c_int]
    lib.orc_set_reclaim.restype = ctypes.c_int
    lib.orc_set_reclaim.argtypes = [vp, ctypes.c_int]
    lib.orc_chain_hist.restype = None
    lib.orc_chain_hist.argtypes = [vp, u64p, ctypes.c_uint]
    lib.orc_dump.restype = u64
    lib.orc_dump.argtypes = [vp, u64p, u64p, u64]
    lib.orc_check_batch.restype = i64
    lib.orc_check_batch.argtypes = [vp, u64, u8p, u64p, u64p, u8p, u64p, u8p, u64p,
                                    ctypes.POINTER(ctypes.c_uint64)]
    _LIB = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _u64(a, n):
    a = np.ascontiguousarray(a).astype(np.uint64, copy=False)
    assert a.shape == (n,)
    return a


class Replay:
    """Sequential replay table: map (is_map=True) or set."""

    def __init__(self, is_map, key_bytes, val_bytes, n_log2, pool_slabs=10000, slab_keys=31,
                 mix_hash=False, reclaim=False):
        self.is_map = bool(is_map)
        self.h = _lib().orc_create(int(is_map), key_bytes, val_bytes if is_map else 0,
                                   n_log2, pool_slabs, slab_keys)
        if not self.h:
            raise ValueError("orc_create rejected the configuration")
        if mix_hash:
            _lib().orc_set_hash(self.h, 1)
        if reclaim and not _lib().orc_set_reclaim(self.h, 1):
            raise MemoryError("orc_set_reclaim")

    def close(self):
        if self.h:
            _lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def apply(self, ops, keys, vals=None):
        """Apply ops in index order; returns (found u8[n], val u64[n], failed)."""
        n = len(keys)
        ops = np.ascontiguousarray(ops, dtype=np.uint8)
        keys = _u64(keys, n)
        vals = None if vals is None else _u64(vals, n)
        found = np.zeros(n, np.uint8)
        out = np.zeros(n, np.uint64)
        failed = _lib().orc_apply(self.h, n, _p(ops), _p(keys), _p(vals), _p(found), _p(out))
        return found, out, int(failed)

    def lookup(self, key):
        v = ctypes.c_uint64(0)
        f = _lib().orc_lookup(self.h, int(key), ctypes.byref(v))
        return bool(f), int(v.value)

    @property
    def present(self):
        return int(_lib().orc_present(self.h))

    @property
    def slabs_used(self):
        return int(_lib().orc_slabs_used(self.h))

    def chain_hist(self, nbins=16):
        h = np.zeros(nbins, np.uint64)
        _lib().orc_chain_hist(self.h, _p(h), nbins)
        return [int(x) for x in h]

    @property
    def max_chain(self):
        return int(_lib().orc_max_chain(self.h))

    def dump(self):
        n = self.present
        k = np.zeros(max(n, 1), np.uint64)
        v = np.zeros(max(n, 1), np.uint64)
        c = _lib().orc_dump(self.h, _p(k), _p(v), n)
        assert c == n
        return k[:n], v[:n]

    def check_batch(self, ops, keys, vals, found, got, post_found=None, post_val=None):
        """Serial-order check of observed results; adopts the end states.

        Returns (number of keys with no valid serial order, first bad op index
        or None)."""
        n = len(keys)
        ops = np.ascontiguousarray(ops, dtype=np.uint8)
        keys = _u64(keys, n)
        vals = _u64(vals if vals is not None else np.zeros(n, np.uint64), n)
        found = np.ascontiguousarray(found, dtype=np.uint8)
        got = _u64(got, n)
        pf = None if post_found is None else np.ascontiguousarray(post_found, dtype=np.uint8)
        pv = None if post_val is None else _u64(post_val, n)
        fb = ctypes.c_uint64(0)
        bad = _lib().orc_check_batch(self.h, n, _p(ops), _p(keys), _p(vals), _p(found),
                                     _p(got), _p(pf), _p(pv), ctypes.byref(fb))
        return int(bad), (None if fb.value == 2**64 - 1 else int(fb.value))


_MT = None


def _mt_lib():
    global _MT
    if _MT is not None:
        return _MT
    path = os.path.join(_HERE, "libmt_baseline.so")
    if not os.path.exists(path):
        build()
    lib = ctypes.CDLL(path)
    vp = ctypes.c_void_p
    lib.mtb_create.restype = vp
    lib.mtb_create.argtypes = [ctypes.c_int, ctypes.c_uint, ctypes.c_int]
    lib.mtb_destroy.argtypes = [vp]
    lib.mtb_run.restype = ctypes.c_double
    lib.mtb_run.argtypes = [vp, ctypes.c_uint64, vp, vp, vp, vp, vp, vp]
    lib.mtb_size.restype = ctypes.c_uint64
    lib.mtb_size.argtypes = [vp]
    lib.mtb_reserve.argtypes = [vp, ctypes.c_uint64]
    _MT = lib
    return lib


class MtBaseline:
    """The multithreaded host baseline (mt_baseline.cpp): std::unordered_map
    shards by bucket bits, a shared_mutex per shard, T threads, a barrier
    between batches.  Timing infrastructure, checked against Replay."""

    def __init__(self, threads, n_log2, is_map=True):
        self._lib = _mt_lib()
        self.threads = threads
        self._h = self._lib.mtb_create(threads, n_log2, 1 if is_map else 0)

    def close(self):
        if self._h:
            self._lib.mtb_destroy(self._h)
            self._h = None

    def __del__(self):
        self.close()

    def run(self, offsets, ops, keys, vals=None, want_results=False):
        """Applies batches [offsets[b], offsets[b+1]); returns (seconds, found, vals)."""
        offsets = np.ascontiguousarray(offsets, dtype=np.uint64)
        n = int(offsets[-1])
        ops = np.ascontiguousarray(ops, dtype=np.uint8)
        keys = _u64(keys, n)
        vals = None if vals is None else _u64(vals, n)
        f = np.zeros(n, np.uint8) if want_results else None
        o = np.zeros(n, np.uint64) if want_results else None
        dt = self._lib.mtb_run(self._h, len(offsets) - 1, _p(offsets), _p(ops), _p(keys), _p(vals),
                               _p(f), _p(o))
        return dt, f, o

    def reserve(self, total):
        """Sizes the shards for `total` keys."""
        self._lib.mtb_reserve(self._h, total)

    @property
    def size(self):
        return int(self._lib.mtb_size(self._h))
