"""Python mirror of lslab::map<K,V> and lslab::set<K> over the C ABI.

Method names, argument order and result meaning follow the reference's host
batch API (include/lslab/map.h:189-220, include/lslab/set.h:92-123).  Device
memory and streams come from torch (plumbing only): inputs and outputs are
CUDA tensors, calls go on torch's current stream.  Keys and values are raw
integer bit patterns: 4-byte widths use int32/uint32 tensors, 8-byte widths
int64/uint64 tensors.
"""
import ctypes

import torch

from . import _lib

_DT = {4: torch.int32, 8: torch.int64}


def _stream(stream):
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _dev(t, nbytes, what):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError(f"{what} must be a CUDA tensor")
    if t.element_size() != nbytes or t.is_floating_point():
        raise TypeError(f"{what} must hold {nbytes}-byte integers, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError(f"{what} must be contiguous")
    return t


def _ops(t, n):
    t = _dev(t, 1, "ops")
    if t.numel() != n:
        raise ValueError("ops and keys differ in length")
    return t


def _mem(t):
    if not isinstance(t, torch.Tensor) or not t.is_cuda:
        raise TypeError("slab_mem must be a CUDA tensor")
    if not t.is_contiguous():
        raise ValueError("slab_mem must be contiguous")
    return ctypes.c_void_p(t.data_ptr()), t.numel() * t.element_size()


def slab_bytes(is_map, key_bytes, value_bytes, n_log2, pool_slabs):
    """Bytes of the slab array (heads + pool) a caller-provided slab_mem must hold
    (lslab_slab_bytes); 0 for an invalid shape."""
    return _lib.load().lslab_slab_bytes(int(bool(is_map)), key_bytes, value_bytes, n_log2,
                                        pool_slabs)


def tuning(lanes_per_op=4, engine=_lib.ENGINE_PLAIN, block_threads=256, max_blocks=0,
           max_subbatch=1 << 20, combine=_lib.COMBINE_BLOCK, hash=0, reclaim=0, spin_ms=0,
           slot_heads=0, runs=0):
    """Launch shapes and table policies (lslab_tuning).  hash and reclaim are
    fixed at creation; 0 in set_tuning keeps them.  spin_ms bounds every device
    spin (0 = 10 s).  slot_heads: SLOT_HEADS_ON / SLOT_HEADS_OFF, 0 = on while
    engine is two-phase.  runs (exec_batches): RUNS_ON / RUNS_OFF, 0 = found out
    per call on the device."""
    return _lib.Tuning(lanes_per_op, engine, block_threads, max_blocks, max_subbatch, combine,
                       hash, reclaim, spin_ms, slot_heads, runs)


class _Table:
    is_map = False

    def __init__(self):
        self._h = None

    def close(self):
        if self._h:
            _lib.check(_lib.load().lslab_destroy(self._h), "destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    @property
    def handle(self):
        return self._h

    def set_tuning(self, tun):
        _lib.check(_lib.load().lslab_set_tuning(self._h, ctypes.byref(tun)), "set_tuning")

    def get_tuning(self):
        t = _lib.Tuning()
        _lib.check(_lib.load().lslab_get_tuning(self._h, ctypes.byref(t)), "get_tuning")
        return t

    def status(self, stream=None):
        """Synchronises and returns the sticky device status (0 = ok), clearing it."""
        return _lib.load().lslab_status(self._h, _stream(stream))

    def stats(self, stream=None):
        st = _lib.Stats()
        _lib.check(_lib.load().lslab_stats_get(self._h, ctypes.byref(st), _stream(stream)), "stats")
        return {f: getattr(st, f) for f, _ in _lib.Stats._fields_}

    def chain_histogram(self, nbins=_lib.CHAIN_BINS, stream=None):
        """Buckets by chain length: entry i counts chains of i+1 slabs (the last
        entry also counts longer chains)."""
        h = (ctypes.c_uint64 * nbins)()
        _lib.check(_lib.load().lslab_chain_histogram(self._h, h, nbins, _stream(stream)),
                   "chain_histogram")
        return [int(x) for x in h]

    def buckets(self):
        return 1 << self.n_log2


class Map(_Table):
    """lslab::map<K,V> (reference include/lslab/map.h:24-234)."""

    is_map = True

    def __init__(self, key_bytes=8, value_bytes=8, n_log2=10, pool_slabs=10000, tun=None,
                 slab_mem=None):
        """slab_mem: optional caller-owned CUDA tensor holding the slab array (at least
        slab_bytes(...) bytes, 256-B aligned); the table keeps a reference to it."""
        super().__init__()
        self.kb, self.vb, self.n_log2 = key_bytes, value_bytes, n_log2
        h = ctypes.c_void_p()
        tp = ctypes.byref(tun) if tun else None
        lib = _lib.load()
        if slab_mem is None:
            rc = lib.lslab_map_create(ctypes.byref(h), key_bytes, value_bytes, n_log2,
                                      pool_slabs, tp)
        else:
            ptr, nbytes = _mem(slab_mem)
            rc = lib.lslab_map_create_at(ctypes.byref(h), key_bytes, value_bytes, n_log2,
                                         pool_slabs, ptr, nbytes, tp)
        _lib.check(rc, "map_create")
        self._slab_mem = slab_mem
        self._h = h

    def _pairs(self, n):
        return torch.empty((n, 2), dtype=_DT[self.vb], device="cuda")

    def split_pairs(self, raw):
        """(found u8[n], value[n]) views of a pair<bool,V> result buffer."""
        n = raw.shape[0]
        found = raw.view(torch.uint8).view(n, 2 * self.vb)[:, 0]
        return found, raw[:, 1]

    def put(self, keys, values, out=None, stream=None):
        """map::put(K*, V*, V* out, n, stream): out[i] = previous value (0 if absent)."""
        n = keys.numel()
        _dev(keys, self.kb, "keys")
        _dev(values, self.vb, "values")
        if out is None:
            out = torch.empty(n, dtype=_DT[self.vb], device="cuda")
        _lib.check(_lib.load().lslab_map_put(self._h, _ptr(keys), _ptr(values), _ptr(out), n,
                                             _stream(stream)), "map_put")
        return out

    def put_pairs(self, pairs, out=None, stream=None):
        """map::put(pair<K,V>*, V* out, n, stream); pairs is a byte tensor [n, 2*max(kb,vb)]."""
        n = pairs.shape[0]
        if out is None:
            out = torch.empty(n, dtype=_DT[self.vb], device="cuda")
        _lib.check(_lib.load().lslab_map_put_pairs(self._h, _ptr(pairs), _ptr(out), n,
                                                   _stream(stream)), "map_put_pairs")
        return out

    def get(self, keys, stream=None, raw=False):
        """map::get(K*, pair<bool,V>* out, n, stream) -> (found, value)."""
        n = keys.numel()
        _dev(keys, self.kb, "keys")
        out = self._pairs(n)
        _lib.check(_lib.load().lslab_map_get(self._h, _ptr(keys), _ptr(out), n, _stream(stream)),
                   "map_get")
        return out if raw else self.split_pairs(out)

    def remove(self, keys, stream=None):
        n = keys.numel()
        _dev(keys, self.kb, "keys")
        out = self._pairs(n)
        _lib.check(_lib.load().lslab_map_remove(self._h, _ptr(keys), _ptr(out), n,
                                                _stream(stream)), "map_remove")
        return self.split_pairs(out)

    def mixed(self, ops, keys, values, stream=None, raw=False):
        n = keys.numel()
        _ops(ops, n)
        _dev(keys, self.kb, "keys")
        _dev(values, self.vb, "values")
        out = self._pairs(n)
        _lib.check(_lib.load().lslab_map_mixed(self._h, _ptr(ops), _ptr(keys), _ptr(values),
                                               _ptr(out), n, _stream(stream)), "map_mixed")
        return out if raw else self.split_pairs(out)

    def exec_batches(self, offsets, ops, keys, values, out=None, stream=None):
        """Runs len(offsets)-1 batches in one persistent launch; returns (found, value)."""
        n = keys.numel()
        _ops(ops, n)
        _dev(keys, self.kb, "keys")
        _dev(values, self.vb, "values")
        _dev(offsets, 8, "offsets")
        if out is None:
            out = self._pairs(n)
        nb = offsets.numel() - 1
        _lib.check(_lib.load().lslab_exec_batches(self._h, nb, _ptr(offsets), _ptr(ops),
                                                  _ptr(keys), _ptr(values), _ptr(out),
                                                  _stream(stream)), "exec_batches")
        return self.split_pairs(out)

    def serve(self, req_offsets, arrival_ticks, ops, keys, values, req_reserve=None,
              max_tick_ops=0, max_tick_reqs=0, stream=None):
        """Device-resident request loop (lslab_serve).  Returns a dict with the
        per-op results, per-request completion ticks and tick ids, and the
        loop's counters."""
        n = keys.numel()
        nreq = arrival_ticks.numel()
        _ops(ops, n)
        _dev(keys, self.kb, "keys")
        _dev(values, self.vb, "values")
        _dev(req_offsets, 8, "req_offsets")
        _dev(arrival_ticks, 8, "arrival_ticks")
        if req_reserve is not None:
            _dev(req_reserve, 4, "req_reserve")
        if req_offsets.numel() != nreq + 1:
            raise ValueError("req_offsets must have nreq + 1 entries")
        out = self._pairs(n)
        complete = torch.zeros(nreq, dtype=torch.int64, device="cuda")
        tick_of = torch.zeros(nreq, dtype=torch.int32, device="cuda")
        p = _lib.ServeParams(nreq, req_offsets.data_ptr(), arrival_ticks.data_ptr(),
                             req_reserve.data_ptr() if req_reserve is not None else None,
                             complete.data_ptr(), tick_of.data_ptr(), max_tick_ops, max_tick_reqs)
        lib = _lib.load()
        _lib.check(lib.lslab_serve(self._h, ctypes.byref(p), _ptr(ops), _ptr(keys), _ptr(values),
                                   _ptr(out), _stream(stream)), "serve")
        st = _lib.ServeStats()
        _lib.check(lib.lslab_serve_result(self._h, ctypes.byref(st), _stream(stream)),
                   "serve_result")
        f, v = self.split_pairs(out)
        return {"found": f, "value": v, "complete_ticks": complete, "tick_of": tick_of,
                "ticks": st.ticks, "rejected": st.rejected, "served_ops": st.served_ops,
                "elapsed_ticks": st.elapsed_ticks}

    def dump(self, stream=None):
        """All present (key, value) pairs, sorted by key (as unsigned)."""
        st = self.stats(stream)
        cap = max(st["present"], 1)
        k = torch.empty(cap, dtype=_DT[self.kb], device="cuda")
        v = torch.empty(cap, dtype=_DT[self.vb], device="cuda")
        c = ctypes.c_uint64()
        _lib.check(_lib.load().lslab_dump(self._h, _ptr(k), _ptr(v), cap, ctypes.byref(c),
                                          _stream(stream)), "dump")
        n = int(c.value)
        return k[:n], v[:n]


class Set(_Table):
    """lslab::set<K> (reference include/lslab/set.h:21-123)."""

    def __init__(self, key_bytes=8, n_log2=10, pool_slabs=10000, tun=None, slab_mem=None):
        """slab_mem: as for Map."""
        super().__init__()
        self.kb, self.n_log2 = key_bytes, n_log2
        h = ctypes.c_void_p()
        tp = ctypes.byref(tun) if tun else None
        lib = _lib.load()
        if slab_mem is None:
            rc = lib.lslab_set_create(ctypes.byref(h), key_bytes, n_log2, pool_slabs, tp)
        else:
            ptr, nbytes = _mem(slab_mem)
            rc = lib.lslab_set_create_at(ctypes.byref(h), key_bytes, n_log2, pool_slabs, ptr,
                                         nbytes, tp)
        _lib.check(rc, "set_create")
        self._slab_mem = slab_mem
        self._h = h

    def _call(self, fn, keys, stream):
        n = keys.numel()
        _dev(keys, self.kb, "keys")
        out = torch.empty(n, dtype=torch.uint8, device="cuda")
        _lib.check(fn(self._h, _ptr(keys), _ptr(out), n, _stream(stream)), fn.__name__)
        return out

    def insert(self, keys, stream=None):
        return self._call(_lib.load().lslab_set_insert, keys, stream)

    def contains(self, keys, stream=None):
        return self._call(_lib.load().lslab_set_contains, keys, stream)

    def remove(self, keys, stream=None):
        return self._call(_lib.load().lslab_set_remove, keys, stream)

    def mixed(self, ops, keys, stream=None):
        n = keys.numel()
        _ops(ops, n)
        _dev(keys, self.kb, "keys")
        out = torch.empty(n, dtype=torch.uint8, device="cuda")
        _lib.check(_lib.load().lslab_set_mixed(self._h, _ptr(ops), _ptr(keys), _ptr(out), n,
                                               _stream(stream)), "set_mixed")
        return out

    def exec_batches(self, offsets, ops, keys, out=None, stream=None):
        n = keys.numel()
        _ops(ops, n)
        _dev(keys, self.kb, "keys")
        _dev(offsets, 8, "offsets")
        if out is None:
            out = torch.empty(n, dtype=torch.uint8, device="cuda")
        nb = offsets.numel() - 1
        _lib.check(_lib.load().lslab_exec_batches(self._h, nb, _ptr(offsets), _ptr(ops),
                                                  _ptr(keys), None, _ptr(out), _stream(stream)),
                   "exec_batches")
        return out

    def dump(self, stream=None):
        st = self.stats(stream)
        cap = max(st["present"], 1)
        k = torch.empty(cap, dtype=_DT[self.kb], device="cuda")
        c = ctypes.c_uint64()
        _lib.check(_lib.load().lslab_dump(self._h, _ptr(k), None, cap, ctypes.byref(c),
                                          _stream(stream)), "dump")
        return k[: int(c.value)]


class Kvcg:
    """KVCG hot/cold split over a Map or Set (lslab_kvcg_*): host-memory batches,
    hot keys served by host workers, the rest by the table.  Arrays are numpy
    (host); keys/values as uint64 bit patterns of the table's widths."""

    def __init__(self, table, hot_capacity, max_batch=1 << 20, threads=0):
        import numpy as np
        self._np = np
        self.table = table
        self.is_map = table.is_map
        self.kb = table.kb
        self.vb = table.vb if table.is_map else 0
        self._kdt = np.uint32 if self.kb == 4 else np.uint64
        self._vdt = np.uint32 if self.vb == 4 else np.uint64
        h = ctypes.c_void_p()
        _lib.check(_lib.load().lslab_kvcg_create(ctypes.byref(h), table._h, hot_capacity,
                                                 max_batch, threads), "kvcg_create")
        self._h = h

    def _host(self, a, dt):
        """a as a contiguous array of dt bit patterns (no copy when it already is one)."""
        np = self._np
        a = np.asarray(a)
        if a.dtype != dt:
            if a.dtype.kind in "iu" and a.dtype.itemsize == np.dtype(dt).itemsize:
                a = a.view(dt)
            else:
                a = a.astype(np.uint64).astype(dt)
        a = np.ascontiguousarray(a)
        return a, ctypes.c_void_p(a.ctypes.data)

    def set_hot(self, keys, stream=None):
        k, kp = self._host(keys, self._kdt)
        _lib.check(_lib.load().lslab_kvcg_set_hot(self._h, kp, len(k), _stream(stream)),
                   "kvcg_set_hot")

    def sample_hot(self, sample_keys, k, stream=None):
        s, sp = self._host(sample_keys, self._kdt)
        _lib.check(_lib.load().lslab_kvcg_sample_hot(self._h, sp, len(s), k, _stream(stream)),
                   "kvcg_sample_hot")

    def mixed(self, ops, keys, values=None, out=None, stream=None, raw=False):
        """Returns (found u8[n], value u64[n]) as numpy arrays (value 0 for sets),
        or with raw=True the result buffer itself (out, if given: (n, 2) of the
        value width for maps, n bytes for sets)."""
        np = self._np
        n = len(keys)
        o = np.ascontiguousarray(np.asarray(ops, dtype=np.uint8))
        if len(o) != n:
            raise ValueError("ops and keys differ in length")
        k, kp = self._host(keys, self._kdt)
        if self.is_map:
            v, vp = self._host(values if values is not None else np.zeros(n, np.uint64),
                               self._vdt)
            shape, odt = (n, 2), self._vdt
        else:
            vp = None
            shape, odt = (n,), np.uint8
        if out is None:
            out = np.empty(shape, dtype=odt)
        elif out.shape != shape or out.dtype != odt or not out.flags.c_contiguous:
            raise ValueError(f"out must be a contiguous {odt.__name__} array of shape {shape}")
        _lib.check(_lib.load().lslab_kvcg_mixed(self._h, ctypes.c_void_p(o.ctypes.data), kp, vp,
                                                ctypes.c_void_p(out.ctypes.data), n,
                                                _stream(stream)), "kvcg_mixed")
        if raw:
            return out
        if self.is_map:
            found = out.view(np.uint8).reshape(n, 2 * self.vb)[:, 0].copy()
            return found, out[:, 1].astype(np.uint64)
        return out, np.zeros(n, np.uint64)

    def flush(self, stream=None):
        _lib.check(_lib.load().lslab_kvcg_flush(self._h, _stream(stream)), "kvcg_flush")

    def stats(self):
        s = _lib.KvcgStats()
        _lib.check(_lib.load().lslab_kvcg_stats_get(self._h, ctypes.byref(s)), "kvcg_stats")
        return {f: getattr(s, f) for f, _ in _lib.KvcgStats._fields_}

    def close(self):
        if self._h:
            _lib.check(_lib.load().lslab_kvcg_destroy(self._h), "kvcg_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Wide:
    """Wide-key map / set (lslab_wide_*): keys of 4..128 bytes, values of 0 (set)
    .. 128 bytes, both multiples of 4.  Keys and values are CUDA uint8 tensors
    of shape (n, key_bytes) and (n, value_bytes); results are (found uint8[n],
    values uint8[n, value_bytes])."""

    def __init__(self, key_bytes, value_bytes=0, n_log2=10, pool_slabs=10000, hash=0):
        self.kb, self.vb, self.n_log2 = key_bytes, value_bytes, n_log2
        self.is_map = value_bytes > 0
        h = ctypes.c_void_p()
        _lib.check(_lib.load().lslab_wide_create(ctypes.byref(h), key_bytes, value_bytes, n_log2,
                                                 pool_slabs, hash), "wide_create")
        self._h = h

    def _rows(self, t, width, what):
        if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.uint8:
            raise TypeError(f"{what} must be a CUDA uint8 tensor")
        if t.dim() != 2 or t.shape[1] != width or not t.is_contiguous():
            raise ValueError(f"{what} must be a contiguous (n, {width}) tensor")
        return t

    def _outs(self, n):
        found = torch.empty(n, dtype=torch.uint8, device="cuda")
        vals = torch.empty((n, self.vb), dtype=torch.uint8, device="cuda") if self.vb else None
        return found, vals

    def _call(self, name, keys, values=None, ops=None, stream=None):
        n = keys.shape[0]
        self._rows(keys, self.kb, "keys")
        if values is not None:
            self._rows(values, self.vb, "values")
        found, vals = self._outs(n)
        lib = _lib.load()
        if name == "mixed":
            _ops(ops, n)
            rc = lib.lslab_wide_mixed(self._h, _ptr(ops), _ptr(keys), _ptr(values), _ptr(found),
                                      _ptr(vals), n, _stream(stream))
        elif name == "put":
            rc = lib.lslab_wide_put(self._h, _ptr(keys), _ptr(values), _ptr(found), _ptr(vals), n,
                                    _stream(stream))
        else:
            f = lib.lslab_wide_get if name == "get" else lib.lslab_wide_remove
            rc = f(self._h, _ptr(keys), _ptr(found), _ptr(vals), n, _stream(stream))
        _lib.check(rc, "wide_" + name)
        return found, vals

    def put(self, keys, values=None, stream=None):
        """Map put (out = previous value) / set insert (found always 1)."""
        return self._call("put", keys, values, stream=stream)

    def get(self, keys, stream=None):
        return self._call("get", keys, stream=stream)

    def remove(self, keys, stream=None):
        return self._call("remove", keys, stream=stream)

    def mixed(self, ops, keys, values=None, stream=None):
        if self.is_map and values is None:
            values = torch.zeros((keys.shape[0], self.vb), dtype=torch.uint8, device="cuda")
        return self._call("mixed", keys, values, ops=ops, stream=stream)

    def status(self, stream=None):
        return _lib.load().lslab_wide_status(self._h, _stream(stream))

    def stats(self, stream=None):
        s = _lib.Stats()
        _lib.check(_lib.load().lslab_wide_stats(self._h, ctypes.byref(s), _stream(stream)),
                   "wide_stats")
        return {f: getattr(s, f) for f, _ in _lib.Stats._fields_}

    def dump(self, stream=None):
        """(keys (m, key_bytes), values (m, value_bytes) or None), unordered."""
        cap = max(self.stats(stream)["present"], 1)
        k = torch.empty((cap, self.kb), dtype=torch.uint8, device="cuda")
        v = torch.empty((cap, self.vb), dtype=torch.uint8, device="cuda") if self.vb else None
        c = ctypes.c_uint64()
        _lib.check(_lib.load().lslab_wide_dump(self._h, _ptr(k), _ptr(v), cap, ctypes.byref(c),
                                               _stream(stream)), "wide_dump")
        m = int(c.value)
        return k[:m], (v[:m] if v is not None else None)

    def close(self):
        if self._h:
            _lib.check(_lib.load().lslab_wide_destroy(self._h), "wide_destroy")
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
