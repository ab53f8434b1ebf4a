"""CPU tests of the KVCG split's host logic (lslab_amd/csrc/kvcg.cpp), built
against a sequential host table (tests/cpp/kvcg_host_stub.cpp) instead of the
GPU.  Both sides then run their ops in batch order, so every batch -- repeated
keys included -- must give exactly the oracle's sequential replay, across
hot-set changes, slices and worker counts; after a flush the host table must
hold the replay's contents."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

from oracle.oracle import OP_GET, OP_PUT, OP_REMOVE, Replay

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "cpp", "libkvcg_host.so")


class KvcgStats(ctypes.Structure):
    _fields_ = [(f, ctypes.c_uint64) for f in
                ("hot_keys", "hot_capacity", "hot_ops", "cold_ops", "batches")]


@pytest.fixture(scope="module")
def lib():
    subprocess.run(["make", "-s", "-C", os.path.join(HERE, "cpp"), "libkvcg_host.so"],
                   check=True)
    lib = ctypes.CDLL(LIB)
    vp, u64 = ctypes.c_void_p, ctypes.c_uint64
    lib.stub_table_create.restype = vp
    lib.stub_table_create.argtypes = [ctypes.c_int] * 3
    lib.stub_table_destroy.argtypes = [vp]
    lib.stub_table_size.restype = u64
    lib.stub_table_size.argtypes = [vp]
    lib.stub_table_dump.restype = u64
    lib.stub_table_dump.argtypes = [vp, vp, vp, u64]
    lib.lslab_kvcg_create.argtypes = [ctypes.POINTER(vp), vp, u64, u64, ctypes.c_int]
    lib.lslab_kvcg_destroy.argtypes = [vp]
    lib.lslab_kvcg_set_hot.argtypes = [vp, vp, u64, vp]
    lib.lslab_kvcg_sample_hot.argtypes = [vp, vp, u64, u64, vp]
    lib.lslab_kvcg_mixed.argtypes = [vp, vp, vp, vp, vp, u64, vp]
    lib.lslab_kvcg_flush.argtypes = [vp, vp]
    lib.lslab_kvcg_stats_get.argtypes = [vp, ctypes.POINTER(KvcgStats)]
    return lib


def _p(a):
    return ctypes.c_void_p(a.ctypes.data) if a is not None else None


class Split:
    def __init__(self, lib, is_map, kb, vb, hot_cap, max_batch, threads):
        self.lib, self.is_map, self.kb, self.vb = lib, is_map, kb, vb if is_map else 0
        self.kdt = np.uint32 if kb == 4 else np.uint64
        self.vdt = np.uint32 if vb == 4 else np.uint64
        self.t = lib.stub_table_create(int(is_map), kb, vb)
        h = ctypes.c_void_p()
        assert lib.lslab_kvcg_create(ctypes.byref(h), self.t, hot_cap, max_batch, threads) == 0
        self.h = h

    def mixed(self, ops, keys, vals):
        n = len(keys)
        o = np.ascontiguousarray(ops, np.uint8)
        k = np.ascontiguousarray(keys.astype(self.kdt))
        if self.is_map:
            v = np.ascontiguousarray(vals.astype(self.vdt))
            out = np.zeros((n, 2), self.vdt)
        else:
            v = None
            out = np.zeros(n, np.uint8)
        assert self.lib.lslab_kvcg_mixed(self.h, _p(o), _p(k), _p(v), _p(out), n, None) == 0
        if self.is_map:
            f = out.view(np.uint8).reshape(n, 2 * self.vb)[:, 0].copy()
            return f, out[:, 1].astype(np.uint64)
        return out, np.zeros(n, np.uint64)

    def set_hot(self, keys):
        k = np.ascontiguousarray(np.asarray(keys).astype(self.kdt))
        return self.lib.lslab_kvcg_set_hot(self.h, _p(k), len(k), None)

    def stats(self):
        s = KvcgStats()
        assert self.lib.lslab_kvcg_stats_get(self.h, ctypes.byref(s)) == 0
        return {f: getattr(s, f) for f, _ in KvcgStats._fields_}

    def table_dump(self):
        n = self.lib.stub_table_size(self.t)
        k = np.zeros(max(n, 1), np.uint64)
        v = np.zeros(max(n, 1), np.uint64)
        m = self.lib.stub_table_dump(self.t, _p(k), _p(v), n)
        o = np.argsort(k[:m])
        return k[:m][o], (v[:m][o] if self.is_map else np.zeros(m, np.uint64))

    def close(self):
        assert self.lib.lslab_kvcg_destroy(self.h) == 0
        self.lib.stub_table_destroy(self.t)


def oracle_sorted(ref):
    k, v = ref.dump()
    o = np.argsort(k)
    return k[o], v[o]


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (True, 8, 4),
                                          (False, 8, 0), (False, 4, 0)])
@pytest.mark.parametrize("threads,max_batch", [(1, 1 << 16), (3, 700), (8, 4096)])
def test_split_equals_sequential_replay(lib, oracle_mod, is_map, kb, vb, threads, max_batch):
    rng = np.random.default_rng(kb * 7 + vb + threads)
    s = Split(lib, is_map, kb, vb, 300, max_batch, threads)
    ref = Replay(is_map, kb, vb, 6, pool_slabs=1 << 14)
    space = np.unique(rng.integers(1, 2**32 if kb == 4 else 2**63, size=3000, dtype=np.uint64))
    p = 1.0 / np.arange(1, len(space) + 1) ** 1.1
    p /= p.sum()
    total = 0
    for step in range(10):
        if step in (1, 3, 6, 8):
            hot = rng.choice(space[:800], size=[250, 300, 0, 120][(1, 3, 6, 8).index(step)],
                             replace=False)
            assert s.set_hot(hot) == 0
        n = 5000
        keys = space[rng.choice(len(space), size=n, p=p)]  # repeats within the batch
        ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE, 7], size=n,
                         p=[0.45, 0.35, 0.15, 0.05]).astype(np.uint8)
        vals = rng.integers(1, 2**32 if vb == 4 else 2**63, size=n, dtype=np.uint64)
        f, v = s.mixed(ops, keys, vals)
        rf, rv, fail = ref.apply(np.where(ops > 2, 0, ops).astype(np.uint8), keys,
                                 vals if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f, rf)
        np.testing.assert_array_equal(v, rv)
        total += n
    st = s.stats()
    assert st["hot_ops"] > 0 and st["hot_ops"] + st["cold_ops"] == total
    assert st["hot_keys"] == 120
    assert s.set_hot(space[:301]) != 0  # over capacity, nothing changes
    assert s.stats()["hot_keys"] == 120
    assert lib.lslab_kvcg_flush(s.h, None) == 0
    k, vv = s.table_dump()
    rk, rvv = oracle_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rvv)
    s.close()


def test_sample_hot_top_k_and_destroy_flushes(lib, oracle_mod):
    s = Split(lib, True, 8, 8, 16, 0, 2)
    keys = np.arange(100, 140, dtype=np.uint64)
    counts = np.concatenate([np.arange(40, 10, -1), np.full(10, 5)])
    sample = np.repeat(keys, counts)
    np.random.default_rng(0).shuffle(sample)
    assert lib.lslab_kvcg_sample_hot(s.h, _p(sample), len(sample), 16, None) == 0
    assert lib.lslab_kvcg_sample_hot(s.h, _p(sample), len(sample), 17, None) != 0
    f, _ = s.mixed(np.full(16, OP_PUT, np.uint8), keys[:16], keys[:16] * 3)
    assert s.stats()["hot_ops"] == 16 and s.stats()["cold_ops"] == 0
    assert lib.stub_table_size(s.t) == 0  # hot puts stay on the host tier...
    h = s.h
    assert lib.lslab_kvcg_destroy(h) == 0  # ...until destroy flushes them
    s.h = None
    assert lib.stub_table_size(s.t) == 16
    k = np.zeros(16, np.uint64)
    v = np.zeros(16, np.uint64)
    lib.stub_table_dump(s.t, _p(k), _p(v), 16)
    o = np.argsort(k)
    np.testing.assert_array_equal(k[o], keys[:16])
    np.testing.assert_array_equal(v[o], keys[:16] * 3)
    lib.stub_table_destroy(s.t)
