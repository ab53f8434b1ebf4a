"""GPU parity of the KVCG hot/cold split (lslab_kvcg_*): host-memory batches
whose hot-key ops run on host workers and the rest on the table.

Batches of distinct keys must give the sequential replay's results op for op,
across hot-set changes (promotion reads the table, demotion writes back), and
after a flush the table must hold the replay's final contents.  Batches with
repeated keys are checked per key against some serial order by the oracle's
checker.
"""
import numpy as np
import pytest

from oracle.oracle import OP_GET, OP_PUT, OP_REMOVE, Replay
from test_gpu_parity import dump_sorted, make, oracle_dump_sorted, run_batch, say

pytestmark = pytest.mark.gpu


def _mask(a, b):
    return a & np.uint64(0xFFFFFFFF) if b == 4 else a


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (True, 8, 4),
                                          (False, 8, 0), (False, 4, 0)])
@pytest.mark.parametrize("engine", [0, 1])
def test_kvcg_distinct_matches_replay(lslab, oracle_mod, is_map, kb, vb, engine):
    say("kvcg_distinct", is_map, kb, vb, engine)
    rng = np.random.default_rng(21 + engine)
    n_log2, pool = 8, 1 << 13
    tbl = make(lslab, is_map, kb, vb, n_log2, pool, engine, lanes=8)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=pool)
    kv = lslab.Kvcg(tbl, hot_capacity=512, max_batch=4096, threads=4)
    space = _mask(rng.permutation(40000).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
                  + np.uint64(7), kb)
    space = np.unique(space)
    total = 0
    for step in range(8):
        if step in (1, 4, 6):
            # hot sets overlap the previous one; some keys are absent from the table
            hot = rng.choice(space[:3000], size=400 if step != 6 else 0, replace=False)
            kv.set_hot(hot)
        n = 6000
        keys = rng.choice(space[:20000], size=n, replace=False)
        if step in (1, 4):
            keys[:300] = hot[:300]  # make sure hot ops occur
            keys = np.unique(keys)
            rng.shuffle(keys)
            n = len(keys)
        ops = (np.full(n, OP_PUT, np.uint8) if step == 0 else
               rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=n, p=[0.4, 0.4, 0.2]).astype(np.uint8))
        vals = _mask(rng.integers(1, 2**63, size=n, dtype=np.uint64), vb if is_map else 8)
        f, v = kv.mixed(ops, keys, vals if is_map else None)
        rf, rv, fail = ref.apply(ops, keys, vals if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f, rf)
        np.testing.assert_array_equal(v, rv)
        total += n
    st = kv.stats()
    assert st["hot_ops"] > 0 and st["hot_ops"] + st["cold_ops"] == total
    assert st["batches"] == 8 and st["hot_keys"] == 0
    kv.set_hot(space[:500])
    kv.flush()
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rvv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rvv)
    kv.close()


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (False, 8, 0)])
@pytest.mark.parametrize("engine", [0, 1])
def test_kvcg_conflicts_serial_order(lslab, oracle_mod, is_map, kb, vb, engine):
    """Zipf keys with repeats inside each batch, hot set from a sample: every
    key's results must match some serial order of its ops."""
    say("kvcg_conflicts", is_map, kb, vb, engine)
    rng = np.random.default_rng(33 + engine)
    n_log2 = 6
    tbl = make(lslab, is_map, kb, vb, n_log2, 1 << 14, engine, lanes=8)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=1 << 14)
    kv = lslab.Kvcg(tbl, hot_capacity=64, max_batch=1 << 14, threads=3)
    space = 3000
    p = 1.0 / np.arange(1, space + 1) ** 1.1
    p /= p.sum()
    perm = rng.permutation(space).astype(np.uint64) * np.uint64(0x9E3779B1) + np.uint64(1)
    kv.sample_hot(perm[rng.choice(space, size=20000, p=p)], 64)
    assert kv.stats()["hot_keys"] == 64
    for bi in range(5):
        n = 20000
        keys = perm[rng.choice(space, size=n, p=p)]
        ops = rng.choice([0, 1, 2], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
        vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
        f, v = kv.mixed(ops, keys, vals if is_map else None)
        pf, pv = kv.mixed(np.zeros(n, np.uint8), keys, vals if is_map else None)
        bad, first = ref.check_batch(ops, keys, vals, f, v, pf, pv)
        assert bad == 0, (bi, first)
    st = kv.stats()
    assert st["hot_ops"] > st["cold_ops"] // 4  # the sample found the hot keys
    kv.flush()
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)
    kv.close()


def test_kvcg_sample_hot_picks_top_k(lslab, oracle_mod):
    """sample_hot keeps the k most frequent keys, ties to the smaller key: a
    batch of those k keys then runs entirely on the host tier, and a batch of
    the next keys entirely on the table."""
    say("kvcg_sample_top_k")
    tbl = make(lslab, True, 8, 8, 6, 1 << 10, 1, lanes=8)
    kv = lslab.Kvcg(tbl, hot_capacity=16, threads=2)
    keys = np.arange(100, 140, dtype=np.uint64)
    counts = np.concatenate([np.arange(40, 10, -1), np.full(10, 5)])  # 40..11, then ties
    sample = np.repeat(keys, counts)
    np.random.default_rng(0).shuffle(sample)
    kv.sample_hot(sample, 16)
    z = np.zeros(16, np.uint8)
    kv.mixed(z, keys[:16], z.astype(np.uint64))
    assert kv.stats()["hot_ops"] == 16 and kv.stats()["cold_ops"] == 0
    kv.mixed(z, keys[16:32], z.astype(np.uint64))
    assert kv.stats()["hot_ops"] == 16 and kv.stats()["cold_ops"] == 16
    with pytest.raises(lslab.LslabError):
        kv.set_hot(np.arange(1, 18, dtype=np.uint64))  # over capacity
    kv.close()
    # ties: 30 keys of distinct counts, then 10 keys of count 5 -> smaller keys first
    kv2 = lslab.Kvcg(tbl, hot_capacity=35, threads=2)
    kv2.sample_hot(sample, 35)
    kv2.mixed(np.zeros(35, np.uint8), keys[:35], np.zeros(35, np.uint64))
    assert kv2.stats()["hot_ops"] == 35 and kv2.stats()["cold_ops"] == 0
    kv2.mixed(np.zeros(5, np.uint8), keys[35:], np.zeros(5, np.uint64))
    assert kv2.stats()["cold_ops"] == 5
    kv2.close()


def test_unknown_opcode_reads_as_get(lslab, oracle_mod):
    """Opcodes outside GET/PUT/REMOVE act as GET on every engine and on a
    reclaim table (none reaches the internal ones)."""
    rng = np.random.default_rng(9)
    for engine, reclaim in ((0, 0), (1, 0), (0, 2), (1, 2)):
        say("unknown_opcode", engine, reclaim)
        tun = lslab.tuning(lanes_per_op=4, engine=engine, max_subbatch=1 << 14, reclaim=reclaim)
        tbl = lslab.Map(8, 8, 4, 1 << 10, tun)
        ref = Replay(True, 8, 8, 4, pool_slabs=1 << 10, reclaim=reclaim == 2)
        keys = np.arange(1, 2001, dtype=np.uint64)
        for ops in (np.full(2000, OP_PUT, np.uint8),
                    rng.choice([0, 1, 2, 3, 7, 255], size=2000).astype(np.uint8)):
            vals = rng.integers(1, 2**40, size=2000, dtype=np.uint64)
            f, v = run_batch(tbl, True, 8, 8, ops, keys, vals)
            rops = np.where(ops > 2, 0, ops).astype(np.uint8)
            rf, rv, fail = ref.apply(rops, keys, vals)
            np.testing.assert_array_equal(f, rf)
            np.testing.assert_array_equal(v, rv)
        assert tbl.status() == 0
