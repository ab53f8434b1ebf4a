"""GPU parity of the wide-key tables (lslab_wide_*) against the wide oracle:
the reference's legacy 128-B-key tests, random streams op for op with the
final contents and slab accounting, keys whose tags collide, pool
exhaustion, and per-key serial order under same-key conflicts."""
import struct

import numpy as np
import pytest
import torch

from oracle.oracle import Replay
from oracle.wide_oracle import OP_GET, OP_PUT, OP_REMOVE, WideReplay, wide_hash
from test_gpu_parity import say

pytestmark = pytest.mark.gpu


def dev_rows(rows, width):
    a = np.frombuffer(b"".join(rows), dtype=np.uint8).reshape(len(rows), width) if rows else \
        np.zeros((0, width), np.uint8)
    return torch.from_numpy(a.copy()).cuda()


def host_rows(t):
    a = t.cpu().numpy()
    return [a[i].tobytes() for i in range(a.shape[0])]


def key128(i):
    return struct.pack("<Q", i) + bytes(120)


def test_wide_put_remove_128b_kats(lslab, oracle_mod):
    """PutRemoveTest_128B and PutRemoveTest_128Bto128B
    (test/slabunified_test.cu.old:386-541) on the device."""
    for vb in (8, 128):
        say("wide_legacy_kat", vb)
        w = lslab.Wide(128, vb, 10, 10000)
        keys = dev_rows([key128(j + 1) for j in range(1000)], 128)
        for rep in range(100):
            if vb == 8:
                vals = [struct.pack("<Q", (rep << 32) | (j + 1)) for j in range(1000)]
            else:
                vals = [key128(rep)] * 1000
            dv = dev_rows(vals, vb)
            f, old = w.put(keys, dv)
            assert int(f.sum()) == 0 and int(old.sum()) == 0, rep
            f, old = w.remove(keys)
            assert int(f.sum()) == 1000, rep
            assert torch.equal(old, dv), rep
        assert w.status() == 0
        st = w.stats()
        assert st["present"] == 0 and st["slabs_used"] == 0
        w.close()


@pytest.mark.parametrize("kb,vb,mix", [(128, 8, False), (128, 128, True), (16, 0, False),
                                       (4, 4, False), (36, 12, True), (64, 0, True)])
def test_wide_random_stream_matches_replay(lslab, oracle_mod, kb, vb, mix):
    """Batches of distinct keys (random get/put/remove) against the replay op
    for op; keys share their first 8 bytes in groups so that identity-bucket
    chains grow long.  Slab accounting after the fill, then present count and
    dump against the oracle."""
    say("wide_stream", kb, vb, mix)
    rng = np.random.default_rng(kb * 131 + vb + mix)
    n_log2, pool = 4, 600
    w = lslab.Wide(kb, vb, n_log2, pool, hash=2 if mix else 0)
    r = WideReplay(kb, vb, n_log2, pool, mix_hash=mix)
    space = []
    for i in range(6000):
        if kb >= 8:
            head = struct.pack("<Q", int(rng.integers(0, 40)))
            tail = rng.integers(0, 256, size=kb - 8, dtype=np.uint8).tobytes()
            space.append(head + tail)
        else:
            space.append(struct.pack("<I", int(rng.integers(0, 2**32))))
    space = list(dict.fromkeys(space))
    for step in range(5):
        idx = rng.choice(len(space), size=2500, replace=False)
        keys = [space[i] for i in idx]
        ops = (np.full(len(keys), OP_PUT, np.uint8) if step == 0 else
               rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=len(keys)).astype(np.uint8))
        vals = [rng.integers(0, 256, size=vb, dtype=np.uint8).tobytes() for _ in keys]
        f, old = w.mixed(torch.from_numpy(ops).cuda(), dev_rows(keys, kb),
                         dev_rows(vals, vb) if vb else None)
        rf, rold, fail = r.apply(ops, keys, vals if vb else None)
        assert fail == 0
        np.testing.assert_array_equal(f.cpu().numpy(), np.array(rf, np.uint8))
        if vb:
            assert host_rows(old) == rold
        # get-only call (lock-free path) agrees too
        f2, v2 = w.get(dev_rows(keys, kb))
        rf2, rv2, _ = r.apply(np.zeros(len(keys), np.uint8), keys)
        np.testing.assert_array_equal(f2.cpu().numpy(), np.array(rf2, np.uint8))
        if vb:
            assert host_rows(v2) == rv2
        if step == 0:
            # Pool use is order-free until an insert follows a remove (which
            # slot a concurrent insert takes depends on the order of the
            # batch's inserts), so it must match exactly after the fill.
            st = w.stats()
            assert (st["slabs_used"], st["max_chain"]) == (r.slabs_used(), r.max_chain())
    assert w.status() == 0
    st = w.stats()
    assert st["present"] == r.present()
    dk, dv = w.dump()
    got = dict(zip(host_rows(dk), host_rows(dv) if vb else [b""] * dk.shape[0]))
    assert got == r.dump()
    w.close()


def test_wide_tag_collisions(lslab, oracle_mod):
    """Keys in one bucket whose 32-bit tags collide: a probe must confirm the
    full key, so each still reads its own value."""
    say("wide_tag_collisions")
    rng = np.random.default_rng(99)
    n = 1 << 18
    tails = rng.integers(0, 2**32, size=(n, 3), dtype=np.uint64)
    keys = [struct.pack("<QIII", 5, *map(int, t)) for t in tails]  # 20-B keys, bucket 5
    tags = {}
    pairs = []
    for k in keys:
        t = wide_hash(k) & 0xFFFFFFFF
        if t in tags and tags[t] != k:
            pairs.append((tags[t], k))
        tags[t] = k
    assert pairs, "no tag collision found"
    coll = [k for p in pairs[:8] for k in p]
    filler = keys[:100]
    allk = list(dict.fromkeys(filler + coll))
    w = lslab.Wide(20, 8, 3, 100)
    r = WideReplay(20, 8, 3, 100)
    vals = [struct.pack("<Q", i + 1) for i in range(len(allk))]
    f, _ = w.put(dev_rows(allk, 20), dev_rows(vals, 8))
    r.apply([OP_PUT] * len(allk), allk, vals)
    f, v = w.get(dev_rows(allk, 20))
    assert int(f.sum()) == len(allk)
    assert host_rows(v) == vals
    # remove one key of each colliding pair: its partner stays
    gone = [p[0] for p in pairs[:8]]
    w.remove(dev_rows(gone, 20))
    f, v = w.get(dev_rows([p[1] for p in pairs[:8]], 20))
    assert int(f.sum()) == len(gone)
    f, _ = w.get(dev_rows(gone, 20))
    assert int(f.sum()) == 0
    assert w.status() == 0


def test_wide_pool_exhaustion(lslab, oracle_mod):
    say("wide_pool_exhaustion")
    w = lslab.Wide(8, 0, 0, 1)
    keys = dev_rows([struct.pack("<Q", i) for i in range(65)], 8)
    f, _ = w.put(keys)
    assert int(f.sum()) == 60  # the 5 failed inserts report false
    assert w.status() == 4     # LSLAB_ERR_POOL_EXHAUSTED
    st = w.stats()
    assert st["present"] == 60 and st["slabs_used"] == 1


@pytest.mark.parametrize("vb", [8, 0])
def test_wide_conflicts_serial_order(lslab, oracle_mod, vb):
    """Zipf-like repeats of 128-B keys in mixed batches: per key, results
    follow some serial order (the bucket locks order them).  Keys map to
    their index for the narrow checker."""
    say("wide_conflicts", vb)
    rng = np.random.default_rng(17)
    space = 500
    keyset = [key128(i)[:8] + bytes([i % 251]) * 120 for i in range(1, space + 1)]
    p = 1.0 / np.arange(1, space + 1) ** 1.2
    p /= p.sum()
    w = lslab.Wide(128, vb, 3, 1 << 10)
    ref = Replay(vb > 0, 8, 8 if vb else 0, 3, pool_slabs=1 << 12)
    for bi in range(4):
        n = 8000
        ids = rng.choice(space, size=n, p=p)
        ops = rng.choice([0, 1, 2], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
        vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
        keys = dev_rows([keyset[i] for i in ids], 128)
        dvals = (torch.from_numpy(vals.view(np.uint8).reshape(n, 8).copy()).cuda()
                 if vb else None)
        f, old = w.mixed(torch.from_numpy(ops).cuda(), keys, dvals)
        pf, pv = w.get(keys)
        f = f.cpu().numpy()
        pf = pf.cpu().numpy()
        ov = old.cpu().numpy().view(np.uint64).reshape(n) if vb else np.zeros(n, np.uint64)
        pvv = pv.cpu().numpy().view(np.uint64).reshape(n) if vb else np.zeros(n, np.uint64)
        bad, first = ref.check_batch(ops, ids.astype(np.uint64) + np.uint64(1), vals, f, ov,
                                     pf, pvv)
        assert bad == 0, (bi, first)
    assert w.status() == 0


def test_wide_rejects_misaligned_records(lslab, oracle_mod):
    """Key and value records are read as 32-bit words: a buffer that is not
    4-byte aligned is refused before any device work."""
    say("wide_misaligned")
    w = lslab.Wide(16, 8, 4, 100)
    raw = torch.zeros(10 * 16 + 1, dtype=torch.uint8, device="cuda")
    with pytest.raises(lslab.LslabError):
        w.get(raw[1:].view(10, 16))
    f, _ = w.get(raw[:160].view(10, 16))
    assert int(f.sum()) == 0
    assert w.status() == 0


@pytest.mark.parametrize("vb", [8, 0])
def test_wide_mixed_gets_see_the_table_before_the_call(lslab, oracle_mod, vb):
    """A mixed call runs its gets against the table as it was before the call
    and its mutations after them: per key the serial order [gets][mutations],
    wherever the get sits in the batch."""
    say("wide_mixed_gets_first", vb)
    w = lslab.Wide(24, vb, 4, 100)
    keys = [struct.pack("<QQQ", i, i * 7, i * 13) for i in range(1, 41)]
    v0 = [struct.pack("<Q", 1000 + i) for i in range(40)] if vb else None
    w.put(dev_rows(keys[:20], 24), dev_rows(v0[:20], 8) if vb else None)  # keys 0..19 present
    # per key: a get before and after a put (keys 0..9 present, 20..29 absent)
    # and a get before and after a remove (keys 10..19 present)
    ops, ks, vs = [], [], []
    for j in list(range(10)) + list(range(20, 30)):
        ops += [OP_GET, OP_PUT, OP_GET]
        ks += [keys[j]] * 3
        vs += [struct.pack("<Q", 5000 + j)] * 3
    for j in range(10, 20):
        ops += [OP_GET, OP_REMOVE, OP_GET]
        ks += [keys[j]] * 3
        vs += [bytes(8)] * 3
    f, out = w.mixed(torch.tensor(ops, dtype=torch.uint8).cuda(), dev_rows(ks, 24),
                     dev_rows(vs, 8) if vb else None)
    f = f.cpu().numpy()
    outs = host_rows(out) if vb else None
    for t in range(30):
        j = (list(range(10)) + list(range(20, 30)) + list(range(10, 20)))[t]
        present = j < 20
        for pos in (0, 2):  # both gets of the key see the state before the call
            assert f[3 * t + pos] == (1 if present else 0), (j, pos)
            if vb:
                assert outs[3 * t + pos] == (v0[j] if present else bytes(8)), (j, pos)
    assert w.status() == 0
    f, out = w.get(dev_rows(keys[:30], 24))
    f = f.cpu().numpy()
    assert list(f) == [1] * 10 + [0] * 10 + [1] * 10  # puts landed, removes too
    if vb:
        got = host_rows(out)
        assert got[:10] == [struct.pack("<Q", 5000 + j) for j in range(10)]
        assert got[20:30] == [struct.pack("<Q", 5000 + j) for j in range(20, 30)]
