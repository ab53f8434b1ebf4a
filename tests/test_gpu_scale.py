"""Full-size configurations, checked through size-independent properties.

- set at scale: set<u64>, 2^24 buckets, 2^28 distinct keys inserted in 16
  launches of 2^24; 2^26 contains (80% present, 20% absent); remove a seeded
  half; every inserted key re-checked.  Expected answers are known from how
  the keys were generated (bijective mix of the index), so every result is
  checked exactly without a replay.
- mixed Zipf at scale: map<u64,u64>, 2^22 buckets prefilled with 2^26 keys;
  two 2^20-op batches (95/5 get/put, Zipf 0.99) through the persistent
  two-phase launch, each checked per key against the oracle after the same
  prefill.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def test_set_at_scale(lslab):
    say("set at scale 2^28")
    from lslab_amd import workload as wl
    s = lslab.Set(8, 24, 1 << 16, lslab.tuning(lanes_per_op=16))
    total = 1 << 28
    chunk = 1 << 24
    for lo in range(0, total, chunk):
        k = wl.keys(chunk, 8, first=lo, seed=0)
        r = s.insert(k)
        assert int(r.sum()) == chunk
        del k, r
    assert s.status() == 0
    st = s.stats()
    assert st["present"] == total
    # chains beyond one slab exist (Poisson(16) tail past 31) and fit the pool
    assert 0 < st["slabs_used"] < (1 << 16) and st["max_chain"] >= 2
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    q = 1 << 26
    hit_idx = torch.randint(0, total, (q * 4 // 5,), device="cuda", generator=g)
    # present keys are mix(i), i < 2^28; absent ones mix(i), i >= 2^28
    present_q = torch.empty(q * 4 // 5, dtype=torch.int64, device="cuda")
    for lo in range(0, present_q.numel(), chunk):
        idx = hit_idx[lo:lo + chunk]
        # gather mix(idx) by generating the keys of a contiguous range and indexing
        present_q[lo:lo + chunk] = _mix64(idx)
    absent_q = wl.keys(q - present_q.numel(), 8, first=total, seed=0)
    queries = torch.cat([present_q, absent_q])
    res = s.contains(queries)
    assert int(res[: present_q.numel()].sum()) == present_q.numel()
    assert int(res[present_q.numel():].sum()) == 0
    del queries, res, present_q, absent_q, hit_idx
    # remove a seeded half: index i is removed iff bit 0 of mix(i ^ seed) is set
    removed_total = 0
    for lo in range(0, total, chunk):
        i = torch.arange(lo, lo + chunk, dtype=torch.int64, device="cuda")
        k = _mix64(i)
        rm = (_mix64(i ^ 0x1234) & 1) == 1
        r = s.remove(k[rm])
        assert bool((r == 1).all())
        removed_total += int(rm.sum())
        del i, k, rm, r
    assert abs(removed_total - total // 2) < total // 100
    for lo in range(0, total, chunk):
        i = torch.arange(lo, lo + chunk, dtype=torch.int64, device="cuda")
        expect = ((_mix64(i ^ 0x1234) & 1) == 0).to(torch.uint8)
        r = s.contains(_mix64(i))
        assert torch.equal(r, expect), lo
        del i, expect, r
    assert s.stats()["present"] == total - removed_total
    assert s.status() == 0


def _mix64(x):
    """murmur3 fmix64 on int64 tensors (the generator's bijective key mix)."""
    x = x.clone()
    m1 = torch.tensor(0xFF51AFD7ED558CCD - (1 << 64), dtype=torch.int64, device=x.device)
    m2 = torch.tensor(0xC4CEB9FE1A85EC53 - (1 << 64), dtype=torch.int64, device=x.device)

    def lsr(v, k):
        return (v >> k) & ((1 << (64 - k)) - 1)

    x = x ^ lsr(x, 33)
    x = x * m1
    x = x ^ lsr(x, 33)
    x = x * m2
    x = x ^ lsr(x, 33)
    return x


def test_mix64_matches_generator(lslab):
    from lslab_amd import workload as wl
    k = wl.keys(1 << 16, 8, first=12345, seed=0)
    i = torch.arange(12345, 12345 + (1 << 16), dtype=torch.int64, device="cuda")
    assert torch.equal(_mix64(i), k)


def test_mixed_zipf_at_scale(lslab, oracle_mod):
    say("mixed zipf at scale 2^26")
    import bench
    from lslab_amd import workload as wl
    from oracle.oracle import Replay
    m = lslab.Map(8, 8, bench.N_LOG2, 1 << 16,
                  lslab.tuning(engine=lslab.ENGINE_TWO_PHASE, max_subbatch=bench.BATCH))
    bench.prefill(m, wl)
    st = m.stats()
    assert st["present"] == 1 << bench.SPACE_LOG2 and st["slabs_used"] == 0
    nb = 2
    ops, keys, vals, offs = bench.make_batches(wl, nb, 77)
    f, v = m.exec_batches(offs, ops, keys, vals)
    assert m.status() == 0
    ref = Replay(True, 8, 8, bench.N_LOG2, pool_slabs=1 << 16)
    n = 1 << bench.SPACE_LOG2
    i = np.arange(n, dtype=np.uint64)
    pv = _mix64(torch.arange(n, dtype=torch.int64, device="cuda") ^ bench.PREFILL_SEED)
    ref.apply(np.ones(n, np.uint8), i, pv.cpu().numpy().view(np.uint64))
    h_ops = ops.cpu().numpy()
    h_keys = keys.cpu().numpy().view(np.uint64)
    h_vals = vals.cpu().numpy().view(np.uint64)
    h_f = f.cpu().numpy()
    h_v = v.contiguous().cpu().numpy().view(np.uint64)
    B = bench.BATCH
    for b in range(nb):
        sl = slice(b * B, (b + 1) * B)
        bad, first = ref.check_batch(h_ops[sl], h_keys[sl], h_vals[sl], h_f[sl], h_v[sl])
        assert bad == 0, (b, first)
    # every get of a key nobody put in its batch returns the table's value
    gets = h_ops == 0
    assert h_f[gets].all()
    # final contents of the keys touched equal the checker's end states
    tk = np.unique(h_keys)
    gf, gv = m.get(torch.from_numpy(tk.view(np.int64)).cuda())
    rf, rv, _ = ref.apply(np.zeros(len(tk), np.uint8), tk)
    np.testing.assert_array_equal(gf.cpu().numpy(), rf)
    np.testing.assert_array_equal(gv.contiguous().cpu().numpy().view(np.uint64), rv)
