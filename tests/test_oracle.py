"""Pins the oracle (sequential replay, oracle/lslab_oracle.c) on the CPU.

- The reference's own known-answer tests, restated:
  GetPutTest (test/slabunified_test.cu:42-89), the same KAT with a foreign
  allocator pool (:156-205) and ContainsInsertRemoveTest (test/set_test.cu:65-106).
- A brute-force Python dict model of the map/set semantics on random streams
  with chains far longer than one slab (which the reference's tests never build).
- Closed forms for slab accounting: 31 keys per slab, inserts only into the
  last slab's lowest empty slot (detail/traverse.h:54-55,98-124), one pool
  slab per overflow, a bump pool that never reuses (detail/device_allocator.h:62-73).
- The serial-order checker: accepts every valid serial order and rejects
  planted mistakes.
"""
import numpy as np
import pytest

from oracle.oracle import OP_GET, OP_PUT, OP_REMOVE, Replay


def test_kat_get_put(oracle_mod):
    # map<int,int>(10): keys 1..1024, 100 reps of put(rep) then get.
    m = Replay(True, 4, 4, 10)
    keys = np.arange(1, 1025, dtype=np.uint64)
    for rep in range(100):
        f, o, fail = m.apply(np.full(1024, OP_PUT, np.uint8), keys, np.full(1024, rep, np.uint64))
        assert fail == 0
        if rep:
            assert (o == rep - 1).all() and f.all()
        else:
            assert (o == 0).all() and not f.any()
        f, o, _ = m.apply(np.zeros(1024, np.uint8), keys)
        assert f.all() and (o == rep).all()


def test_kat_get_put_foreign_pool(oracle_mod):
    # Same KAT with a pool of a different capacity (the GroupAllocator test
    # swaps only where slabs come from); one slab per bucket is never exceeded.
    m = Replay(True, 4, 4, 10, pool_slabs=2097152 // 288)
    keys = np.arange(1, 1025, dtype=np.uint64)
    for rep in range(5):
        m.apply(np.full(1024, OP_PUT, np.uint8), keys, np.full(1024, rep, np.uint64))
        f, o, _ = m.apply(np.zeros(1024, np.uint8), keys)
        assert f.all() and (o == rep).all()
    assert m.slabs_used == 0


def test_kat_contains_insert_remove(oracle_mod):
    s = Replay(False, 4, 0, 10)
    keys = np.arange(1, 1025, dtype=np.uint64)
    for _ in range(100):
        f, _, fail = s.apply(np.full(1024, OP_PUT, np.uint8), keys)
        assert fail == 0 and f.all()
        f, _, _ = s.apply(np.zeros(1024, np.uint8), keys)
        assert f.all()
        f, _, _ = s.apply(np.full(1024, OP_REMOVE, np.uint8), keys)
        assert f.all()
        f, _, _ = s.apply(np.zeros(1024, np.uint8), keys)
        assert not f.any()


def _dict_model(is_map, kmask, vmask, ops, keys, vals):
    d = {}
    found = np.zeros(len(keys), np.uint8)
    out = np.zeros(len(keys), np.uint64)
    for i, (op, k) in enumerate(zip(ops, keys)):
        k = int(k) & kmask
        v = int(vals[i]) & vmask if vals is not None else 0
        present = k in d
        if op == OP_GET:
            found[i] = present
            out[i] = d[k] if (present and is_map) else 0
        elif op == OP_PUT:
            found[i] = present or not is_map
            out[i] = d[k] if (present and is_map) else 0
            d[k] = v
        else:
            found[i] = present
            out[i] = d[k] if (present and is_map) else 0
            d.pop(k, None)
    return d, found, out


@pytest.mark.parametrize("is_map,kb,vb", [(True, 4, 4), (True, 4, 8), (True, 8, 4), (True, 8, 8),
                                          (False, 4, 0), (False, 8, 0)])
def test_brute_force_vs_dict(oracle_mod, is_map, kb, vb):
    rng = np.random.default_rng(1234 + kb * 10 + vb)
    n = 20000
    # 4 buckets and ~400 live keys: chains of ~4 slabs, removes and re-inserts.
    keyspace = rng.integers(0, 2**64, size=400, dtype=np.uint64)
    keyspace[:4] = [0, 2**32 - 1, 2**64 - 1, 2**32]  # width-edge keys
    ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=n, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    keys = keyspace[rng.integers(0, len(keyspace), size=n)]
    vals = rng.integers(0, 2**64, size=n, dtype=np.uint64)
    r = Replay(is_map, kb, vb, 2, pool_slabs=100000)
    f, o, fail = r.apply(ops, keys, vals)
    assert fail == 0
    kmask = 2**32 - 1 if kb == 4 else 2**64 - 1
    vmask = 2**32 - 1 if vb == 4 else 2**64 - 1
    d, f2, o2 = _dict_model(is_map, kmask, vmask, ops, keys, vals)
    np.testing.assert_array_equal(f, f2)
    np.testing.assert_array_equal(o, o2)
    dk, dv = r.dump()
    assert dict(zip(dk.tolist(), dv.tolist())) == ({k: v for k, v in d.items()} if is_map
                                                    else {k: 0 for k in d})
    assert r.max_chain > 1


def test_slab_accounting_closed_forms(oracle_mod):
    # m distinct keys into one bucket: ceil((m-31)/31) pool slabs.
    for m in (1, 31, 32, 62, 63, 100, 1000):
        r = Replay(True, 8, 8, 0, pool_slabs=1000)
        keys = np.arange(m, dtype=np.uint64)
        r.apply(np.full(m, OP_PUT, np.uint8), keys, keys)
        assert r.slabs_used == max(0, -(-(m - 31) // 31)), m
        assert r.max_chain == 1 + r.slabs_used
    # An empty slot in a non-last slab is never reused; one in the last is.
    r = Replay(True, 8, 8, 0, pool_slabs=10)
    keys = np.arange(62, dtype=np.uint64)
    r.apply(np.full(62, OP_PUT, np.uint8), keys, keys)
    assert r.slabs_used == 1
    r.apply(np.array([OP_REMOVE], np.uint8), np.array([0], np.uint64))  # head slab slot 0
    r.apply(np.array([OP_PUT], np.uint8), np.array([100], np.uint64), np.array([7], np.uint64))
    assert r.slabs_used == 2  # last slab was full -> a new slab despite the hole
    r.apply(np.array([OP_REMOVE], np.uint8), np.array([100], np.uint64))  # last slab slot 0
    r.apply(np.array([OP_PUT], np.uint8), np.array([101], np.uint64), np.array([8], np.uint64))
    assert r.slabs_used == 2  # reused the last slab's hole


def test_chain_histogram_closed_form(oracle_mod):
    # Insert-only: bucket i holding c_i keys has a chain of max(1, ceil(c_i/31)) slabs.
    counts = [0, 1, 31, 32, 62, 63, 200, 500]
    r = Replay(True, 8, 8, 3, pool_slabs=100)
    keys = np.concatenate([np.arange(c, dtype=np.uint64) * np.uint64(8) + np.uint64(i)
                           for i, c in enumerate(counts)])
    r.apply(np.full(len(keys), OP_PUT, np.uint8), keys, keys)
    want = [0] * 16
    for c in counts:
        want[min(max(1, -(-c // 31)), 16) - 1] += 1
    assert r.chain_hist(16) == want
    short = r.chain_hist(4)  # the last bin collects the longer chains
    assert short[:3] == want[:3] and short[3] == sum(want[3:])


def test_put_on_reused_slot_returns_zero(oracle_mod):
    # A removed slot keeps its stale value; a later insert of another key there
    # must still report 0 as the previous value.
    r = Replay(True, 8, 8, 0, pool_slabs=10)
    r.apply(np.array([OP_PUT], np.uint8), np.array([5], np.uint64), np.array([99], np.uint64))
    r.apply(np.array([OP_REMOVE], np.uint8), np.array([5], np.uint64))
    f, o, _ = r.apply(np.array([OP_PUT], np.uint8), np.array([6], np.uint64),
                      np.array([1], np.uint64))
    assert f[0] == 0 and o[0] == 0


def test_pool_exhaustion_count(oracle_mod):
    # 1 bucket, pool of 2: room for 31 * 3 keys; the rest fail.
    r = Replay(False, 8, 0, 0, pool_slabs=2)
    keys = np.arange(31 * 3 + 5, dtype=np.uint64)
    f, _, fail = r.apply(np.full(len(keys), OP_PUT, np.uint8), keys)
    assert fail == 5 and r.present == 93 and r.slabs_used == 2
    assert f[:93].all() and not f[93:].any()


def _replay_in_order(is_map, order, ops, keys, vals, pre):
    """Results of applying ops in `order` to a fresh replay holding `pre`."""
    r = Replay(is_map, 8, 8, 3, pool_slabs=10000)
    if len(pre[0]):
        r.apply(np.full(len(pre[0]), OP_PUT, np.uint8), pre[0], pre[1])
    f = np.zeros(len(keys), np.uint8)
    o = np.zeros(len(keys), np.uint64)
    for i in order:
        ff, oo, _ = r.apply(ops[i:i + 1], keys[i:i + 1], vals[i:i + 1])
        f[i], o[i] = ff[0], oo[0]
    post_f, post_v, _ = r.apply(np.zeros(len(keys), np.uint8), keys, vals)
    return f, o, post_f, post_v


@pytest.mark.parametrize("is_map", [True, False])
def test_checker_accepts_any_serial_order(oracle_mod, is_map):
    rng = np.random.default_rng(7)
    n = 3000
    keys = rng.integers(0, 12, size=n).astype(np.uint64)  # heavy same-key conflicts
    ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    pre = (np.arange(0, 12, 2, dtype=np.uint64), np.arange(100, 106, dtype=np.uint64))
    for order in (np.arange(n), np.arange(n)[::-1], rng.permutation(n),
                  # the two-phase engine's order: every get first, mutations shuffled
                  np.concatenate([np.flatnonzero(ops == OP_GET),
                                  rng.permutation(np.flatnonzero(ops != OP_GET))])):
        f, o, pf, pv = _replay_in_order(is_map, order, ops, keys, vals, pre)
        chk = Replay(is_map, 8, 8, 3, pool_slabs=10000)
        chk.apply(np.full(len(pre[0]), OP_PUT, np.uint8), pre[0], pre[1])
        bad, first = chk.check_batch(ops, keys, vals, f, o, pf, pv)
        assert bad == 0, first
        # the checker moved the table to the observed end state
        f2, o2, _ = chk.apply(np.zeros(n, np.uint8), keys)
        np.testing.assert_array_equal(f2, pf)
        if is_map:
            np.testing.assert_array_equal(o2, pv)


@pytest.mark.parametrize("is_map", [True, False])
def test_checker_rejects_planted_mistakes(oracle_mod, is_map):
    rng = np.random.default_rng(11)
    n = 400
    keys = rng.integers(0, 5, size=n).astype(np.uint64)
    ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    pre = (np.array([0, 2], np.uint64), np.array([50, 51], np.uint64))
    f, o, pf, pv = _replay_in_order(is_map, np.arange(n), ops, keys, vals, pre)

    def bad_count(f_, o_, pf_=pf, pv_=pv):
        chk = Replay(is_map, 8, 8, 3, pool_slabs=10000)
        chk.apply(np.full(2, OP_PUT, np.uint8), pre[0], pre[1])
        return chk.check_batch(ops, keys, vals, f_, o_, pf_, pv_)[0]

    assert bad_count(f, o) == 0
    gets = np.flatnonzero((ops == OP_GET) & (f == 1))
    rems = np.flatnonzero((ops == OP_REMOVE) & (f == 1))
    if is_map:
        puts = np.flatnonzero((ops == OP_PUT) & (f == 1))
        o1 = o.copy()
        o1[puts[3]] ^= 1  # a previous value nobody ever wrote
        assert bad_count(f, o1) >= 1
        o2 = o.copy()
        o2[gets[5]] = 12345  # a get that saw a value never written
        assert bad_count(f, o2) >= 1
        f3 = f.copy()
        f3[rems[0]] = 0  # a remove that missed a present key (value then must be 0)
        o3 = o.copy()
        o3[rems[0]] = 0
        assert bad_count(f3, o3) >= 1
        pv4 = pv.copy()
        pv4[:] = pv4 ^ 1  # an end state no chain reaches
        assert bad_count(f, o, pf, pv4) >= 1
    else:
        # a set insert never reports false
        f2 = f.copy()
        f2[np.flatnonzero(ops == OP_PUT)[0]] = 0
        assert bad_count(f2, o) >= 1

        def small(ops_, found_, post=None, pre_present=False):
            chk = Replay(False, 8, 8, 3, pool_slabs=10)
            if pre_present:
                chk.apply(np.array([OP_PUT], np.uint8), np.array([9], np.uint64))
            k = np.full(len(ops_), 9, np.uint64)
            pf_ = None if post is None else np.full(len(ops_), post, np.uint8)
            return chk.check_batch(np.array(ops_, np.uint8), k, np.zeros(len(ops_), np.uint64),
                                   np.array(found_, np.uint8), np.zeros(len(ops_), np.uint64),
                                   pf_, None)[0]

        # two removes reporting "was present" need two present phases: one insert gives one
        assert small([OP_PUT, OP_REMOVE, OP_REMOVE], [1, 1, 1]) == 1
        assert small([OP_PUT, OP_REMOVE, OP_REMOVE], [1, 1, 0]) == 0
        # absent key, no insert: contains=true is impossible
        assert small([OP_GET, OP_REMOVE], [1, 0]) == 1
        # present key, no remove reporting true: a remove/contains seeing absent is impossible
        assert small([OP_GET, OP_PUT], [0, 1], pre_present=True) == 1
        assert small([OP_REMOVE, OP_REMOVE], [0, 0], pre_present=True) == 1
        # end state: insert then a true remove cannot end present
        assert small([OP_PUT, OP_REMOVE], [1, 1], post=1) == 1
        assert small([OP_PUT, OP_REMOVE], [1, 1], post=0) == 0


@pytest.mark.parametrize("is_map", [True, False])
def test_mt_baseline_matches_replay(oracle_mod, is_map):
    """The multithreaded host baseline (timed beside the sequential replay in
    bench.py) gives the replay's results op by op on batches whose ops touch
    distinct keys, with batch order kept by its barrier."""
    from oracle.oracle import MtBaseline
    rng = np.random.default_rng(91 + is_map)
    nb, space = 12, 6000
    sizes = rng.integers(1, 3000, size=nb)
    keys = np.concatenate([rng.permutation(space)[:s] for s in sizes]).astype(np.uint64)
    keys = keys * np.uint64(0x9E3779B97F4A7C15)
    n = len(keys)
    ops = rng.choice([0, 1, 2], size=n, p=[0.4, 0.45, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    mt = MtBaseline(4, 10, is_map)
    dt, f, v = mt.run(offs, ops, keys, vals if is_map else None, want_results=True)
    assert dt > 0
    r = Replay(is_map, 8, 8 if is_map else 0, 10, pool_slabs=1 << 12)
    for b in range(nb):
        sl = slice(int(offs[b]), int(offs[b + 1]))
        rf, rv, fail = r.apply(ops[sl], keys[sl], vals[sl] if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f[sl], rf, err_msg=f"batch {b}")
        if is_map:
            np.testing.assert_array_equal(v[sl], rv, err_msg=f"batch {b}")
    assert mt.size == r.present
    mt.close()


def _fmix64(x):
    x = np.asarray(x, np.uint64)
    with np.errstate(over="ignore"):
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xFF51AFD7ED558CCD)
        x = x ^ (x >> np.uint64(33))
        x = x * np.uint64(0xC4CEB9FE1A85EC53)
        x = x ^ (x >> np.uint64(33))
    return x


def test_mix_hash_buckets_closed_form(oracle_mod):
    """A MIX table buckets by murmur3's fmix64 (published constants): its chain
    histogram equals the one computed from fmix64(key) & mask in numpy, and
    results do not depend on the hash."""
    n_log2 = 6
    keys = np.arange(1, 3001, dtype=np.uint64) << np.uint64(n_log2)  # all in bucket 0 by identity
    for mix in (False, True):
        r = Replay(True, 8, 8, n_log2, pool_slabs=1000, mix_hash=mix)
        f, v, fail = r.apply(np.full(len(keys), OP_PUT, np.uint8), keys, keys + np.uint64(5))
        assert fail == 0 and not f.any()
        f, v, _ = r.apply(np.zeros(len(keys), np.uint8), keys)
        assert f.all()
        np.testing.assert_array_equal(v, keys + np.uint64(5))
        b = (_fmix64(keys) if mix else keys) & np.uint64((1 << n_log2) - 1)
        counts = np.bincount(b.astype(np.int64), minlength=1 << n_log2)
        want = [0] * 16
        for c in counts:
            want[min(max(1, -(-int(c) // 31)), 16) - 1] += 1
        assert r.chain_hist(16) == want
    # the identity table piles everything into one long chain; the mixed one does not
    assert want[0] + want[1] + want[2] == 1 << n_log2


def test_reclaim_policy_closed_forms(oracle_mod):
    """One bucket, 31-key slabs, groups of 31 keys.  Each cycle removes the
    group inserted two cycles earlier (never in the last slab) and inserts a
    new group: the reference policy needs a new slab every cycle (holes
    outside the last slab are never reused) until the pool is exhausted; the
    reclaim policy refills holes and frees the emptied pool slab, so one pool
    slab suffices."""
    def run(reclaim, cycles):
        r = Replay(True, 8, 8, 0, pool_slabs=4, reclaim=reclaim)
        put = lambda ks: r.apply(np.full(len(ks), OP_PUT, np.uint8), ks, ks)  # noqa: E731
        rem = lambda ks: r.apply(np.full(len(ks), OP_REMOVE, np.uint8), ks)  # noqa: E731
        groups = [np.arange(31 * g, 31 * g + 31, dtype=np.uint64) for g in range(cycles + 2)]
        assert put(np.concatenate(groups[:2]))[2] == 0 and r.slabs_used == 1
        fails = 0
        for c in range(cycles):
            rem(groups[c])
            fails += put(groups[c + 2])[2]
        return r, fails, groups
    r, fails, _ = run(False, 6)
    assert fails > 0 and r.slabs_used == 4  # one new slab per cycle: pool of 4 exhausted
    r, fails, groups = run(True, 6)
    assert fails == 0 and r.slabs_used == 1 and r.present == 62
    # odd cycles refill the pool slab: after 6 cycles it holds group 7 (the head
    # group 6); removing group 7 frees it and leaves the head alone
    r.apply(np.full(31, OP_REMOVE, np.uint8), groups[7])
    assert r.slabs_used == 0 and r.chain_hist(4) == [1, 0, 0, 0]
    # the next overflow takes the freed slab again
    r.apply(np.full(5, OP_PUT, np.uint8), np.arange(5000, 5005, dtype=np.uint64),
            np.arange(5, dtype=np.uint64))
    assert r.slabs_used == 1 and r.present == 36


@pytest.mark.parametrize("is_map", [True, False])
def test_reclaim_matches_dict_and_frees_everything(oracle_mod, is_map):
    """Reclaim changes placement only: results equal a dict model under random
    churn, the pool never exceeds what the live keys need, and removing every
    key returns every pool slab."""
    rng = np.random.default_rng(8)
    n_log2, pool = 2, 400
    r = Replay(is_map, 8, 8 if is_map else 0, n_log2, pool_slabs=pool, reclaim=True)
    d = {}
    for step in range(40):
        keys = rng.integers(0, 3000, size=700).astype(np.uint64)
        ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=700, p=[0.2, 0.45, 0.35]).astype(np.uint8)
        vals = rng.integers(1, 2**40, size=700, dtype=np.uint64)
        f, v, fail = r.apply(ops, keys, vals if is_map else None)
        assert fail == 0
        for i, (op, k) in enumerate(zip(ops, keys)):
            k = int(k)
            had = k in d
            assert f[i] == (1 if (had or (op == OP_PUT and not is_map)) else 0)
            if is_map:
                assert v[i] == (d[k] if had else 0)
            if op == OP_PUT:
                d[k] = int(vals[i]) if is_map else 0
            elif op == OP_REMOVE:
                d.pop(k, None)
        per_bucket = np.bincount(np.array(list(d.keys()), np.int64) & 3, minlength=4)
        assert r.present == len(d)
        assert r.slabs_used <= sum(max(0, int(c) - 31) for c in per_bucket) + pool  # sanity
    alive = np.array(sorted(d), np.uint64)
    r.apply(np.full(len(alive), OP_REMOVE, np.uint8), alive)
    assert r.present == 0 and r.slabs_used == 0
    assert r.chain_hist(4) == [4, 0, 0, 0]
