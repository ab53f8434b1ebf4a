"""GPU parity: the HIP map/set against the oracle (sequential replay).

Phase-separated batches (no two ops of a batch on one key) must match the
oracle bit for bit, op by op, and so must the final table contents and the
slab accounting.  Batches with same-key conflicts are checked per key against
SOME serial order by the oracle's checker.  Every case names itself on stdout
before it touches the GPU.
"""
import numpy as np
import pytest
import torch

from oracle.oracle import OP_GET, OP_PUT, OP_REMOVE, Replay

pytestmark = pytest.mark.gpu

TYPES = [(True, 4, 4), (True, 4, 8), (True, 8, 4), (True, 8, 8), (False, 4, 0), (False, 8, 0)]
ENGINES = [0, 1]
LANES = [1, 2, 4, 8, 16, 32]


def say(*a):
    print("[case]", *a, flush=True)


def _np_dt(b):
    return np.int32 if b == 4 else np.int64


def to_dev(a, b):
    """uint64 numpy bit patterns -> CUDA int tensor of width b."""
    a = np.asarray(a, dtype=np.uint64)
    if b == 4:
        a = (a & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32)
    else:
        a = a.view(np.int64)
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def to_host(t, b):
    a = t.cpu().numpy()
    return a.view(np.uint32).astype(np.uint64) if b == 4 else a.view(np.uint64)


def make(lslab, is_map, kb, vb, n_log2, pool, engine=0, lanes=16, subbatch=1 << 16, combine=2,
         slot_heads=0):
    tun = lslab.tuning(lanes_per_op=lanes, engine=engine, max_subbatch=subbatch, combine=combine,
                       slot_heads=slot_heads)
    if is_map:
        return lslab.Map(kb, vb, n_log2, pool, tun)
    return lslab.Set(kb, n_log2, pool, tun)


def run_batch(tbl, is_map, kb, vb, ops, keys, vals):
    """Runs a mixed batch; returns (found u8, val u64) on the host."""
    n = len(keys)
    d_ops = torch.from_numpy(np.ascontiguousarray(ops, dtype=np.uint8)).cuda()
    d_keys = to_dev(keys, kb)
    if is_map:
        d_vals = to_dev(vals, vb)
        f, v = tbl.mixed(d_ops, d_keys, d_vals)
        return f.cpu().numpy().astype(np.uint8), to_host(v.contiguous(), vb)
    f = tbl.mixed(d_ops, d_keys)
    return f.cpu().numpy().astype(np.uint8), np.zeros(n, np.uint64)


def dump_sorted(tbl, is_map, kb, vb):
    if is_map:
        k, v = tbl.dump()
        k, v = to_host(k, kb), to_host(v, vb)
    else:
        k = to_host(tbl.dump(), kb)
        v = np.zeros(len(k), np.uint64)
    o = np.argsort(k, kind="stable")
    return k[o], v[o]


def oracle_dump_sorted(r):
    k, v = r.dump()
    o = np.argsort(k, kind="stable")
    return k[o], v[o]


@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("lanes", [8, 16])
def test_kat_get_put(lslab, engine, lanes):
    # test/slabunified_test.cu:42-89: map<int,int>(10), keys 1..1024, 100 reps.
    say("kat_get_put", engine, lanes)
    m = make(lslab, True, 4, 4, 10, 10000, engine, lanes)
    keys = torch.arange(1, 1025, dtype=torch.int32, device="cuda")
    for rep in range(100):
        vals = torch.full((1024,), rep, dtype=torch.int32, device="cuda")
        old = m.put(keys, vals)
        if rep:
            assert bool((old == rep - 1).all()), rep
        f, v = m.get(keys)
        assert bool((f == 1).all()) and bool((v == rep).all()), rep
    assert m.status() == 0


def test_kat_get_put_pairs(lslab):
    # map::put(pair<K,V>*, ...) (map.h:189-192) with the pair layout.
    say("kat_get_put_pairs")
    m = make(lslab, True, 4, 8, 10, 10000)
    n = 1024
    for rep in range(3):
        raw = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
        raw.view(torch.int32)[:, 0] = torch.arange(1, n + 1, dtype=torch.int32, device="cuda")
        raw[:, 1] = rep + (1 << 40)
        old = m.put_pairs(raw.view(torch.uint8).view(n, 16))
        if rep:
            assert bool((old == rep - 1 + (1 << 40)).all())
        f, v = m.get(torch.arange(1, n + 1, dtype=torch.int32, device="cuda"))
        assert bool((f == 1).all()) and bool((v == rep + (1 << 40)).all())


@pytest.mark.parametrize("engine", ENGINES)
def test_kat_contains_insert_remove(lslab, engine):
    # test/set_test.cu:65-106: set<int>(10), keys 1..1024, 100 reps.
    say("kat_contains_insert_remove", engine)
    s = make(lslab, False, 4, 0, 10, 10000, engine)
    keys = torch.arange(1, 1025, dtype=torch.int32, device="cuda")
    for rep in range(100):
        assert bool((s.insert(keys) == 1).all())
        assert bool((s.contains(keys) == 1).all())
        assert bool((s.remove(keys) == 1).all())
        assert bool((s.contains(keys) == 0).all())
    assert s.status() == 0


@pytest.mark.parametrize("is_map,kb,vb", TYPES)
@pytest.mark.parametrize("engine", ENGINES)
@pytest.mark.parametrize("lanes", LANES)
def test_phase_separated_bit_exact(lslab, oracle_mod, is_map, kb, vb, engine, lanes):
    """Distinct keys per batch, 16 buckets (chains of many slabs), puts,
    gets of present and absent keys, removes, re-puts into freed slots."""
    say("phase_separated", is_map, kb, vb, engine, lanes)
    rng = np.random.default_rng(100 + kb + vb + lanes + 7 * engine)
    n_log2 = 4
    tbl = make(lslab, is_map, kb, vb, n_log2, 4096, engine, lanes)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=4096)
    universe = rng.permutation(np.arange(1, 6001, dtype=np.uint64) * np.uint64(2654435761))
    if kb == 4:
        universe = universe & np.uint64(0xFFFFFFFF)
        universe = np.unique(universe)
        rng.shuffle(universe)
    batches = []
    a = universe[:3000]
    batches.append((np.full(3000, OP_PUT, np.uint8), a))
    batches.append((np.zeros(4000, np.uint8), np.concatenate([a[:2000], universe[3000:5000]])))
    batches.append((np.full(1500, OP_REMOVE, np.uint8), rng.permutation(a)[:1500]))
    batches.append((np.zeros(3000, np.uint8), a))
    b = np.concatenate([a[:1000], universe[5000:5500]])  # updates + new keys into freed slots
    batches.append((np.full(len(b), OP_PUT, np.uint8), b))
    mixed_keys = rng.permutation(universe[:5600])[:2500]  # one op per key
    batches.append((rng.choice([0, 1, 2], size=2500).astype(np.uint8), mixed_keys))
    for bi, (ops, keys) in enumerate(batches):
        vals = rng.integers(1, 2**63, size=len(keys), dtype=np.uint64)
        if vb == 4:
            vals &= np.uint64(0xFFFFFFFF)
        f_ref, v_ref, fail = ref.apply(ops, keys, vals)
        assert fail == 0
        f, v = run_batch(tbl, is_map, kb, vb, ops, keys, vals)
        bad = np.flatnonzero((f != f_ref) | (v != v_ref))
        assert len(bad) == 0, (bi, bad[:5], f[bad[:5]], f_ref[bad[:5]], v[bad[:5]], v_ref[bad[:5]])
        if bi < 4:
            # Which slot a key gets depends on the order of concurrent inserts,
            # so after removes (batch 2) the holes left in each chain's last
            # slab -- and the slabs later inserts need -- may differ from the
            # sequential replay.  Until an insert follows a remove, pool use is
            # order-free and must match exactly.
            assert tbl.stats()["slabs_used"] == ref.slabs_used, bi
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)
    st = tbl.stats()
    assert st["present"] == ref.present
    assert st["max_chain"] >= 2


# (engine, lanes per op, combining scope, slot heads): the plain engine, and
# the two-phase engine with wavefront-only and block-wide combining of
# same-key mutations, chained at per-slot heads (0: on with the two-phase
# engine) or at combining-table entries only (2: off).
CONFLICT_MODES = [(0, 16, 2, 0), (1, 16, 1, 0), (1, 16, 2, 0), (1, 1, 1, 0), (1, 1, 2, 0),
                  (1, 4, 2, 0), (1, 16, 1, 2), (1, 1, 1, 2), (1, 4, 2, 2)]


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (False, 8, 0)])
@pytest.mark.parametrize("engine,lanes,combine,heads", CONFLICT_MODES)
def test_conflicting_batches_serial_order(lslab, oracle_mod, is_map, kb, vb, engine, lanes,
                                          combine, heads):
    """Zipf-like hot keys with get/put/remove conflicts inside each batch:
    every key's results must match some serial order of its ops."""
    say("conflicting", is_map, kb, vb, engine, lanes, combine, heads)
    rng = np.random.default_rng(5 + engine)
    n_log2 = 5
    tbl = make(lslab, is_map, kb, vb, n_log2, 1 << 14, engine, lanes, combine=combine,
               slot_heads=heads)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=1 << 14)
    space = 3000
    p = 1.0 / np.arange(1, space + 1) ** 1.1
    p /= p.sum()
    perm = rng.permutation(space).astype(np.uint64) * np.uint64(0x9E3779B1) + np.uint64(1)
    if kb == 4:
        perm &= np.uint64(0xFFFFFFFF)
    for bi in range(6):
        n = 20000
        keys = perm[rng.choice(space, size=n, p=p)]
        ops = rng.choice([0, 1, 2], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
        vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
        if vb == 4:
            vals &= np.uint64(0xFFFFFFFF)
        f, v = run_batch(tbl, is_map, kb, vb, ops, keys, vals)
        pf, pv = run_batch(tbl, is_map, kb, vb, np.zeros(n, np.uint8), keys, vals)
        bad, first = ref.check_batch(ops, keys, vals, f, v, pf, pv)
        assert bad == 0, (bi, first)
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)


def test_pool_exhaustion_status(lslab, oracle_mod):
    # 1 bucket, pool of 2 slabs: 93 keys fit, 5 inserts fail with a status.
    say("pool_exhaustion")
    s = make(lslab, False, 8, 0, 0, 2)
    keys = torch.arange(98, dtype=torch.int64, device="cuda")
    res = s.insert(keys)
    assert s.status() == 4
    assert int(res.sum()) == 93
    st = s.stats()
    assert st["present"] == 93 and st["slabs_used"] == 2
    assert s.status() == 0  # cleared


@pytest.mark.parametrize("lanes", [1, 4, 32])
@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (False, 4, 0)])
def test_two_phase_inserts_one_bucket_pair(lslab, oracle_mod, is_map, kb, vb, lanes):
    """Thousands of distinct keys inserted into 2 buckets in one two-phase
    sub-batch: phase W's inserts race for slots of the same last slab and for
    linking the next one.  Slab count and chain length equal the sequential
    replay's, every key reads back its value; with a pool too small, the table
    stops at its exact capacity with POOL_EXHAUSTED."""
    say("two_phase_inserts_one_bucket_pair", is_map, kb, lanes)
    rng = np.random.default_rng(kb + lanes)
    n = 6000
    keys = np.unique(rng.integers(1, 2**31, size=n + 100, dtype=np.uint64))[:n]
    rng.shuffle(keys)
    vals = rng.integers(1, 2**31, size=n, dtype=np.uint64)
    cap = 31  # keys per slab, every layout
    tbl = make(lslab, is_map, kb, vb, 1, 400, 1, lanes)
    ref = Replay(is_map, kb, vb, 1, pool_slabs=400)
    ops = np.full(n, OP_PUT, np.uint8)
    run_batch(tbl, is_map, kb, vb, ops, keys, vals)
    ref.apply(ops, keys, vals if is_map else None)
    assert tbl.status() == 0
    st = tbl.stats()
    assert (st["present"], st["slabs_used"], st["max_chain"]) == (n, ref.slabs_used,
                                                                  ref.max_chain)
    f, v = run_batch(tbl, is_map, kb, vb, np.zeros(n, np.uint8), keys, vals)
    assert f.all()
    if is_map:
        np.testing.assert_array_equal(v, vals)
    small = make(lslab, is_map, kb, vb, 1, 20, 1, lanes)
    run_batch(small, is_map, kb, vb, ops, keys, vals)
    assert small.status() == 4
    st = small.stats()
    assert (st["present"], st["slabs_used"]) == (cap * 22, 20)
    f, v = run_batch(small, is_map, kb, vb, np.zeros(n, np.uint8), keys, vals)
    assert int(f.sum()) == cap * 22
    if is_map:
        np.testing.assert_array_equal(v[f == 1], vals[f == 1])


@pytest.mark.parametrize("engine", ENGINES)
def test_plain_small_config(lslab, oracle_mod, engine):
    """2^20 distinct uint32 keys, 2^16 buckets, one put batch, then 2^20 gets
    (half present, half absent): bit-exact against the replay."""
    say("plain_small", engine)
    from lslab_amd import workload
    n = 1 << 20
    keys = workload.keys(n, 4, first=0, seed=0)
    vals = workload.values(n, 4, first=0, seed=1)
    absent = workload.keys(n // 2, 4, first=n, seed=0)
    present = keys[torch.randperm(n, device="cuda")[: n // 2]]
    q = torch.cat([present, absent])[torch.randperm(n, device="cuda")]
    m = make(lslab, True, 4, 4, 16, 10000, engine)
    old = m.put(keys, vals)
    f, v = m.get(q)
    assert m.status() == 0
    ref = Replay(True, 4, 4, 16)
    hk, hv, hq = to_host(keys, 4), to_host(vals, 4), to_host(q, 4)
    rf, rold, fail = ref.apply(np.full(n, OP_PUT, np.uint8), hk, hv)
    assert fail == 0
    np.testing.assert_array_equal(to_host(old, 4), rold)
    gf, gv, _ = ref.apply(np.zeros(n, np.uint8), hq)
    np.testing.assert_array_equal(f.cpu().numpy(), gf)
    np.testing.assert_array_equal(to_host(v.contiguous(), 4), gv)
    assert int(gf.sum()) == n // 2
    st = m.stats()
    assert st["present"] == n
    if engine == 0:
        assert st["slabs_used"] == ref.slabs_used


@pytest.mark.parametrize("combine", [1, 2])
@pytest.mark.parametrize("is_map", [True, False])
def test_exec_batches_uneven(lslab, oracle_mod, is_map, combine):
    """Uneven batches in one persistent launch; each batch checked in order."""
    say("exec_batches", is_map, combine)
    rng = np.random.default_rng(77)
    kb, vb = 8, 8
    tbl = make(lslab, is_map, kb, vb, 6, 1 << 14, 1, 16, subbatch=1 << 12, combine=combine)
    ref = Replay(is_map, kb, vb, 6, pool_slabs=1 << 14)
    sizes = np.exp(rng.uniform(np.log(1), np.log(9000), size=40)).astype(np.int64)
    sizes[3] = 0  # an empty batch
    n = int(sizes.sum())
    space = 5000
    p = 1.0 / np.arange(1, space + 1) ** 0.99
    p /= p.sum()
    if is_map:
        ranks = rng.choice(space, size=n, p=p)
    else:
        # A set's state after a batch is not implied by its results, so each
        # set batch touches distinct keys (results are then deterministic).
        ranks = np.concatenate([rng.permutation(20000)[:s] for s in sizes])
    keys = ranks.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.7, 0.25, 0.05]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    if is_map:
        f, v = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb), to_dev(vals, vb))
        f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    else:
        f = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb)).cpu().numpy()
        v = np.zeros(n, np.uint64)
    assert tbl.status() == 0
    # Replay batch by batch; the end state of each key after a batch is not
    # observable mid-stream for sets, so sets use the checker without it.
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    if is_map:
        k, vv = dump_sorted(tbl, is_map, kb, vb)
        rk, rv = oracle_dump_sorted(ref)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("lanes,combine", [(2, 1), (4, 2), (1, 1)])
@pytest.mark.parametrize("is_map", [True, False])
def test_exec_batches_all_alone(lslab, oracle_mod, is_map, lanes, combine):
    """A stream in which no two consecutive batches fit one run (every
    adjacent pair holds more than the 2^12-op run cap): the launcher's
    reset kernel picks the kernel without the run path.  Batches larger
    than a sub-batch, empty batches and Zipf conflicts; each batch checked
    in order."""
    say("exec_batches_all_alone", is_map, lanes, combine)
    rng = np.random.default_rng(4097)
    kb, vb = 8, 8
    sub = 1 << 12
    tbl = make(lslab, is_map, kb, vb, 6, 1 << 14, 1, lanes, subbatch=sub, combine=combine)
    ref = Replay(is_map, kb, vb, 6, pool_slabs=1 << 14)
    sizes = np.array([5000, 0, 9000, sub + 1, 0, 4500, 12000, sub + 1, 6000], np.int64)
    assert all(sizes[i] + sizes[i + 1] > sub for i in range(len(sizes) - 1))
    n = int(sizes.sum())
    space = 8000
    p = 1.0 / np.arange(1, space + 1) ** 0.99
    p /= p.sum()
    if is_map:
        ranks = rng.choice(space, size=n, p=p)
    else:
        ranks = np.concatenate([rng.permutation(30000)[:s] for s in sizes])
    keys = ranks.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    if is_map:
        f, v = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb), to_dev(vals, vb))
        f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    else:
        f = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb)).cpu().numpy()
        v = np.zeros(n, np.uint64)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    if is_map:
        k, vv = dump_sorted(tbl, is_map, kb, vb)
        rk, rv = oracle_dump_sorted(ref)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("lanes,max_blocks", [(8, 0), (8, 128), (1, 0)])
def test_exec_batches_many_sizes(lslab, oracle_mod, lanes, max_blocks):
    """Thousands of uneven batches in one persistent launch: each block's
    list of final-mutation candidates must stay intact while the sub-batch
    size changes from batch to batch (a list overwritten by a stale copy
    leaves a combining entry behind, seen as a wrong result or a CORRUPT
    status)."""
    say("exec_batches_many_sizes", lanes, max_blocks)
    rng = np.random.default_rng(4242 + lanes + max_blocks)
    kb = vb = 8
    tbl = make(lslab, True, kb, vb, 12, 1 << 16, 1, lanes, subbatch=1 << 16, combine=1)
    tbl.set_tuning(lslab.tuning(lanes_per_op=lanes, engine=1, max_blocks=max_blocks,
                                max_subbatch=1 << 16, combine=1))
    ref = Replay(True, kb, vb, 12, pool_slabs=1 << 16)
    nb = 1500
    sizes = np.exp(rng.uniform(np.log(1), np.log(40000), size=nb)).astype(np.int64)
    n = int(sizes.sum())
    space = 1 << 17
    p = 1.0 / np.arange(1, space + 1) ** 0.99
    p /= p.sum()
    keys = rng.choice(space, size=n, p=p).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.7, 0.25, 0.05]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    f, v = tbl.exec_batches(torch.from_numpy(offs).cuda(), torch.from_numpy(ops).cuda(),
                            to_dev(keys, kb), to_dev(vals, vb))
    f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    assert tbl.status() == 0
    for b in range(nb):
        lo, hi = offs[b], offs[b + 1]
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    k, vv = dump_sorted(tbl, True, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (False, 8, 0),
                                         (False, 4, 0)])
def test_chain_histogram_matches_oracle(lslab, oracle_mod, is_map, kb, vb):
    """Insert-only table: the device chain-length histogram equals the
    oracle's (chains are then determined by the per-bucket key counts)."""
    say("chain_histogram", is_map, kb, vb)
    rng = np.random.default_rng(17)
    n_log2 = 8
    tbl = make(lslab, is_map, kb, vb, n_log2, 1 << 12)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=1 << 12)
    # skewed bucket loads: bucket b gets about b/4 keys, some chains span many slabs
    keys = np.concatenate([np.arange(rng.integers(0, 300), dtype=np.uint64) * np.uint64(256)
                           + np.uint64(b) for b in range(256)])
    rng.shuffle(keys)
    ops = np.full(len(keys), OP_PUT, np.uint8)
    run_batch(tbl, is_map, kb, vb, ops, keys, keys)
    ref.apply(ops, keys, keys if is_map else None)
    assert tbl.status() == 0
    assert tbl.chain_histogram(16) == ref.chain_hist(16)
    assert tbl.chain_histogram(3) == ref.chain_hist(3)
    assert tbl.stats()["max_chain"] == ref.max_chain


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (False, 4, 0)])
@pytest.mark.parametrize("engine", ENGINES)
def test_mix_hash_table(lslab, oracle_mod, is_map, kb, vb, engine):
    """A table created with LSLAB_HASH_MIX: keys that share their low bits
    (one bucket under the identity hash) spread over the buckets exactly as
    the oracle's fmix64 places them, and every result is the replay's."""
    say("mix_hash", is_map, kb, vb, engine)
    n_log2 = 7
    tun = lslab.tuning(lanes_per_op=4, engine=engine, max_subbatch=1 << 14, hash=lslab.HASH_MIX)
    tbl = (lslab.Map(kb, vb, n_log2, 1 << 12, tun) if is_map
           else lslab.Set(kb, n_log2, 1 << 12, tun))
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=1 << 12, mix_hash=True)
    keys = np.arange(1, 6001, dtype=np.uint64) << np.uint64(n_log2)
    rng = np.random.default_rng(3)
    rng.shuffle(keys)
    for ops in (np.full(len(keys), OP_PUT, np.uint8),
                np.zeros(len(keys), np.uint8),
                rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=len(keys)).astype(np.uint8)):
        vals = rng.integers(1, 2**31, size=len(keys), dtype=np.uint64)
        f, v = run_batch(tbl, is_map, kb, vb, ops, keys, vals)
        rf, rv, fail = ref.apply(ops, keys, vals if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f, rf)
        np.testing.assert_array_equal(v, rv)
        if (ops == OP_PUT).all():
            assert tbl.chain_histogram(16) == ref.chain_hist(16)
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rvv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rvv)
    # the table's hash cannot change after creation; 0 keeps it
    with pytest.raises(Exception):
        tbl.set_tuning(lslab.tuning(lanes_per_op=4, engine=engine, hash=lslab.HASH_IDENTITY))
    tbl.set_tuning(lslab.tuning(lanes_per_op=2, engine=engine, max_subbatch=1 << 14, hash=0))
    f, _ = run_batch(tbl, is_map, kb, vb, np.zeros(len(keys), np.uint8), keys, keys)
    rf, _, _ = ref.apply(np.zeros(len(keys), np.uint8), keys)
    np.testing.assert_array_equal(f, rf)


def make_reclaim(lslab, is_map, kb, vb, n_log2, pool, engine, lanes, reclaim):
    tun = lslab.tuning(lanes_per_op=lanes, engine=engine, max_subbatch=1 << 14,
                       reclaim=lslab.RECLAIM_ON if reclaim else lslab.RECLAIM_OFF)
    if is_map:
        return lslab.Map(kb, vb, n_log2, pool, tun)
    return lslab.Set(kb, n_log2, pool, tun)


@pytest.mark.parametrize("engine,lanes", [(0, 4), (1, 1), (1, 4), (1, 32)])
def test_reclaim_churn(lslab, oracle_mod, engine, lanes):
    """One bucket, groups of 31 keys; each batch removes the group inserted two
    batches earlier and inserts a new one.  The reference policy exhausts a
    4-slab pool (status POOL_EXHAUSTED); the reclaim policy keeps one pool
    slab, and its results, slab count and chains equal the oracle's."""
    say("reclaim_churn", engine, lanes)
    groups = [np.arange(31 * g, 31 * g + 31, dtype=np.uint64) * np.uint64(3) for g in range(10)]
    for reclaim in (False, True):
        tbl = make_reclaim(lslab, True, 8, 8, 0, 4, engine, lanes, reclaim)
        ref = Replay(True, 8, 8, 0, pool_slabs=4, reclaim=reclaim)
        # One group per batch: a batch's inserts into one bucket land in lock
        # order, so only whole-group batches make the slab of every key, and
        # with it the pool use, independent of that order.
        for g in groups[:2]:
            run_batch(tbl, True, 8, 8, np.full(31, OP_PUT, np.uint8), g, g)
            ref.apply(np.full(31, OP_PUT, np.uint8), g, g)
        failed = 0
        for c in range(8):
            for ops, ks in ((np.full(31, OP_REMOVE, np.uint8), groups[c]),
                            (np.full(31, OP_PUT, np.uint8), groups[c + 2])):
                f, v = run_batch(tbl, True, 8, 8, ops, ks, ks + np.uint64(c))
                rf, rv, fail = ref.apply(ops, ks, ks + np.uint64(c))
                failed += fail
                if not fail:
                    np.testing.assert_array_equal(f, rf)
                    np.testing.assert_array_equal(v, rv)
        st = tbl.status()
        if reclaim:
            assert failed == 0 and st == 0
            assert tbl.stats()["slabs_used"] == ref.slabs_used == 1
            assert tbl.chain_histogram(4) == ref.chain_hist(4) == [0, 1, 0, 0]
            k, vv = dump_sorted(tbl, True, 8, 8)
            rk, rvv = oracle_dump_sorted(ref)
            np.testing.assert_array_equal(k, rk)
            np.testing.assert_array_equal(vv, rvv)
            # remove the group in the pool slab: the slab is freed
            run_batch(tbl, True, 8, 8, np.full(31, OP_REMOVE, np.uint8), groups[9], groups[9])
            assert tbl.stats()["slabs_used"] == 0 and tbl.chain_histogram(4) == [1, 0, 0, 0]
        else:
            assert failed > 0 and st == 4  # LSLAB_ERR_POOL_EXHAUSTED
            assert tbl.stats()["slabs_used"] == 4


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 4), (False, 8, 0),
                                         (False, 4, 0)])
@pytest.mark.parametrize("engine,lanes", [(0, 8), (1, 1), (1, 8)])
def test_reclaim_random_stream(lslab, oracle_mod, is_map, kb, vb, engine, lanes):
    """Phase-separated batches (distinct keys per batch) with many removes on
    a small table under the reclaim policy: results bit-exact against the
    oracle's reclaim replay, equal contents, and removing every key returns
    every pool slab."""
    say("reclaim_random", is_map, kb, vb, engine, lanes)
    rng = np.random.default_rng(31 + kb + lanes + engine)
    n_log2, pool = 3, 2048
    tbl = make_reclaim(lslab, is_map, kb, vb, n_log2, pool, engine, lanes, True)
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=pool, reclaim=True)
    universe = np.arange(1, 20001, dtype=np.uint64) * np.uint64(2654435761)
    if kb == 4:
        universe &= np.uint64(0xFFFFFFFF)
    for bi in range(30):
        keys = rng.permutation(universe)[:rng.integers(200, 4000)]
        ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=len(keys),
                         p=[0.2, 0.4, 0.4]).astype(np.uint8)
        vals = rng.integers(1, 2**31, size=len(keys), dtype=np.uint64)
        f, v = run_batch(tbl, is_map, kb, vb, ops, keys, vals)
        rf, rv, fail = ref.apply(ops, keys, vals if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f, rf, err_msg=f"batch {bi}")
        np.testing.assert_array_equal(v, rv, err_msg=f"batch {bi}")
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rvv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rvv)
    st = tbl.stats()
    assert st["present"] == ref.present
    # chains hold no empty pool slab: at most one partly filled slab more per
    # key range than the keys need (a bound, placement is order-dependent)
    assert st["slabs_used"] <= st["present"]
    run_batch(tbl, is_map, kb, vb, np.full(len(k), OP_REMOVE, np.uint8), k, k)
    assert tbl.status() == 0
    st = tbl.stats()
    assert st["present"] == 0 and st["slabs_used"] == 0
    assert tbl.chain_histogram(4) == [1 << n_log2, 0, 0, 0]


@pytest.mark.parametrize("is_map,kb,vb", [(True, 8, 8), (True, 4, 8), (False, 4, 0)])
@pytest.mark.parametrize("engine", ENGINES)
def test_caller_slab_memory(lslab, oracle_mod, is_map, kb, vb, engine):
    """A table whose slab array lives in a caller-owned buffer (the reference's
    Allocator argument): the library zeroes the garbage it is given, results are
    the replay's, a short or misaligned buffer is refused, the bytes are the
    table's, and destroying the table leaves the buffer to its owner."""
    say("caller_slab_memory", is_map, kb, vb, engine)
    n_log2, pool = 6, 1 << 10
    need = lslab.slab_bytes(is_map, kb, vb, n_log2, pool)
    # a u32 set's slab is its 128-B key line plus a 4-B next link
    slab = {(True, 8, 8): 512, (True, 4, 8): 384, (False, 4, 0): 132}[(is_map, kb, vb)]
    assert need == ((1 << n_log2) + pool) * slab
    arena = torch.full((need + 512,), 0xAB, dtype=torch.uint8, device="cuda")
    tun = lslab.tuning(lanes_per_op=8, engine=engine, max_subbatch=1 << 14)

    def mk(mem):
        return (lslab.Map(kb, vb, n_log2, pool, tun, slab_mem=mem) if is_map
                else lslab.Set(kb, n_log2, pool, tun, slab_mem=mem))

    with pytest.raises(lslab.LslabError):
        mk(arena[:need - 1])
    with pytest.raises(lslab.LslabError):
        mk(arena[1:need + 1])
    tbl = mk(arena[256:256 + need])
    ref = Replay(is_map, kb, vb, n_log2, pool_slabs=pool)
    rng = np.random.default_rng(11)
    for step in range(6):
        keys = rng.integers(1, 20000, size=5000, dtype=np.uint64)
        ops = (np.full(len(keys), OP_PUT, np.uint8) if step == 0 else
               rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=len(keys)).astype(np.uint8))
        keys = np.unique(keys)
        ops = ops[:len(keys)]
        vals = rng.integers(1, 2**31, size=len(keys), dtype=np.uint64)
        f, v = run_batch(tbl, is_map, kb, vb, ops, keys, vals)
        rf, rv, fail = ref.apply(ops, keys, vals if is_map else None)
        assert fail == 0
        np.testing.assert_array_equal(f, rf)
        np.testing.assert_array_equal(v, rv)
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rvv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rvv)
    torch.cuda.synchronize()
    host = arena.cpu()
    # bytes outside the table's range are untouched; inside, the table wrote
    assert (host[:256] == 0xAB).all() and (host[256 + need:] == 0xAB).all()
    assert not (host[256:256 + need] == 0xAB).all()
    tbl.close()
    arena.fill_(1)
    torch.cuda.synchronize()
    assert int(arena.sum()) == need + 512


@pytest.mark.parametrize("is_map", [True, False])
def test_exec_batches_run_boundaries(lslab, oracle_mod, is_map):
    """Batches around the run limits (runs hold at most the sub-batch size,
    2^18 ops here): batches of 2^16 ops that fill runs exactly, a batch that
    cannot join the run before it, a batch larger than the sub-batch (runs
    alone in sub-batches), empty batches inside runs.  Each batch is checked
    in order against the oracle's serial-order checker, then the final
    contents."""
    say("exec_batches_run_boundaries", is_map)
    rng = np.random.default_rng(41)
    kb, vb = 8, 8
    tbl = make(lslab, is_map, kb, vb, 8, 1 << 15, 1, 4, subbatch=1 << 18, combine=1)
    ref = Replay(is_map, kb, vb, 8, pool_slabs=1 << 15)
    sizes = np.array([100, 1 << 16, 50, 0, (1 << 16) + 1, 7, 300000, 5, 1 << 16, 1 << 16,
                      1 << 16, 1 << 16, 3], np.int64)
    n = int(sizes.sum())
    space = 40000
    p = 1.0 / np.arange(1, space + 1) ** 0.99
    p /= p.sum()
    if is_map:
        ranks = rng.choice(space, size=n, p=p)
    else:
        ranks = np.concatenate([rng.permutation(400000)[:s] for s in sizes])
    keys = ranks.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.7, 0.25, 0.05]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    if is_map:
        f, v = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb), to_dev(vals, vb))
        f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    else:
        f = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb)).cpu().numpy()
        v = np.zeros(n, np.uint64)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    if is_map:
        k, vv = dump_sorted(tbl, is_map, kb, vb)
        rk, rv = oracle_dump_sorted(ref)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("runs", [0, 1, 2])
def test_exec_batches_runs_modes(lslab, oracle_mod, runs):
    """lslab_tuning.runs: auto, run kernel only, all-alone kernel only.  The
    same stream (small batches that form runs, batches larger than a
    sub-batch, an empty batch) gives oracle-correct results in every mode."""
    say("exec_batches_runs_modes", runs)
    rng = np.random.default_rng(3003)
    kb = vb = 8
    tbl = lslab.Map(kb, vb, 6, 1 << 14, lslab.tuning(lanes_per_op=2, engine=1,
                                                     max_subbatch=1 << 12, combine=1, runs=runs))
    ref = Replay(True, kb, vb, 6, pool_slabs=1 << 14)
    sizes = np.array([30, 500, 7, 0, 9000, 64, 1, 4096, 200, 5000, 40], np.int64)
    n = int(sizes.sum())
    space = 2000
    p = 1.0 / np.arange(1, space + 1) ** 0.99
    p /= p.sum()
    keys = rng.choice(space, size=n, p=p).astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.6, 0.3, 0.1]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    f, v = tbl.exec_batches(torch.from_numpy(offs).cuda(), torch.from_numpy(ops).cuda(),
                            to_dev(keys, kb), to_dev(vals, vb))
    f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    k, vv = dump_sorted(tbl, True, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("combine", [1, 2])
@pytest.mark.parametrize("is_map", [True, False])
def test_exec_batches_long_runs(lslab, oracle_mod, is_map, combine):
    """Runs of the full 128 batches (small batches, far below the 2^16-op run
    cap) over a small, Zipf-hot key space, so that one key is mutated in
    batches on both sides of batch 64 of a run: the high word of the key
    table's batch bitmap carries the nearest earlier batch and the key's
    last batch.  Each batch checked in order, then the final contents."""
    say("exec_batches_long_runs", is_map, combine)
    rng = np.random.default_rng(128 + combine)
    kb, vb = 8, 8
    tbl = make(lslab, is_map, kb, vb, 6, 1 << 14, 1, 2, subbatch=1 << 16, combine=combine)
    ref = Replay(is_map, kb, vb, 6, pool_slabs=1 << 14)
    sizes = rng.integers(20, 200, size=400).astype(np.int64)
    sizes[[5, 70, 130, 300]] = 0  # empty batches inside runs
    n = int(sizes.sum())
    space = 300
    p = 1.0 / np.arange(1, space + 1) ** 1.1
    p /= p.sum()
    if is_map:
        ranks = rng.choice(space, size=n, p=p)
    else:
        ranks = np.concatenate([rng.permutation(space)[:s] for s in sizes])
    keys = ranks.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    ops = rng.choice([0, 1, 2], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    if is_map:
        f, v = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb), to_dev(vals, vb))
        f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    else:
        f = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb)).cpu().numpy()
        v = np.zeros(n, np.uint64)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    if is_map:
        k, vv = dump_sorted(tbl, is_map, kb, vb)
        rk, rv = oracle_dump_sorted(ref)
        np.testing.assert_array_equal(k, rk)
        np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("lanes", [1, 4])
@pytest.mark.parametrize("is_map", [True, False])
def test_exec_batches_crowded_key_table(lslab, oracle_mod, is_map, lanes):
    """Runs whose key table is crowded: sub-batches of 64 ops give runs of at
    most 64 ops and a key table of 128 entries, and the runs mutate up to 64
    distinct keys of a wide key space (the all-ones key among them, which has
    its own entry), so the table fills to half and probes wrap and run long.
    Each batch checked in order, then the final contents."""
    say("exec_batches_crowded_key_table", is_map, lanes)
    rng = np.random.default_rng(6400 + lanes)
    kb, vb = 8, 8
    tbl = make(lslab, is_map, kb, vb, 6, 1 << 14, 1, lanes, subbatch=64, combine=1)
    ref = Replay(is_map, kb, vb, 6, pool_slabs=1 << 14)
    sizes = rng.integers(1, 12, size=3000).astype(np.int64)
    sizes[[10, 200, 1999]] = 0
    sizes[[50, 51, 900]] = [64, 63, 65]  # a full run, a near-full one, one above the cap
    n = int(sizes.sum())
    space = 6000
    if is_map:
        ranks = rng.integers(0, space, size=n)
        hot = rng.random(n) < 0.3  # a hot share over few keys, so keys repeat across batches
        ranks[hot] = rng.integers(0, 40, size=int(hot.sum()))
    else:
        ranks = np.concatenate([rng.permutation(space)[:s] for s in sizes])
    keys = ranks.astype(np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    keys[ranks == 7] = np.uint64(0xFFFFFFFFFFFFFFFF)
    ops = rng.choice([0, 1, 2], size=n, p=[0.4, 0.4, 0.2]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    if is_map:
        f, v = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb), to_dev(vals, vb))
        f, v = f.cpu().numpy(), to_host(v.contiguous(), vb)
    else:
        f = tbl.exec_batches(d_offs, d_ops, to_dev(keys, kb)).cpu().numpy()
        v = np.zeros(n, np.uint64)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    k, vv = dump_sorted(tbl, is_map, kb, vb)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    if is_map:
        np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("combine", [1, 2])
def test_exec_batches_reclaim_runs(lslab, oracle_mod, combine):
    """The batched kernel on a reclaim table: many small batches (runs sharing
    one phase W) that insert and remove heavily in few buckets, so pool slabs
    are emptied, unlinked and reused inside runs.  Each batch follows a serial
    order, the final contents are the replay's, and the pool slabs in use are
    at least what the final contents need."""
    say("exec_batches_reclaim_runs", combine)
    rng = np.random.default_rng(1234 + combine)
    tun = lslab.tuning(lanes_per_op=4, engine=1, max_subbatch=1 << 14, combine=combine,
                       reclaim=lslab.RECLAIM_ON)
    tbl = lslab.Map(8, 8, 3, 1 << 12, tun)
    ref = Replay(True, 8, 8, 3, pool_slabs=1 << 12, reclaim=True)
    sizes = np.exp(rng.uniform(np.log(1), np.log(3000), size=120)).astype(np.int64)
    n = int(sizes.sum())
    keys = rng.integers(1, 3000, size=n).astype(np.uint64)
    ops = rng.choice([0, 1, 2], size=n, p=[0.3, 0.4, 0.3]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    f, v = tbl.exec_batches(torch.from_numpy(offs).cuda(), torch.from_numpy(ops).cuda(),
                            to_dev(keys, 8), to_dev(vals, 8))
    f, v = f.cpu().numpy(), to_host(v.contiguous(), 8)
    assert tbl.status() == 0
    for b in range(len(sizes)):
        lo, hi = offs[b], offs[b + 1]
        if hi == lo:
            continue
        bad, first = ref.check_batch(ops[lo:hi], keys[lo:hi], vals[lo:hi], f[lo:hi], v[lo:hi])
        assert bad == 0, (b, first)
    k, vv = dump_sorted(tbl, True, 8, 8)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)
    st = tbl.stats()
    assert st["present"] == len(rk)
    per_bucket = np.bincount((rk & np.uint64(7)).astype(np.int64), minlength=8)
    need = int(sum(max(0, -(-(int(c) - 31) // 31)) for c in per_bucket))
    assert need <= st["slabs_used"] <= 1 << 12


def test_set_tuning_resizes_scratch(lslab, oracle_mod):
    """Changing max_subbatch re-sizes the two-phase scratch (device_bytes
    follows it) and the table keeps its contents and results."""
    say("set_tuning_resize")
    tbl = make(lslab, True, 8, 8, 6, 1 << 10, 1, 4, subbatch=1 << 12)
    ref = Replay(True, 8, 8, 6, pool_slabs=1 << 10)
    rng = np.random.default_rng(8)
    before = tbl.stats()["device_bytes"]
    for sub in (1 << 16, 1 << 10, 1 << 12):
        tbl.set_tuning(lslab.tuning(lanes_per_op=4, engine=1, max_subbatch=sub))
        keys = rng.choice(np.arange(1, 20000, dtype=np.uint64), size=3000, replace=False)
        ops = rng.choice([0, 1, 2], size=3000).astype(np.uint8)
        vals = rng.integers(1, 2**63, size=3000, dtype=np.uint64)
        f, v = run_batch(tbl, True, 8, 8, ops, keys, vals)
        rf, rv, fail = ref.apply(ops, keys, vals)
        assert fail == 0
        np.testing.assert_array_equal(f, rf)
        np.testing.assert_array_equal(v, rv)
        if sub == 1 << 16:
            assert tbl.stats()["device_bytes"] > before
    assert tbl.stats()["device_bytes"] == before
    assert tbl.status() == 0


def test_slot_heads_follow_tuning(lslab, oracle_mod):
    """Per-slot chain heads are allocated while the tuning asks for them (AUTO:
    with the two-phase engine) and freed when it stops; device_bytes follows
    (4 B per key slot), and results stay the replay's in every mode, on a
    stream that mutates present and absent keys."""
    say("slot_heads_follow_tuning")
    n_log2, pool = 6, 1 << 10
    tbl = make(lslab, True, 8, 8, n_log2, pool, 0, 4, subbatch=1 << 14)
    ref = Replay(True, 8, 8, n_log2, pool_slabs=pool)
    heads = ((1 << n_log2) + pool) * 32 * 4
    b0 = tbl.stats()["device_bytes"]
    rng = np.random.default_rng(21)
    modes = [(1, 0, heads), (1, lslab.SLOT_HEADS_OFF, 0), (0, lslab.SLOT_HEADS_ON, heads),
             (0, 0, 0), (1, lslab.SLOT_HEADS_ON, heads)]
    for engine, sh, extra in modes:
        tbl.set_tuning(lslab.tuning(lanes_per_op=4, engine=engine, max_subbatch=1 << 14,
                                    combine=lslab.COMBINE_WAVE, slot_heads=sh))
        assert tbl.stats()["device_bytes"] == b0 + extra, (engine, sh)
        for _ in range(3):
            keys = rng.integers(1, 4000, size=6000, dtype=np.uint64)
            ops = rng.choice([OP_GET, OP_PUT, OP_REMOVE], size=len(keys),
                             p=[0.5, 0.35, 0.15]).astype(np.uint8)
            vals = rng.integers(1, 2**63, size=len(keys), dtype=np.uint64)
            f, v = run_batch(tbl, True, 8, 8, ops, keys, vals)
            pf, pv = run_batch(tbl, True, 8, 8, np.zeros(len(keys), np.uint8), keys, vals)
            bad, first = ref.check_batch(ops, keys, vals, f, v, pf, pv)
            assert bad == 0, (engine, sh, first)
    assert tbl.status() == 0
    k, vv = dump_sorted(tbl, True, 8, 8)
    rk, rv = oracle_dump_sorted(ref)
    np.testing.assert_array_equal(k, rk)
    np.testing.assert_array_equal(vv, rv)


@pytest.mark.parametrize("runs", [0, 1, 2])
def test_exec_batches_refuses_decreasing_offsets(lslab, oracle_mod, runs):
    """Batch offsets that decrease are refused on the device before any op is
    read: the call leaves the table as it was and the status reads
    LSLAB_ERR_INVALID; the next valid call runs normally."""
    say("exec_batches_refuses_decreasing_offsets", runs)
    kb = vb = 8
    tbl = lslab.Map(kb, vb, 6, 1 << 10, lslab.tuning(lanes_per_op=2, engine=1,
                                                     max_subbatch=1 << 10, runs=runs))
    keys = (np.arange(1, 201, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15))
    vals = np.arange(1, 201, dtype=np.uint64)
    ops = np.ones(200, np.uint8)
    d_ops = torch.from_numpy(ops).cuda()
    d_keys, d_vals = to_dev(keys, kb), to_dev(vals, vb)
    bad = torch.tensor([0, 100, 50, 200], dtype=torch.int64, device="cuda")
    tbl.exec_batches(bad, d_ops, d_keys, d_vals)
    assert tbl.status() == lslab._lib.ERR_INVALID
    assert tbl.stats()["present"] == 0
    good = torch.tensor([0, 100, 150, 200], dtype=torch.int64, device="cuda")
    f, v = tbl.exec_batches(good, d_ops, d_keys, d_vals)
    assert tbl.status() == 0
    assert tbl.stats()["present"] == 200
    k, vv = dump_sorted(tbl, True, kb, vb)
    o = np.argsort(keys)
    np.testing.assert_array_equal(k, keys[o])
    np.testing.assert_array_equal(vv, vals[o])
