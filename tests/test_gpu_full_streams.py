"""GPU parity of the batched and served forms on the benchmark's own streams.

- The batched form captured in a HIP graph (torch.cuda.graph): a small stream
  of distinct-key batches, runs of small batches among them, replayed three
  times with eager calls and synchronisations between the replays; every
  replay's results equal the oracle's sequential replay op for op.
- The batched form at full size: the prefilled 2^22-bucket map<u64,u64>
  (2^26 keys), the 4096 uneven batches of the benchmark (log-uniform sizes in
  [32, 65536], 70/25/5 get/put/remove, Zipf 0.99 over 2^26), one
  graph-captured launch; every batch is checked in order by the oracle's
  per-key serial-order checker, then every touched key's final state.
- The served form at full size: the benchmark's request stream (U{1..4096}
  ops, 50/50 get/put, Zipf 1.2 over 2^30) at its 2^18-op ticks with every
  request arriving at once, and a Poisson-arrival run at 70% of capacity;
  each tick the loop formed is checked in order by the serial-order checker.

Each case prints its name before it touches the GPU and its progress while
the oracle checks.
"""
import ctypes
import time

import numpy as np
import pytest
import torch

import bench
from oracle.oracle import OP_PUT, Replay

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _h64(t):
    return t.contiguous().cpu().numpy().view(np.uint64)


def _exec_captured(lib, m, offs, ops, keys, vals, out):
    """Captures one lslab_exec_batches call into a graph and returns it."""
    from lslab_amd import _lib
    nb = offs.numel() - 1
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _lib.check(lib.lslab_exec_batches(
            m.handle, nb, ctypes.c_void_p(offs.data_ptr()), ctypes.c_void_p(ops.data_ptr()),
            ctypes.c_void_p(keys.data_ptr()), ctypes.c_void_p(vals.data_ptr()),
            ctypes.c_void_p(out.data_ptr()),
            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)), "exec_batches")
    return g


@pytest.mark.parametrize("combine", [1, 2])
def test_graph_captured_exec_batches_replays(lslab, oracle_mod, combine):
    say("graph_captured_exec_batches", combine)
    from lslab_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(600 + combine)
    n_log2, pool = 8, 1 << 14
    m = lslab.Map(8, 8, n_log2, pool, lslab.tuning(lanes_per_op=4, engine=1,
                                                   max_subbatch=1 << 16, combine=combine))
    ref = Replay(True, 8, 8, n_log2, pool_slabs=pool)
    # sizes: small batches that share runs, one batch of exactly 2^16 ops, one
    # larger than the sub-batch (runs alone in sub-batches), an empty batch
    sizes = np.concatenate([rng.integers(1, 3000, size=30), [1 << 16, 0, 70000, 5, 900]])
    universe = np.arange(1, 200001, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    keys = np.concatenate([rng.permutation(universe)[:s] for s in sizes])  # distinct per batch
    n = len(keys)
    ops = rng.choice([0, 1, 2], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    offs = torch.from_numpy(offs_h).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    d_keys = torch.from_numpy(keys.view(np.int64)).cuda()
    d_vals = torch.from_numpy(vals.view(np.int64)).cuda()
    out = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
    g = _exec_captured(lib, m, offs, d_ops, d_keys, d_vals, out)
    probe = universe[:5000]
    for rep in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert m.status() == 0, rep
        f = out.view(torch.uint8).view(n, 16)[:, 0].cpu().numpy()
        v = _h64(out[:, 1])
        for b in range(len(sizes)):
            lo, hi = offs_h[b], offs_h[b + 1]
            rf, rv, fail = ref.apply(ops[lo:hi], keys[lo:hi], vals[lo:hi])
            assert fail == 0
            np.testing.assert_array_equal(f[lo:hi], rf, err_msg=f"replay {rep} batch {b}")
            np.testing.assert_array_equal(v[lo:hi], rv, err_msg=f"replay {rep} batch {b}")
        # eager work on the same table between replays: gets and a put batch
        ef, ev = m.get(torch.from_numpy(probe.view(np.int64)).cuda())
        rf, rv, _ = ref.apply(np.zeros(len(probe), np.uint8), probe)
        np.testing.assert_array_equal(ef.cpu().numpy(), rf)
        np.testing.assert_array_equal(_h64(ev), rv)
        pk = rng.permutation(universe)[:2000]
        pv = rng.integers(1, 2**63, size=2000, dtype=np.uint64)
        old = m.put(torch.from_numpy(pk.view(np.int64)).cuda(),
                    torch.from_numpy(pv.view(np.int64)).cuda())
        _, rold, _ = ref.apply(np.full(2000, OP_PUT, np.uint8), pk, pv)
        np.testing.assert_array_equal(_h64(old), rold)
    k, vv = m.dump()
    rk, rv = ref.dump()
    o1, o2 = np.argsort(_h64(k)), np.argsort(rk)
    np.testing.assert_array_equal(_h64(k)[o1], rk[o2])
    np.testing.assert_array_equal(_h64(vv)[o1], rv[o2])
    m.close()


def _prefill_both(m, ref, wl):
    """The benchmark's prefill (keys 0..2^26-1, device-generated values) into
    the table and the same pairs into the oracle."""
    n, chunk = 1 << bench.SPACE_LOG2, 1 << 24
    for lo in range(0, n, chunk):
        k = torch.arange(lo, lo + chunk, dtype=torch.int64, device="cuda")
        v = wl.values(chunk, 8, first=lo, seed=bench.PREFILL_SEED)
        m.put(k, v)
        _, _, fail = ref.apply(np.full(chunk, OP_PUT, np.uint8), _h64(k), _h64(v))
        assert fail == 0
    assert m.status() == 0


@pytest.mark.timeout(900)
def test_batched_full_stream_graph_captured(lslab, oracle_mod):
    """Config 4 at full size, one graph-captured persistent launch."""
    say("batched_full_stream_graph_captured")
    from lslab_amd import _lib, autotune
    from lslab_amd import workload as wl
    lib = _lib.load()
    tun = autotune.load("batched")
    m = lslab.Map(8, 8, bench.N_LOG2, 1 << 16,
                  lslab.tuning(lanes_per_op=tun["lanes_per_op"], engine=1,
                               max_blocks=tun["max_blocks"], max_subbatch=bench.BATCH,
                               combine=tun["combine"]))
    ref = Replay(True, 8, 8, bench.N_LOG2, pool_slabs=1 << 16)
    t0 = time.time()
    _prefill_both(m, ref, wl)
    print(f"  prefill of table and oracle: {time.time() - t0:.1f} s", flush=True)
    offs_h, ops, keys, vals = bench.gen_batched(wl, 4096)
    n = int(offs_h[-1])
    offs = torch.from_numpy(offs_h).cuda()
    out = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
    g = _exec_captured(lib, m, offs, ops, keys, vals, out)
    g.replay()
    torch.cuda.synchronize()
    assert m.status() == 0
    print(f"  {n} ops in {len(offs_h) - 1} batches replayed", flush=True)
    h_ops = ops.cpu().numpy()
    h_keys, h_vals = _h64(keys), _h64(vals)
    f = out.view(torch.uint8).view(n, 16)[:, 0].cpu().numpy()
    v = _h64(out[:, 1])
    del out
    t0 = time.time()
    for b in range(len(offs_h) - 1):
        lo, hi = offs_h[b], offs_h[b + 1]
        bad, first = ref.check_batch(h_ops[lo:hi], h_keys[lo:hi], h_vals[lo:hi], f[lo:hi],
                                     v[lo:hi])
        assert bad == 0, (b, first)
        if b % 512 == 511:
            print(f"  checked {b + 1} batches ({time.time() - t0:.1f} s)", flush=True)
    touched = np.unique(h_keys)
    tf, tv = m.get(torch.from_numpy(touched.view(np.int64)).cuda())
    rf, rv, _ = ref.apply(np.zeros(len(touched), np.uint8), touched)
    np.testing.assert_array_equal(tf.cpu().numpy(), rf)
    np.testing.assert_array_equal(_h64(tv), rv)
    assert m.stats()["present"] == ref.present
    print(f"  {len(touched)} touched keys agree; {ref.present} keys present", flush=True)
    m.close()


def _check_ticks(ref, res, offs_h, h_ops, h_keys, h_vals):
    tick_of = res["tick_of"].cpu().numpy().view(np.uint32)
    assert (tick_of != 0xFFFFFFFF).all(), "no request of this stream may be rejected"
    assert (np.diff(tick_of.astype(np.int64)) >= 0).all()  # admitted in arrival order
    f = res["found"].cpu().numpy()
    v = _h64(res["value"])
    starts = np.flatnonzero(np.diff(np.concatenate([[-1], tick_of.astype(np.int64)])))
    ends = np.concatenate([starts[1:], [len(tick_of)]])
    assert len(starts) == res["ticks"]
    for tk, (r0, r1) in enumerate(zip(starts, ends)):
        lo, hi = offs_h[r0], offs_h[r1]
        bad, first = ref.check_batch(h_ops[lo:hi], h_keys[lo:hi], h_vals[lo:hi], f[lo:hi],
                                     v[lo:hi])
        assert bad == 0, (tk, first)
    return len(starts)


@pytest.mark.timeout(900)
def test_served_full_stream_ticks(lslab, oracle_mod):
    """Config 5 at the benchmark's tick size: every tick checked in order."""
    say("served_full_stream_ticks")
    from lslab_amd import autotune
    from lslab_amd import workload as wl
    tun = autotune.load("served")
    for nreq, load in ((20000, None), (4000, 0.7)):
        offs, ops, keys, vals, reserve, rng = bench.gen_served(wl, torch, nreq)
        offs_h = offs.cpu().numpy()
        arrival = np.zeros(nreq, np.int64)
        if load is not None:
            cap = bench._serve_run(lslab, torch, tun["lanes_per_op"], tun["tick_ops"],
                                   tun["max_blocks"], tun["combine"], offs, ops, keys, vals,
                                   reserve, arrival)
            rate = load * nreq / (cap["elapsed_ticks"] / 1e8)
            arrival = np.floor(np.cumsum(rng.exponential(1.0 / rate, size=nreq)) * 1e8)
            arrival = (arrival - arrival[0]).astype(np.int64)
        res = bench._serve_run(lslab, torch, tun["lanes_per_op"], tun["tick_ops"],
                               tun["max_blocks"], tun["combine"], offs, ops, keys, vals, reserve,
                               arrival)
        assert res["rejected"] == 0 and res["served_ops"] == int(offs_h[-1])
        print(f"  {nreq} requests, {int(offs_h[-1])} ops in {res['ticks']} ticks", flush=True)
        ref = Replay(True, 8, 8, 24, pool_slabs=1 << 22)
        t0 = time.time()
        nt = _check_ticks(ref, res, offs_h, ops.cpu().numpy(), _h64(keys), _h64(vals))
        print(f"  {nt} ticks checked ({time.time() - t0:.1f} s)", flush=True)
        ref.close()
