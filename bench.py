"""Benchmark: LSlab map on one MI355X, mixed get/put Zipf 0.99 over 2^26 keys.

Headline workload (a step = one batch of the hot path): map<u64,u64> with
2^22 buckets prefilled with keys 0..2^26-1 (16 keys per bucket), batches of
2^20 ops, 95% get / 5% put, keys = a fixed bijection of Zipf(0.99) ranks over
2^26, put values unique per op.  Inputs are generated on the device and are
resident in HBM before timing starts.

- plain: one per-batch launch of the wave64 plain kernel per batch (bucket
  reader/writer locks held per op), batches back to back on one stream;
- fused: the same batches in ONE persistent launch of the two-phase engine
  (lock-free reads, per-key exchange chains, grid barriers between phases and
  batches).  value = fused Mops/s.

The JSON line also carries the plain rate, the roofline of the fused kernel
(algorithmic bytes / HIP-event kernel time), the same batches arriving in
host memory (the table behind PCIe copies vs the KVCG hot/cold split), the batched
form (4096 uneven batches, one graph-captured launch vs one launch per batch),
the served form (device-resident request loop: capacity, then p50/p99 at 70%
of it), the reference's own harness workload (1024-key get and put calls,
launch-inclusive latency) and the host-core baseline (the oracle's sequential
replay of a bounded sample).  The plain rates use the plain engine's own tuned
lanes per op.
Launch shapes come from lslab_amd/tuning.json (python -m lslab_amd.autotune).
"""
import argparse
import ctypes
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mops/s mixed get/put Zipf0.99 2^26 keys, fused vs plain; p99 request latency"
N_LOG2 = 22
SPACE_LOG2 = 26
BATCH = 1 << 20
THETA = 0.99
P_GET, P_PUT = 0.95, 0.05
SEED = 0x5EED
PREFILL_SEED = 0xF111
# Algorithmic bytes per op (u64 keys and values, reference slab layout):
# request 8 B key (+8 B value for a put), result pair<bool,u64> 16 B, one slab
# probe of valid + 31 keys + next = 264 B, the value touched: 8 B read for a
# get, read + write for a put.  get 8+16+264+8 = 296 B; put 16+16+264+8 = 304 B.
B_GET, B_PUT = 296.0, 304.0
B_OP = P_GET * B_GET + P_PUT * B_PUT  # 296.4
HBM_PEAK_GBS = 8000.0


PROFILE = os.path.join(ROOT, "profiles", "k_batches_map_u64_zipf099_headline.md")


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


def measured_traffic_per_op():
    """HBM bytes per op of the fused kernel from the committed rocprofv3 PMC
    summary (FETCH_SIZE x calibration + WRITE_SIZE, per op), or None."""
    try:
        for line in open(PROFILE):
            if line.startswith("traffic_bytes_per_op:"):
                return float(line.split(":", 1)[1])
    except (OSError, ValueError):
        pass
    return None


def make_batches(wl, nb, first_batch):
    import torch
    ops, keys, vals = wl.ops(nb * BATCH, THETA, SPACE_LOG2, P_GET, P_PUT, 8, 8, seed=SEED,
                             first_op=first_batch * BATCH)
    offs = torch.arange(0, (nb + 1) * BATCH, BATCH, dtype=torch.int64, device="cuda")
    return ops, keys, vals, offs


def prefill(m, wl):
    import torch
    n = 1 << SPACE_LOG2
    chunk = 1 << 24
    for lo in range(0, n, chunk):
        k = torch.arange(lo, lo + chunk, dtype=torch.int64, device="cuda")
        v = wl.values(chunk, 8, first=lo, seed=PREFILL_SEED)
        m.put(k, v)
    st = m.status()
    if st:
        raise RuntimeError(f"prefill status {st}")


def gen_batched(wl, nbatches=4096):
    """Config 4 stream: batch sizes log-uniform in [32, 65536], 70/25/5
    get/put/remove, Zipf 0.99 over 2^26.  Returns (host offsets, ops, keys, vals)."""
    import numpy as np
    rng = np.random.default_rng(4096)
    sizes = np.rint(np.exp(rng.uniform(np.log(32), np.log(65536), size=nbatches))).astype(np.int64)
    offs_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ops, keys, vals = wl.ops(int(offs_h[-1]), THETA, SPACE_LOG2, 0.70, 0.25, 8, 8, seed=SEED + 4,
                             first_op=1 << 40)
    return offs_h, ops, keys, vals


def gen_served(wl, torch, nreq):
    """Config 5 stream: requests of U{1..4096} ops, 50/50 get/put, Zipf 1.2
    over 2^30.  Returns (device offsets, ops, keys, vals, reserve, host rng)."""
    import numpy as np
    rng = np.random.default_rng(1200)
    sizes = rng.integers(1, 4097, size=nreq)
    offs_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    ops, keys, vals = wl.ops(int(offs_h[-1]), 1.2, 30, 0.5, 0.5, 8, 8, seed=SEED + 5,
                             first_op=1 << 41)
    offs = torch.from_numpy(offs_h).cuda()
    cs = torch.cumsum((ops == 1).to(torch.int64), 0)
    cs = torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), cs])
    reserve = (cs[offs[1:]] - cs[offs[:-1]]).to(torch.int32)  # one slab per put
    return offs, ops, keys, vals, reserve, rng


def _serve_run(lslab_amd, torch, lanes, tick, max_blocks, combine, offs, ops, keys, vals, reserve,
               arrival_h):
    import numpy as np
    mm = lslab_amd.Map(8, 8, 24, 1 << 22,
                       lslab_amd.tuning(lanes_per_op=lanes, engine=lslab_amd.ENGINE_TWO_PHASE,
                                        max_blocks=max_blocks, max_subbatch=tick, combine=combine))
    arr = torch.from_numpy(arrival_h.astype(np.int64)).cuda()
    torch.cuda.synchronize()
    res = mm.serve(offs, arr, ops, keys, vals, req_reserve=reserve, max_tick_ops=tick)
    st = mm.status()
    mm.close()
    if st:
        raise RuntimeError(f"serve status {st}")
    return res


def served_capacity(lslab_amd, wl, torch, lanes, nreq, tick, max_blocks, combine):
    """Mops/s of the serving loop when every request arrives at t=0."""
    import numpy as np
    offs, ops, keys, vals, reserve, _ = gen_served(wl, torch, nreq)
    res = _serve_run(lslab_amd, torch, lanes, tick, max_blocks, combine, offs, ops, keys, vals,
                     reserve, np.zeros(nreq, np.int64))
    return round(res["served_ops"] / (res["elapsed_ticks"] / 1e8) / 1e6, 1)


def served(lslab_amd, wl, torch, tun, nreq, load_frac=0.7):
    """Config 5: capacity (all arrivals at t=0), then open-loop Poisson
    arrivals at load_frac of that rate, p50/p99 of completion - arrival (both
    on the device's 100 MHz clock)."""
    import numpy as np
    offs, ops, keys, vals, reserve, rng = gen_served(wl, torch, nreq)
    n = int(offs[-1])
    lanes, tick, mb, cb = tun["lanes_per_op"], tun["tick_ops"], tun["max_blocks"], tun["combine"]
    res0 = _serve_run(lslab_amd, torch, lanes, tick, mb, cb, offs, ops, keys, vals, reserve,
                      np.zeros(nreq, np.int64))
    cap_s = res0["elapsed_ticks"] / 1e8
    rate = load_frac * nreq / cap_s  # requests per second
    arrival = np.floor(np.cumsum(rng.exponential(1.0 / rate, size=nreq)) * 1e8).astype(np.int64)
    arrival -= arrival[0]
    res1 = _serve_run(lslab_amd, torch, lanes, tick, mb, cb, offs, ops, keys, vals, reserve, arrival)
    lat_us = (res1["complete_ticks"].cpu().numpy() - arrival) / 100.0
    return {"requests": nreq, "ops": n,
            "capacity_mops": round(res0["served_ops"] / cap_s / 1e6, 2),
            "offered_load": load_frac, "offered_req_per_s": round(rate, 1),
            "mops": round(res1["served_ops"] / (res1["elapsed_ticks"] / 1e8) / 1e6, 2),
            "p50_us": round(float(np.percentile(lat_us, 50)), 2),
            "p99_us": round(float(np.percentile(lat_us, 99)), 2),
            "ticks": int(res1["ticks"]), "rejected": int(res1["rejected"]),
            "lanes_per_op": lanes, "max_tick_ops": tick, "max_blocks": mb, "combine": cb}


def time_plain(m, ops, keys, vals, nb, torch):
    """Per-batch launches of the plain kernel; returns (seconds, per-batch ms list)."""
    from lslab_amd import _lib
    lib = _lib.load()
    out = torch.empty((BATCH, 2), dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    evs = [torch.cuda.Event(enable_timing=True) for _ in range(nb + 1)]
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    evs[0].record(s)
    for b in range(nb):
        lo = b * BATCH
        rc = lib.lslab_map_mixed(m.handle, ctypes.c_void_p(ops[lo:].data_ptr()),
                                 ctypes.c_void_p(keys[lo:].data_ptr()),
                                 ctypes.c_void_p(vals[lo:].data_ptr()),
                                 ctypes.c_void_p(out.data_ptr()), BATCH, sp)
        _lib.check(rc, "map_mixed")
        evs[b + 1].record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per = [evs[i].elapsed_time(evs[i + 1]) for i in range(nb)]
    return wall, per


def time_fused(m, ops, keys, vals, offs, nb, torch):
    """One persistent launch over nb batches; returns (wall s, kernel s from
    HIP events recorded on the launch stream)."""
    from lslab_amd import _lib
    lib = _lib.load()
    out = torch.empty((nb * BATCH, 2), dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(s)
    rc = lib.lslab_exec_batches(m.handle, nb, ctypes.c_void_p(offs.data_ptr()),
                                ctypes.c_void_p(ops.data_ptr()), ctypes.c_void_p(keys.data_ptr()),
                                ctypes.c_void_p(vals.data_ptr()), ctypes.c_void_p(out.data_ptr()), sp)
    _lib.check(rc, "exec_batches")
    e1.record(s)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, e0.elapsed_time(e1) / 1e3


def batched(lslab_amd, wl, torch, m, tun, plain_lanes, nbatches=4096):
    """Config 4 on the prefilled table: one graph-captured persistent launch
    vs one plain launch per batch (at the plain engine's own tuned lanes per
    op)."""
    from lslab_amd import _lib
    lib = _lib.load()
    offs_h, ops, keys, vals = gen_batched(wl, nbatches)
    n = int(offs_h[-1])
    sizes = offs_h[1:] - offs_h[:-1]
    offs = torch.from_numpy(offs_h).cuda()
    out = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    m.set_tuning(lslab_amd.tuning(lanes_per_op=plain_lanes, engine=lslab_amd.ENGINE_PLAIN,
                                  max_subbatch=BATCH))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for b in range(nbatches):
        lo = int(offs_h[b])
        _lib.check(lib.lslab_map_mixed(m.handle, ctypes.c_void_p(ops[lo:].data_ptr()),
                                       ctypes.c_void_p(keys[lo:].data_ptr()),
                                       ctypes.c_void_p(vals[lo:].data_ptr()),
                                       ctypes.c_void_p(out[lo:].data_ptr()), int(sizes[b]), sp),
                   "map_mixed")
    torch.cuda.synchronize()
    plain_s = time.perf_counter() - t0
    m.set_tuning(lslab_amd.tuning(lanes_per_op=tun["lanes_per_op"],
                                  engine=lslab_amd.ENGINE_TWO_PHASE,
                                  max_blocks=tun["max_blocks"], max_subbatch=BATCH,
                                  combine=tun["combine"]))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _lib.check(lib.lslab_exec_batches(m.handle, nbatches, ctypes.c_void_p(offs.data_ptr()),
                                          ctypes.c_void_p(ops.data_ptr()),
                                          ctypes.c_void_p(keys.data_ptr()),
                                          ctypes.c_void_p(vals.data_ptr()),
                                          ctypes.c_void_p(out.data_ptr()),
                                          ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "exec_batches")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    g.replay()
    torch.cuda.synchronize()
    fused_s = time.perf_counter() - t0
    st = m.status()
    if st:
        raise RuntimeError(f"batched status {st}")
    return {"batches": nbatches, "ops": n, "plain_mops": round(n / plain_s / 1e6, 2),
            "fused_graph_mops": round(n / fused_s / 1e6, 2),
            "fused_vs_plain": round(plain_s / fused_s, 3),
            "lanes_per_op": tun["lanes_per_op"], "max_blocks": tun["max_blocks"],
            "combine": tun["combine"], "plain_lanes_per_op": plain_lanes}


def host_input(lslab_amd, torch, m, tun, ops, keys, vals, nb, hot_keys=1 << 16):
    """The headline batches arriving in HOST memory (the KVCG setting): (1) the
    table alone, each batch copied in from pinned memory, run and copied back
    on one stream (two- and three-stream pipelines measured no faster: the
    copies did not overlap), (2) the KVCG hot/cold split with the hot_keys most
    frequent keys of the first batch on host workers, (3) the split with an
    empty hot set (its staging alone)."""
    import numpy as np
    from lslab_amd import _lib
    lib = _lib.load()
    n = nb * BATCH
    m.set_tuning(lslab_amd.tuning(lanes_per_op=tun["lanes_per_op"],
                                  engine=lslab_amd.ENGINE_TWO_PHASE, max_blocks=tun["max_blocks"],
                                  max_subbatch=BATCH, combine=tun["combine"]))
    p_ops = torch.empty(n, dtype=torch.uint8, pin_memory=True)
    p_keys = torch.empty(n, dtype=torch.int64, pin_memory=True)
    p_vals = torch.empty(n, dtype=torch.int64, pin_memory=True)
    p_out = torch.empty((n, 2), dtype=torch.int64, pin_memory=True)
    p_ops.copy_(ops[:n])
    p_keys.copy_(keys[:n])
    p_vals.copy_(vals[:n])
    d_ops = torch.empty(BATCH, dtype=torch.uint8, device="cuda")
    d_keys = torch.empty(BATCH, dtype=torch.int64, device="cuda")
    d_vals = torch.empty(BATCH, dtype=torch.int64, device="cuda")
    d_out = torch.empty((BATCH, 2), dtype=torch.int64, device="cuda")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()

    def serial(nbat):
        for b in range(nbat):
            sl = slice(b * BATCH, (b + 1) * BATCH)
            d_ops.copy_(p_ops[sl], non_blocking=True)
            d_keys.copy_(p_keys[sl], non_blocking=True)
            d_vals.copy_(p_vals[sl], non_blocking=True)
            _lib.check(lib.lslab_map_mixed(m.handle, ctypes.c_void_p(d_ops.data_ptr()),
                                           ctypes.c_void_p(d_keys.data_ptr()),
                                           ctypes.c_void_p(d_vals.data_ptr()),
                                           ctypes.c_void_p(d_out.data_ptr()), BATCH, sp),
                       "map_mixed")
            p_out[sl].copy_(d_out, non_blocking=True)
        torch.cuda.synchronize()

    serial(2)
    t0 = time.perf_counter()
    serial(nb)
    pipe_s = time.perf_counter() - t0
    # raw PCIe rate of pinned copies, both directions
    big = torch.empty(1 << 26, dtype=torch.int64, pin_memory=True)
    dbig = torch.empty(1 << 26, dtype=torch.int64, device="cuda")
    dbig.copy_(big)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    dbig.copy_(big, non_blocking=True)
    torch.cuda.synchronize()
    h2d = big.numel() * 8 / (time.perf_counter() - t0) / 1e9
    t0 = time.perf_counter()
    big.copy_(dbig, non_blocking=True)
    torch.cuda.synchronize()
    d2h = big.numel() * 8 / (time.perf_counter() - t0) / 1e9
    del big, dbig
    h_ops = np.ascontiguousarray(p_ops.numpy())
    h_keys = np.ascontiguousarray(p_keys.numpy().view(np.uint64))
    h_vals = np.ascontiguousarray(p_vals.numpy().view(np.uint64))
    h_out = np.empty((BATCH, 2), dtype=np.uint64)
    threads = min(16, os.cpu_count() or 1)
    res = {"batches": nb, "ops": n, "pcie_h2d_gbs": round(h2d, 1), "pcie_d2h_gbs": round(d2h, 1),
           "table_via_pcie_mops": round(n / pipe_s / 1e6, 2)}
    for label, k in (("kvcg", hot_keys), ("kvcg_cold_only", 0)):
        kv = lslab_amd.Kvcg(m, hot_capacity=max(k, 1), max_batch=BATCH, threads=threads)
        if k:
            kv.sample_hot(h_keys[:BATCH], k)
        kv.mixed(h_ops[:BATCH], h_keys[:BATCH], h_vals[:BATCH], out=h_out, raw=True)  # warm
        st0 = kv.stats()
        t0 = time.perf_counter()
        for b in range(nb):
            sl = slice(b * BATCH, (b + 1) * BATCH)
            kv.mixed(h_ops[sl], h_keys[sl], h_vals[sl], out=h_out, raw=True)
        dt = time.perf_counter() - t0
        st = kv.stats()
        hot = st["hot_ops"] - st0["hot_ops"]
        res[label + "_mops"] = round(n / dt / 1e6, 2)
        if k:
            res["hot_keys"] = k
            res["hot_fraction"] = round(hot / n, 4)
        kv.close()
    res["threads"] = threads
    st = m.status()
    if st:
        raise RuntimeError(f"host-input status {st}")
    return res


def wide_keys(lslab_amd, torch, nb=8):
    """Wide-key map, 128-B keys -> 8-B values (the legacy 128-B-key tests'
    shape): 2^18 buckets prefilled with 2^22 keys (16 per bucket), then
    batches of 2^20 ops over the keys, uniform: get-only calls (lock-free) and
    95/5 get/put mixed calls (bucket locks).  Algorithmic bytes per get: the
    128-B key in, its bucket's 128-B tag line, one 128-B key-arena entry, the
    8-B value and a 9-B result = 401 B."""
    nkeys, n_log2 = 1 << 22, 18
    w = lslab_amd.Wide(128, 8, n_log2, 1 << 16)
    mult = (torch.arange(16, dtype=torch.int64, device="cuda") * 2 + 1) * 0x9E3779B1

    def rows(ids):
        k = ids[:, None] * mult[None, :]
        k[:, 0] = ids
        return k.view(torch.uint8).reshape(ids.numel(), 128)

    for lo in range(0, nkeys, BATCH):
        ids = torch.arange(lo, lo + BATCH, dtype=torch.int64, device="cuda")
        w.put(rows(ids), ids.view(torch.uint8).reshape(BATCH, 8))
    g = torch.Generator(device="cuda")
    g.manual_seed(0x3D1)
    batches = []
    for b in range(nb):
        ids = torch.randint(0, nkeys, (BATCH,), device="cuda", generator=g)
        ops = (torch.rand(BATCH, device="cuda", generator=g) < P_PUT).to(torch.uint8)
        batches.append((rows(ids), ops, ids.view(torch.uint8).reshape(BATCH, 8)))
    torch.cuda.synchronize()
    w.get(batches[0][0])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for k, _, _ in batches:
        w.get(k)
    torch.cuda.synchronize()
    get_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    for k, o, v in batches:
        w.mixed(o, k, v)
    torch.cuda.synchronize()
    mixed_s = time.perf_counter() - t0
    st = w.status()
    if st:
        raise RuntimeError(f"wide status {st}")
    n = nb * BATCH
    res = {"workload": "wide map 128-B keys -> 8-B values, 2^18 buckets, 2^22 keys, "
                       "batches of 2^20 uniform ops",
           "get_mops": round(n / get_s / 1e6, 2), "mixed_95_5_mops": round(n / mixed_s / 1e6, 2),
           "bytes_per_get": 401, "get_gbs": round(401 * n / get_s / 1e9, 1)}
    w.close()
    return res


def _timed(torch, fn):
    """Runs fn() between two events on the current stream; returns seconds."""
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / 1e3


def other_configs(lslab_amd, wl, torch):
    """The two configurations that are parity cases rather than the headline,
    timed with events: map<u32,u32> on 2^16 buckets (put 2^20 distinct keys in
    one call, then 2^20 gets, half of them absent), and set<u64> on 2^24 buckets
    (2^28 inserts in 16 calls, 2^26 contains with 80% present, then a remove of
    half the keys in 16 calls).  Throughputs in millions of ops per second, each
    call's answers checked against what the key generation implies.  Lanes per
    op as measured best for each op on these shapes: 8 for plain-engine
    mutations, 4 for two-phase puts, 2 for lookups."""
    res = {}
    n = 1 << 20
    keys = wl.keys(n, 4, first=0, seed=0)
    vals = wl.values(n, 4, first=0, seed=1)
    absent = wl.keys(n // 2, 4, first=n, seed=0)
    queries = torch.cat([keys[: n // 2], absent])
    for engine, name, lanes in ((lslab_amd.ENGINE_PLAIN, "plain", 8),
                                (lslab_amd.ENGINE_TWO_PHASE, "two_phase", 4)):
        for _rep in range(2):  # the first table warms the launch path up; the second is timed
            m = lslab_amd.Map(4, 4, 16, 1 << 14,
                              lslab_amd.tuning(lanes_per_op=lanes, engine=engine, max_subbatch=n))
            put_s = _timed(torch, lambda: m.put(keys, vals))
            m.set_tuning(lslab_amd.tuning(lanes_per_op=2, engine=engine, max_subbatch=n))
            got = {}
            get_s = _timed(torch, lambda: got.setdefault("r", m.get(queries)))
            found = got["r"][0]
            if m.status() or int(found[: n // 2].sum()) != n // 2 or int(found[n // 2:].sum()):
                raise RuntimeError("plain small config: wrong results")
            m.close()
        res[f"u32_map_put_{name}_mops"] = round(n / put_s / 1e6, 1)
        res[f"u32_map_get_{name}_mops"] = round(n / get_s / 1e6, 1)
    del keys, vals, absent, queries
    look = lslab_amd.tuning(lanes_per_op=2)
    for engine, name, lanes in ((lslab_amd.ENGINE_PLAIN, "plain", 8),
                                (lslab_amd.ENGINE_TWO_PHASE, "two_phase", 4)):
        mut = lslab_amd.tuning(lanes_per_op=lanes, engine=engine, max_subbatch=1 << 24)
        s = lslab_amd.Set(8, 24, 1 << 16, mut)
        total, chunk = 1 << 28, 1 << 24
        ins_s = 0.0
        for lo in range(0, total, chunk):
            k = wl.keys(chunk, 8, first=lo, seed=0)
            ins_s += _timed(torch, lambda: s.insert(k))
            del k
        q = 1 << 26
        qk = torch.cat([wl.keys(q * 4 // 5, 8, first=0, seed=0),
                        wl.keys(q - q * 4 // 5, 8, first=total, seed=0)])
        s.set_tuning(look)
        got = {}
        con_s = _timed(torch, lambda: got.setdefault("r", s.contains(qk)))
        if int(got["r"].sum()) != q * 4 // 5:
            raise RuntimeError("set at scale: wrong contains results")
        del qk, got
        s.set_tuning(mut)
        rm_s, removed = 0.0, 0
        for lo in range(0, total, chunk):
            k = wl.keys(chunk, 8, first=lo, seed=0)[::2].contiguous()  # half of the keys
            rm_s += _timed(torch, lambda: s.remove(k))
            removed += k.numel()
            del k
        st = s.stats()
        if s.status() or st["present"] != total - removed:
            raise RuntimeError("set at scale: wrong contents")
        s.close()
        res.update({f"set_insert_{name}_mops": round(total / ins_s / 1e6, 1),
                    f"set_contains_mops": round(q / con_s / 1e6, 1),
                    f"set_remove_{name}_mops": round(removed / rm_s / 1e6, 1)})
    return res


def reference_harness(lslab_amd, torch, reps=100):
    """The reference's own benchmark (benchmark/benchmark.cu:18-177):
    map<u64,u32> with 2^20 buckets, keys 1..2^19 put with value 0, then
    batches of 1024 keys uniform over [0, 2^21] (a seeded generator; the
    reference draws rand()), first gets, then puts of value 0, one call per
    batch timed by wall clock around the launch and the stream
    synchronisation, as the reference times it; the kernel time from HIP
    events on the same stream is reported beside it.  reps batches per
    series (the reference prints 10); puts with both engines."""
    import numpy as np
    from lslab_amd import _lib
    lib = _lib.load()
    pop, rng_hi, batch = 1 << 19, 1 << 21, 1024
    res = {"workload": "map<u64,u32> 2^20 buckets, keys 1..2^19 prefilled; batches of 1024 "
                       "keys uniform over [0, 2^21]; one call per batch",
           "reps": reps}
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    g = np.random.default_rng(0xBE7C)
    kb = [torch.from_numpy(g.integers(0, rng_hi + 1, size=batch, dtype=np.int64)).cuda()
          for _ in range(reps + 3)]
    zero = torch.zeros(batch, dtype=torch.int32, device="cuda")
    pairs = torch.empty((batch, 2), dtype=torch.int32, device="cuda")
    old = torch.empty(batch, dtype=torch.int32, device="cuda")

    def series(call):
        wall, kern = [], []
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for i in range(reps + 3):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            e0.record(s)
            _lib.check(call(kb[i]), "reference harness")
            e1.record(s)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            if i >= 3:  # the first calls warm the launch path
                wall.append(dt)
                kern.append(e0.elapsed_time(e1) / 1e3)
        w = np.array(wall)
        return {"latency_us_p50": round(float(np.percentile(w, 50)) * 1e6, 2),
                "latency_us_p99": round(float(np.percentile(w, 99)) * 1e6, 2),
                "mops_p50": round(batch / float(np.percentile(w, 50)) / 1e6, 2),
                "kernel_us_p50": round(float(np.percentile(kern, 50)) * 1e6, 2)}

    for engine, name in ((lslab_amd.ENGINE_PLAIN, "plain"), (lslab_amd.ENGINE_TWO_PHASE, "two_phase")):
        m = lslab_amd.Map(8, 4, 20, 10000, lslab_amd.tuning(lanes_per_op=8, engine=engine,
                                                           max_subbatch=1 << 19))
        m.put(torch.arange(1, pop + 1, dtype=torch.int64, device="cuda"),
              torch.zeros(pop, dtype=torch.int32, device="cuda"))
        if name == "plain":
            m.set_tuning(lslab_amd.tuning(lanes_per_op=2, engine=engine, max_subbatch=1 << 19))
            res["get"] = series(lambda k: lib.lslab_map_get(
                m.handle, ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(pairs.data_ptr()), batch, sp))
            m.set_tuning(lslab_amd.tuning(lanes_per_op=8, engine=engine, max_subbatch=1 << 19))
        res[f"put_{name}"] = series(lambda k: lib.lslab_map_put(
            m.handle, ctypes.c_void_p(k.data_ptr()), ctypes.c_void_p(zero.data_ptr()),
            ctypes.c_void_p(old.data_ptr()), batch, sp))
        if m.status():
            raise RuntimeError("reference harness status")
        m.close()
    # The same workload from C++ over the header-only API (tests/cpp/ref_harness,
    # built with the library), without Python's per-call overhead.
    exe = os.path.join(ROOT, "tests", "cpp", "ref_harness")
    if os.path.exists(exe):
        import subprocess
        r = subprocess.run([exe, str(reps)], capture_output=True, text=True, timeout=120)
        if r.returncode != 0:
            raise RuntimeError("ref_harness: " + r.stderr[-400:])
        res["cpp"] = json.loads(r.stdout.strip().splitlines()[-1])
    return res


def cpu_baseline(ops, keys, vals, nb_sample):
    """The oracle's sequential replay of nb_sample batches after the same prefill."""
    import numpy as np
    from oracle.oracle import Replay
    r = Replay(True, 8, 8, N_LOG2, pool_slabs=1 << 16)
    n = 1 << SPACE_LOG2
    # Same prefill values as the device (murmur3 fmix64 of i ^ seed).
    i = np.arange(n, dtype=np.uint64) ^ np.uint64(PREFILL_SEED)
    with np.errstate(over="ignore"):
        z = i ^ (i >> np.uint64(33))
        z = z * np.uint64(0xFF51AFD7ED558CCD)
        z ^= z >> np.uint64(33)
        z = z * np.uint64(0xC4CEB9FE1A85EC53)
        z ^= z >> np.uint64(33)
    t0 = time.perf_counter()
    r.apply(np.full(n, 1, np.uint8), np.arange(n, dtype=np.uint64), z)
    prefill_s = time.perf_counter() - t0
    m = nb_sample * BATCH
    h_ops = ops[:m].cpu().numpy()
    h_keys = keys[:m].cpu().numpy().view(np.uint64)
    h_vals = vals[:m].cpu().numpy().view(np.uint64)
    t0 = time.perf_counter()
    r.apply(h_ops, h_keys, h_vals)
    dt = time.perf_counter() - t0
    r.close()
    return {"value": round(m / dt / 1e6, 3), "unit": "Mops/s", "cores": 1, "kind": "port",
            "sample": f"{nb_sample} batches x 2^20 ops of the same stream after a 2^26-key "
                      f"prefill ({prefill_s:.1f} s, untimed), oracle sequential replay",
            "cpu": _cpu_model()}


def cpu_baseline_mt(ops, keys, vals, nb_sample, threads):
    """The multithreaded host baseline (sharded std::unordered_map, one
    shared_mutex per shard, a barrier between batches) on the same prefill
    and the same nb_sample batches."""
    import numpy as np
    from oracle.oracle import MtBaseline
    mt = MtBaseline(threads, N_LOG2, True)
    n = 1 << SPACE_LOG2
    mt.reserve(n)
    i = np.arange(n, dtype=np.uint64) ^ np.uint64(PREFILL_SEED)
    with np.errstate(over="ignore"):
        z = i ^ (i >> np.uint64(33))
        z = z * np.uint64(0xFF51AFD7ED558CCD)
        z ^= z >> np.uint64(33)
        z = z * np.uint64(0xC4CEB9FE1A85EC53)
        z ^= z >> np.uint64(33)
    pre_s, _, _ = mt.run(np.array([0, n], np.uint64), np.full(n, 1, np.uint8),
                         np.arange(n, dtype=np.uint64), z)
    del z, i
    m = nb_sample * BATCH
    offs = np.arange(0, m + 1, BATCH, dtype=np.uint64)
    dt, _, _ = mt.run(offs, ops[:m].cpu().numpy(), keys[:m].cpu().numpy().view(np.uint64),
                      vals[:m].cpu().numpy().view(np.uint64))
    mt.close()
    return {"value": round(m / dt / 1e6, 3), "unit": "Mops/s", "cores": threads, "kind": "port",
            "sample": f"{nb_sample} batches x 2^20 ops of the same stream after a 2^26-key "
                      f"prefill ({pre_s:.1f} s, untimed), std::unordered_map in "
                      f"{16 * threads} shards, {threads} threads, barrier between batches",
            "cpu": _cpu_model()}


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" x{os.cpu_count()}"
    except OSError:
        pass
    return "unknown"


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--lanes", type=int, default=0, help="override the tuned lanes per op")
    ap.add_argument("--combine", type=int, default=0,
                    help="override the tuned combining scope (1 wavefront, 2 block)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-batches", type=int, default=2)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--no-host-input", action="store_true")
    ap.add_argument("--no-wide", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-reference-harness", action="store_true")
    ap.add_argument("--served-requests", type=int, default=20000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first, as the HBM byte-counter reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(a multi-GPU table would shard by bucket; this benchmark runs one GPU's share)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import lslab_amd
    from lslab_amd import autotune
    from lslab_amd import workload as wl

    torch.cuda.set_device(0)
    K, W = args.steps, args.warmup
    tuned = {k: autotune.load(k) for k in ("headline", "batched", "served", "headline_plain",
                                           "batched_plain")}
    for v in tuned.values():
        v.setdefault("combine", lslab_amd.COMBINE_BLOCK)
        if args.lanes:
            v["lanes_per_op"] = args.lanes
        if args.combine:
            v["combine"] = args.combine
    if args.copy_calibration:
        # 1 GiB streaming copy: the byte counters' scale for streaming access
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        # random 128-B rows: 2^22 of 2^24 rows of a 2 GiB tensor (512 MiB read at
        # random, 32 MiB of indices, 512 MiB written): the scale for random lines
        rows = torch.zeros((1 << 24, 16), dtype=torch.int64, device="cuda")
        g = torch.Generator(device="cuda")
        g.manual_seed(7)
        idx = torch.randint(0, 1 << 24, (1 << 22,), device="cuda", generator=g)
        picked = torch.index_select(rows, 0, idx)
        torch.cuda.synchronize()
        del rows, idx, picked
    th = tuned["headline"]
    plain_lanes = tuned["headline_plain"]["lanes_per_op"]
    m = lslab_amd.Map(8, 8, N_LOG2, 1 << 16,
                      lslab_amd.tuning(lanes_per_op=plain_lanes,
                                       engine=lslab_amd.ENGINE_PLAIN, max_subbatch=BATCH))
    t0 = time.perf_counter()
    prefill(m, wl)
    torch.cuda.synchronize()
    log(f"prefill 2^26 keys: {time.perf_counter() - t0:.2f} s", m.stats())

    ops, keys, vals, offs = make_batches(wl, W + K, 0)
    torch.cuda.synchronize()

    time_plain(m, ops, keys, vals, W, torch)
    wall_p, per_p = time_plain(m, ops[W * BATCH:], keys[W * BATCH:], vals[W * BATCH:], K, torch)
    plain_mops = K * BATCH / wall_p / 1e6
    log(f"plain: {wall_p * 1e3:.2f} ms for {K} batches -> {plain_mops:.1f} Mops/s")
    st = m.status()
    if st:
        raise RuntimeError(f"plain status {st}")

    # Batches of max_subbatch ops never share a run: the call skips the
    # device-side check and launches the all-alone kernel only.
    m.set_tuning(lslab_amd.tuning(lanes_per_op=th["lanes_per_op"], engine=lslab_amd.ENGINE_TWO_PHASE,
                                  max_blocks=th["max_blocks"], max_subbatch=BATCH,
                                  combine=th["combine"], runs=lslab_amd.RUNS_OFF))
    time_fused(m, ops, keys, vals, offs[: W + 1], W, torch)
    offs_k = offs[W:] - W * BATCH
    wall_f, kern_f = time_fused(m, ops[W * BATCH:], keys[W * BATCH:], vals[W * BATCH:], offs_k, K,
                                torch)
    st = m.status()
    if st:
        raise RuntimeError(f"fused status {st}")
    fused_mops = K * BATCH / wall_f / 1e6
    log(f"fused: {wall_f * 1e3:.2f} ms wall, {kern_f * 1e3:.2f} ms kernel -> {fused_mops:.1f} Mops/s")

    achieved = B_OP * K * BATCH / kern_f / 1e9
    tpo = measured_traffic_per_op()
    per_sorted = sorted(per_p)
    res = {
        "metric": METRIC,
        "value": round(fused_mops, 2),
        "unit": "Mops/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": round(wall_f * 1e3 / K, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64",
        "data": "synthetic",
        "config": {
            "workload": "map<u64,u64> 2^22 buckets prefilled with 2^26 keys; batches of 2^20 "
                        "ops, 95% get / 5% put, Zipf 0.99 over 2^26 keys",
            "engine": "two-phase, one persistent launch for all timed batches",
            "lanes_per_op": th["lanes_per_op"],
            "max_blocks": th["max_blocks"],
            "combine": th["combine"],
            "global_batch": BATCH,
            "parallelism": "single GPU",
        },
        "plain_mops": round(plain_mops, 2),
        "fused_vs_plain": round(fused_mops / plain_mops, 3),
        "plain_best_lanes": plain_lanes,
        "plain_lanes_sweep": {str(e["lanes_per_op"]): e["mops"]
                              for e in tuned["headline_plain"].get("all", [])},
        "plain_batch_ms_p50": round(per_sorted[len(per_sorted) // 2], 4),
        "plain_batch_ms_p99": round(per_sorted[min(len(per_sorted) - 1,
                                                   int(0.99 * len(per_sorted)))], 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": round(tpo * K * BATCH) if tpo else None,
            # the HBM byte rate the launch actually moved (rocprof bytes per op x
            # ops/s): the kernel is bound by dependent random accesses, not bytes
            "measured_gbs": round(tpo * K * BATCH / kern_f / 1e9, 1) if tpo else None,
            "measured_frac": round(tpo * K * BATCH / kern_f / 1e9 / HBM_PEAK_GBS, 4) if tpo else None,
            "achieved_is": "algorithmic bytes (reference slab layout) / kernel time",
            "traffic_source": "rocprofv3 FETCH_SIZE (x1.063, random-line calibration) + WRITE_SIZE per op "
                              "of k_batches, profiles/k_batches_map_u64_zipf099_headline.md, "
                              "times the ops of this launch",
            "kernel": "k_batches<Layout<u64,u64,map>, lanes_per_op>",
            "bytes_per_op": B_OP,
        },
    }
    if not args.no_host_input:
        log("host-memory batches: table over PCIe vs KVCG hot/cold split ...")
        res["host_input"] = host_input(lslab_amd, torch, m, th, ops[W * BATCH:],
                                       keys[W * BATCH:], vals[W * BATCH:], min(K, 8))
        log(res["host_input"])
    if not args.no_configs:
        log("plain small map and set at scale ...")
        res["other_configs"] = other_configs(lslab_amd, wl, torch)
        log(res["other_configs"])
    if not args.no_wide:
        log("wide keys ...")
        res["wide"] = wide_keys(lslab_amd, torch)
        log(res["wide"])
    if not args.no_batched:
        log("batched form (4096 uneven batches) ...")
        res["batched"] = batched(lslab_amd, wl, torch, m, tuned["batched"],
                                 tuned["batched_plain"]["lanes_per_op"])
        log(res["batched"])
    m.close()
    if not args.no_reference_harness:
        log("reference harness workload (1024-key batches) ...")
        res["reference_harness"] = reference_harness(lslab_amd, torch)
        log(res["reference_harness"])
    if not args.no_served:
        log("served form ...")
        sv = served(lslab_amd, wl, torch, tuned["served"], args.served_requests)
        res["served"] = sv
        res["p50_us"], res["p99_us"] = sv["p50_us"], sv["p99_us"]
        log(sv)
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(ops[W * BATCH:], keys[W * BATCH:], vals[W * BATCH:],
                                           args.cpu_batches)
        log("multithreaded cpu baseline ...")
        res["cpu_baseline_mt"] = cpu_baseline_mt(ops[W * BATCH:], keys[W * BATCH:],
                                                 vals[W * BATCH:], 4 * args.cpu_batches,
                                                 min(16, os.cpu_count() or 1))
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
