"""Benchmark: LSlab map on one MI355X, mixed get/put Zipf 0.99 over 2^26 keys.

Workload (a step = one batch of the hot path): map<u64,u64> with 2^22
buckets prefilled with keys 0..2^26-1 (16 keys per bucket), batches of 2^20
ops, 95% get / 5% put, keys = a fixed bijection of Zipf(0.99) ranks over 2^26,
put values unique per op.  Inputs are generated on the device and resident in
HBM before timing starts.

- plain: one per-batch launch of the wave64 plain kernel per batch (bucket
  reader/writer locks held per op), batches back to back on one stream;
- fused: the same number of batches in ONE persistent launch of the
  two-phase engine (lock-free reads, per-key exchange chains, grid barriers
  between phases and batches).

value = fused Mops/s.  The JSON line also carries the plain rate, the
roofline of the fused kernel (algorithmic bytes / HIP-event kernel time), and
the host-core baseline (the oracle's sequential replay of a bounded sample of
the same workload).
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
# request 8 B key (+8 B value and 1 B op for puts), result pair<bool,u64> 16 B,
# one slab probe of valid + 31 keys + next = 264 B, value 8 B read (get hit)
# or read+write 16 B (put): get 8+16+264+8 = 296 B, put 8+8+16+264+16 = 312 B
# minus the 8 B value already counted in the request = 304 B.
B_GET, B_PUT = 296.0, 304.0
B_OP = P_GET * B_GET + P_PUT * B_PUT  # 296.4
HBM_PEAK_GBS = 8000.0


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


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


def cpu_baseline(ops, keys, vals, nb_sample):
    """The oracle's sequential replay of nb_sample batches after the same prefill."""
    import numpy as np
    from oracle.oracle import Replay
    r = Replay(True, 8, 8, N_LOG2, pool_slabs=1 << 16)
    n = 1 << SPACE_LOG2
    # Same prefill values as the device (mix64(i ^ seed), murmur3 finaliser).
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


def _cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip() + f" x{os.cpu_count()}"
    except OSError:
        pass
    return "unknown"


def batched(lslab_amd, wl, torch, m, lanes, nbatches=4096):
    """Config 4: uneven batches (log-uniform sizes in [32, 65536]), 70/25/5
    get/put/remove, Zipf 0.99 over 2^26, on the prefilled table: one fused
    persistent launch vs one plain launch per batch."""
    import numpy as np
    from lslab_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(4096)
    sizes = np.rint(np.exp(rng.uniform(np.log(32), np.log(65536), size=nbatches))).astype(np.int64)
    offs_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs_h[-1])
    ops, keys, vals = wl.ops(n, THETA, SPACE_LOG2, 0.70, 0.25, 8, 8, seed=SEED + 4,
                             first_op=1 << 40)
    offs = torch.from_numpy(offs_h).cuda()
    out = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    s = torch.cuda.current_stream()
    sp = ctypes.c_void_p(s.cuda_stream)
    # plain: one launch per batch
    m.set_tuning(lslab_amd.tuning(lanes_per_op=lanes, engine=lslab_amd.ENGINE_PLAIN,
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
    # fused: one persistent launch, captured in a hipGraph
    m.set_tuning(lslab_amd.tuning(lanes_per_op=lanes, engine=lslab_amd.ENGINE_TWO_PHASE,
                                  max_subbatch=BATCH))
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
            "fused_vs_plain": round(plain_s / fused_s, 3)}


def served(lslab_amd, wl, torch, lanes, nreq, load_frac=0.7, tick_ops=1 << 18):
    """Config 5: Poisson arrivals of requests of U{1..4096} ops, 50/50
    get/put, Zipf 1.2 over 2^30, empty table, device-resident loop.  First the
    loop's capacity (every request arrives at t=0), then an open-loop run at
    load_frac of that rate for p50/p99 latency."""
    import numpy as np
    rng = np.random.default_rng(1200)
    sizes = rng.integers(1, 4097, size=nreq)
    offs_h = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offs_h[-1])
    ops, keys, vals = wl.ops(n, 1.2, 30, 0.5, 0.5, 8, 8, seed=SEED + 5, first_op=1 << 41)
    offs = torch.from_numpy(offs_h).cuda()
    cs = torch.cumsum((ops == 1).to(torch.int64), 0)
    cs = torch.cat([torch.zeros(1, dtype=torch.int64, device="cuda"), cs])
    reserve = (cs[offs[1:]] - cs[offs[:-1]]).to(torch.int32)

    def run(arrival_h):
        mm = lslab_amd.Map(8, 8, 24, 1 << 22,
                           lslab_amd.tuning(lanes_per_op=lanes, engine=lslab_amd.ENGINE_TWO_PHASE,
                                            max_subbatch=tick_ops))
        arr = torch.from_numpy(arrival_h.astype(np.int64)).cuda()
        torch.cuda.synchronize()
        res = mm.serve(offs, arr, ops, keys, vals, req_reserve=reserve, max_tick_ops=tick_ops)
        st = mm.status()
        mm.close()
        if st:
            raise RuntimeError(f"serve status {st}")
        comp = res["complete_ticks"].cpu().numpy()
        lat_us = (comp - arrival_h) / 100.0
        return res, lat_us

    res0, _ = run(np.zeros(nreq, np.int64))
    cap_s = res0["elapsed_ticks"] / 1e8
    cap_mops = res0["served_ops"] / cap_s / 1e6
    rate = load_frac * nreq / cap_s  # requests per second
    gaps = rng.exponential(1.0 / rate, size=nreq)
    arrival = np.floor(np.cumsum(gaps) * 1e8).astype(np.int64)
    arrival -= arrival[0]
    res1, lat = run(arrival)
    span = (res1["elapsed_ticks"]) / 1e8
    return {"requests": nreq, "ops": n, "capacity_mops": round(cap_mops, 2),
            "offered_load": load_frac, "offered_req_per_s": round(rate, 1),
            "mops": round(res1["served_ops"] / span / 1e6, 2),
            "p50_us": round(float(np.percentile(lat, 50)), 2),
            "p99_us": round(float(np.percentile(lat, 99)), 2),
            "ticks": int(res1["ticks"]), "rejected": int(res1["rejected"]),
            "max_tick_ops": tick_ops}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=32)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--lanes", type=int, default=4)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-batches", type=int, default=2)
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-served", action="store_true")
    ap.add_argument("--served-requests", type=int, default=20000)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run a 1 GiB device copy first, as the HBM byte-counter reference")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one GPU; see DESIGN.md for how the table would shard)",
              file=sys.stderr)
        sys.exit(2)
    import torch
    import lslab_amd
    from lslab_amd import workload as wl

    torch.cuda.set_device(0)
    K, W = args.steps, args.warmup
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.int64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    tun = lslab_amd.tuning(lanes_per_op=args.lanes, engine=lslab_amd.ENGINE_PLAIN,
                           max_subbatch=BATCH)
    m = lslab_amd.Map(8, 8, N_LOG2, 1 << 16, tun)
    t0 = time.perf_counter()
    prefill(m, wl)
    torch.cuda.synchronize()
    log(f"prefill 2^26 keys: {time.perf_counter() - t0:.2f} s", m.stats())

    ops, keys, vals, offs = make_batches(wl, W + K, 0)
    torch.cuda.synchronize()
    log("op stream ready")

    # plain: warmup then K timed per-batch launches
    time_plain(m, ops, keys, vals, W, torch)
    wall_p, per_p = time_plain(m, ops[W * BATCH:], keys[W * BATCH:], vals[W * BATCH:], K, torch)
    plain_mops = K * BATCH / wall_p / 1e6
    log(f"plain: {wall_p * 1e3:.2f} ms for {K} batches -> {plain_mops:.1f} Mops/s")
    st = m.status()
    if st:
        raise RuntimeError(f"plain status {st}")

    # fused: warmup launch over W batches, then one launch over the K timed batches
    m.set_tuning(lslab_amd.tuning(lanes_per_op=args.lanes, engine=lslab_amd.ENGINE_TWO_PHASE,
                                  max_subbatch=BATCH))
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
            "lanes_per_op": args.lanes,
            "global_batch": BATCH,
            "parallelism": "single GPU",
        },
        "plain_mops": round(plain_mops, 2),
        "fused_vs_plain": round(fused_mops / plain_mops, 3),
        "plain_batch_ms_p50": round(per_sorted[len(per_sorted) // 2], 4),
        "plain_batch_ms_p99": round(per_sorted[min(len(per_sorted) - 1,
                                                   int(0.99 * len(per_sorted)))], 4),
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "kernel": "k_batches<Layout<u64,u64,map>, G>",
            "bytes_per_op": B_OP,
        },
    }
    if not args.no_batched:
        log("batched form (4096 uneven batches) ...")
        res["batched"] = batched(lslab_amd, wl, torch, m, args.lanes)
        log(res["batched"])
    if not args.no_served:
        log("served form ...")
        sv = served(lslab_amd, wl, torch, args.lanes, args.served_requests)
        res["served"] = sv
        res["p50_us"], res["p99_us"] = sv["p50_us"], sv["p99_us"]
        log(sv)
    if not args.no_cpu:
        log("cpu baseline ...")
        res["cpu_baseline"] = cpu_baseline(ops[W * BATCH:], keys[W * BATCH:], vals[W * BATCH:],
                                           args.cpu_batches)
    print(json.dumps(res), flush=True)
    m.close()


if __name__ == "__main__":
    main()
