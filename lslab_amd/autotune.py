"""Launch-shape autotuner for the persistent forms.

For each timed workload it sweeps, on a short sample of that workload:
  - lanes_per_op : lanes that probe one slab per op (1, 2, 4, 8, ...).  This
                   is the slab-probe width: 32 lanes read a u64 slab's 256-B key
                   block one slot per lane (the reference's warp-per-op shape on
                   a half wave); 2 lanes read 128 B each and keep 32 ops in
                   flight per wavefront.
  - max_blocks   : persistent grid size (one 1024-thread block per CU at
                   most); fewer blocks make each grid barrier cheaper, which
                   pays off for small batches.
  - combine      : combining scope of same-key mutations before the global
                   exchange (1 = within a wavefront, 2 = across the block);
  - tick ops     : (served form) ops admitted per tick.
and keeps the setting with the highest Mops/s.  It also sweeps lanes_per_op of
the plain per-batch engine (bucket locks, 1..32 lanes, 32 being the
reference's warp-per-op shape) on the headline and batched streams, the
baselines the fused forms are compared with.  Results go to
lslab_amd/tuning.json under the workload's name; bench.py reads them.

python -m lslab_amd.autotune
"""
import ctypes
import itertools
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_JSON = os.path.join(HERE, "tuning.json")

DEFAULTS = {
    "headline": {"lanes_per_op": 1, "max_blocks": 0, "combine": 1},
    "batched": {"lanes_per_op": 1, "max_blocks": 0, "combine": 1},
    "served": {"lanes_per_op": 2, "max_blocks": 128, "tick_ops": 1 << 18, "combine": 2},
    # the plain per-batch engine (bucket locks) on the same two streams
    "headline_plain": {"lanes_per_op": 2},
    "batched_plain": {"lanes_per_op": 2},
}
PLAIN_LANES = (1, 2, 4, 8, 16, 32)  # 32: one half-wave per op, the reference's warp per op


def load(name):
    """Tuned entry for a workload, or its default."""
    try:
        with open(TUNING_JSON) as f:
            e = json.load(f).get(name)
            if e:
                return e
    except (OSError, ValueError):
        pass
    return dict(DEFAULTS[name])


def _timed(fn, reps=2):
    import torch
    best = None
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    return best


def _exec(table, offs, ops, keys, vals, out, nb):
    import torch
    from . import _lib
    lib = _lib.load()
    rc = lib.lslab_exec_batches(table.handle, nb, ctypes.c_void_p(offs.data_ptr()),
                                ctypes.c_void_p(ops.data_ptr()), ctypes.c_void_p(keys.data_ptr()),
                                ctypes.c_void_p(vals.data_ptr()), ctypes.c_void_p(out.data_ptr()),
                                ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _lib.check(rc, "exec_batches")


def search_batches(table, offs, ops, keys, vals, nb, n_ops, max_subbatch, lanes, grids, log):
    import torch
    from . import _lib, table as tb
    out = torch.empty((n_ops, 2), dtype=torch.int64, device="cuda")
    res = []
    for g, mb, cb in itertools.product(lanes, grids, (_lib.COMBINE_WAVE, _lib.COMBINE_BLOCK)):
        table.set_tuning(tb.tuning(lanes_per_op=g, engine=_lib.ENGINE_TWO_PHASE, max_blocks=mb,
                                   max_subbatch=max_subbatch, combine=cb))
        dt = _timed(lambda: _exec(table, offs, ops, keys, vals, out, nb))
        st = table.status()
        if st:
            raise RuntimeError(f"status {st} during autotune")
        res.append({"lanes_per_op": g, "max_blocks": mb, "combine": cb,
                    "mops": round(n_ops / dt / 1e6, 1)})
        log(f"[autotune] lanes={g} blocks={mb or 'all'} combine={cb}: {res[-1]['mops']} Mops/s")
    return max(res, key=lambda r: r["mops"]), res


def search_plain(table, offs_h, ops, keys, vals, nb, log, name):
    """Lanes per op of the plain engine: each batch one locked launch, back to
    back on one stream, as the benchmark times it."""
    import torch
    from . import _lib, table as tb
    lib = _lib.load()
    n = int(offs_h[nb])
    out = torch.empty((n, 2), dtype=torch.int64, device="cuda")
    sp = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    res = []
    for g in PLAIN_LANES:
        table.set_tuning(tb.tuning(lanes_per_op=g, engine=_lib.ENGINE_PLAIN, max_subbatch=1 << 20))

        def run():
            for b in range(nb):
                lo = int(offs_h[b])
                _lib.check(lib.lslab_map_mixed(table.handle, ctypes.c_void_p(ops[lo:].data_ptr()),
                                               ctypes.c_void_p(keys[lo:].data_ptr()),
                                               ctypes.c_void_p(vals[lo:].data_ptr()),
                                               ctypes.c_void_p(out[lo:].data_ptr()),
                                               int(offs_h[b + 1] - lo), sp), "map_mixed")
        dt = _timed(run, reps=1)
        if table.status():
            raise RuntimeError("status during the plain sweep")
        res.append({"lanes_per_op": g, "mops": round(n / dt / 1e6, 1)})
        log(f"[autotune] {name} plain lanes={g}: {res[-1]['mops']} Mops/s")
    return max(res, key=lambda r: r["mops"]), res


def main():
    sys.path.insert(0, os.path.dirname(HERE))
    import numpy as np
    import torch

    import bench
    import lslab_amd
    from lslab_amd import workload as wl

    log = lambda *a: print(*a, flush=True)  # noqa: E731
    out = {"device": torch.cuda.get_device_name(0)}
    m = lslab_amd.Map(8, 8, bench.N_LOG2, 1 << 16,
                      lslab_amd.tuning(engine=lslab_amd.ENGINE_TWO_PHASE, max_subbatch=bench.BATCH))
    bench.prefill(m, wl)
    # headline: 4 batches of 2^20 ops
    nb = 4
    ops, keys, vals, offs = bench.make_batches(wl, nb, 1000)
    top, res = search_batches(m, offs, ops, keys, vals, nb, nb * bench.BATCH, bench.BATCH,
                              (1, 2, 4, 8), (0, 192, 128), log)
    out["headline"] = dict(top, all=res)
    top, res = search_plain(m, offs.cpu().numpy(), ops, keys, vals, 3, log, "headline")
    out["headline_plain"] = dict(top, all=res)
    # batched: the first 1024 of the uneven batches
    offs_h, ops, keys, vals = bench.gen_batched(wl, 1024)
    offs = torch.from_numpy(offs_h).cuda()
    top, res = search_batches(m, offs, ops, keys, vals, 1024, int(offs_h[-1]), bench.BATCH,
                              (1, 2, 4, 8), (0, 128, 64), log)
    out["batched"] = dict(top, all=res)
    top, res = search_plain(m, offs_h, ops, keys, vals, 256, log, "batched")
    out["batched_plain"] = dict(top, all=res)
    m.close()
    # served: loop capacity (every request arrives at t=0) on a short stream
    res = []
    for g, tick, mb, cb in itertools.product((1, 2, 4), (1 << 16, 1 << 17, 1 << 18), (0, 128),
                                             (lslab_amd.COMBINE_WAVE, lslab_amd.COMBINE_BLOCK)):
        sv = bench.served_capacity(lslab_amd, wl, torch, g, 4000, tick, mb, cb)
        res.append({"lanes_per_op": g, "tick_ops": tick, "max_blocks": mb, "combine": cb,
                    "mops": sv})
        log(f"[autotune] served lanes={g} tick={tick} blocks={mb or 'all'} combine={cb}: "
            f"{sv} Mops/s")
    top = max(res, key=lambda r: r["mops"])
    out["served"] = dict(top, all=res)
    with open(TUNING_JSON, "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print(json.dumps({k: ({kk: vv for kk, vv in v.items() if kk != "all"} if isinstance(v, dict) else v)
                      for k, v in out.items()}))


if __name__ == "__main__":
    main()
