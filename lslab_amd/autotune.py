"""Launch-shape autotuner for the batched (persistent, two-phase) form.

Sweeps, on a short sample of the workload it is given:
  - lanes_per_op   : lanes that probe one slab per op (4, 8, 16, 32) -- the
                     slab-probe width: 32 lanes read a u64 slab's 256-B key
                     block one slot per lane (the reference's warp-per-op
                     shape on a half wave), 4 lanes read 64 B each and keep
                     16 ops in flight per wavefront
  - block_threads  : 64 .. 256
  - grid fraction  : the persistent grid as 1, 1/2 or 1/4 of full residency
                     (fewer blocks make each grid barrier cheaper)
and keeps the fastest setting (Mops/s of one exec_batches launch).  Results
are stored in lslab_amd/tuning.json keyed by workload name, which the
benchmark reads.

python -m lslab_amd.autotune   (runs the headline workload's search)
"""
import ctypes
import itertools
import json
import os
import time

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_JSON = os.path.join(HERE, "tuning.json")


def load(name):
    try:
        with open(TUNING_JSON) as f:
            return json.load(f).get(name)
    except (OSError, ValueError):
        return None


def store(name, entry):
    data = {}
    try:
        with open(TUNING_JSON) as f:
            data = json.load(f)
    except (OSError, ValueError):
        pass
    data[name] = entry
    with open(TUNING_JSON, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)


def _occupancy_blocks(torch):
    props = torch.cuda.get_device_properties(0)
    return props.multi_processor_count


def search(table, offsets, ops, keys, vals, nbatches, n_ops, max_subbatch,
           lanes=(4, 8, 16, 32), blocks=(256, 128), grid_fracs=(1.0, 0.5, 0.25), reps=2,
           log=print):
    """Times exec_batches over the given batches for each setting; returns
    (best tuning dict, all results).  The table's state changes as batches
    run; every setting sees the same op stream on a table of the same shape."""
    import torch

    from . import _lib, table as tb
    lib = _lib.load()
    out = torch.empty((n_ops, 2), dtype=torch.int64, device="cuda")
    cus = _occupancy_blocks(torch)
    results = []
    for g, bt, frac in itertools.product(lanes, blocks, grid_fracs):
        # per-CU residency is probed by launching with max_blocks = 0 first
        max_blocks = 0 if frac == 1.0 else max(1, int(cus * 8 * (256 // bt) * frac))
        tun = tb.tuning(lanes_per_op=g, engine=_lib.ENGINE_TWO_PHASE, block_threads=bt,
                        max_blocks=max_blocks, max_subbatch=max_subbatch)
        table.set_tuning(tun)
        s = torch.cuda.current_stream()
        best = None
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            rc = lib.lslab_exec_batches(table.handle, nbatches, ctypes.c_void_p(offsets.data_ptr()),
                                        ctypes.c_void_p(ops.data_ptr()),
                                        ctypes.c_void_p(keys.data_ptr()),
                                        ctypes.c_void_p(vals.data_ptr()) if vals is not None else None,
                                        ctypes.c_void_p(out.data_ptr()),
                                        ctypes.c_void_p(s.cuda_stream))
            _lib.check(rc, "exec_batches")
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            best = dt if best is None else min(best, dt)
        st = table.status()
        if st:
            raise RuntimeError(f"status {st} during autotune")
        mops = n_ops / best / 1e6
        results.append({"lanes_per_op": g, "block_threads": bt, "max_blocks": max_blocks,
                        "grid_frac": frac, "mops": round(mops, 2)})
        log(f"[autotune] lanes={g} block={bt} grid={frac}: {mops:.1f} Mops/s")
    top = max(results, key=lambda r: r["mops"])
    return top, results


def main():
    import sys
    sys.path.insert(0, os.path.dirname(HERE))
    import torch

    import lslab_amd
    from lslab_amd import workload as wl
    import bench

    m = lslab_amd.Map(8, 8, bench.N_LOG2, 1 << 16,
                      lslab_amd.tuning(engine=lslab_amd.ENGINE_TWO_PHASE, max_subbatch=bench.BATCH))
    bench.prefill(m, wl)
    nb = 4
    ops, keys, vals, offs = bench.make_batches(wl, nb, 1000)
    top, res = search(m, offs, ops, keys, vals, nb, nb * bench.BATCH, bench.BATCH)
    entry = dict(top)
    entry["all"] = res
    entry["device"] = torch.cuda.get_device_name(0)
    store("map_u64_u64_zipf099_95get", entry)
    print(json.dumps(top))


if __name__ == "__main__":
    main()
