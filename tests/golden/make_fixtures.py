"""Writes the small golden fixtures in this folder (inputs and expected
outputs of op streams, produced by the oracle's sequential replay).

The reference holds no golden vectors or fixtures, so these are the build's
own: each case is replayed in op order by oracle/lslab_oracle.c, whose own
pins are the reference's known-answer tests, a dict model and closed forms
(tests/test_oracle.py).  Every batch of a case touches distinct keys, so the
expected results are the only valid ones for any engine.

python tests/golden/make_fixtures.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle.oracle import Replay  # noqa: E402

CASES = [
    # name, is_map, key bytes, value bytes, n_log2, pool, seed
    ("map_u32_u32_chains", True, 4, 4, 3, 256, 1),
    ("map_u64_u64_chains", True, 8, 8, 3, 256, 2),
    ("map_u32_u64_wide_keys", True, 4, 8, 5, 256, 3),
    ("map_u64_u32_edge_keys", True, 8, 4, 2, 256, 4),
    ("set_u64_chains", False, 8, 0, 3, 256, 5),
    ("set_u32_chains", False, 4, 0, 3, 256, 6),
]


def make_case(is_map, kb, vb, n_log2, pool, seed):
    rng = np.random.default_rng(seed)
    kmax = 2**32 if kb == 4 else 2**64
    universe = np.unique(rng.integers(0, kmax, size=900, dtype=np.uint64))
    universe[:3] = [0, kmax - 1, kmax >> 1]  # width edges, incl. the all-ones key
    universe = np.unique(universe)
    rng.shuffle(universe)
    cap = 31  # keys per slab, every layout (detail/slab_node.h:13-35)
    r = Replay(is_map, kb, vb, n_log2, pool_slabs=pool, slab_keys=cap)
    batches = []
    for b in range(8):
        n = int(rng.integers(50, 400))
        keys = rng.permutation(universe)[:n]  # distinct within the batch
        p = [0.2, 0.7, 0.1] if b < 2 else [0.4, 0.35, 0.25]
        ops = rng.choice([0, 1, 2], size=n, p=p).astype(np.uint8)
        vals = rng.integers(0, 2**64, size=n, dtype=np.uint64)
        if vb == 4:
            vals &= np.uint64(0xFFFFFFFF)
        f, v, fail = r.apply(ops, keys, vals)
        assert fail == 0
        batches.append((ops, keys, vals, f, v))
    dk, dv = r.dump()
    o = np.argsort(dk)
    offs = np.cumsum([0] + [len(x[0]) for x in batches]).astype(np.int64)
    return dict(
        is_map=np.array(is_map), key_bytes=np.array(kb), val_bytes=np.array(vb),
        n_log2=np.array(n_log2), pool=np.array(pool), slab_keys=np.array(cap),
        offsets=offs,
        ops=np.concatenate([x[0] for x in batches]),
        keys=np.concatenate([x[1] for x in batches]),
        vals=np.concatenate([x[2] for x in batches]),
        found=np.concatenate([x[3] for x in batches]),
        out=np.concatenate([x[4] for x in batches]),
        final_keys=dk[o], final_vals=dv[o], slabs_used=np.array(r.slabs_used),
    )


def main():
    for name, *args in CASES:
        d = make_case(*args)
        np.savez_compressed(os.path.join(HERE, name + ".npz"), **d)
        print(name, len(d["keys"]), "ops,", len(d["final_keys"]), "keys at the end")


if __name__ == "__main__":
    main()
