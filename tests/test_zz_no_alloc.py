"""lslab_exec_batches_noalloc: the batched form with its scratch in the
library's static device array, called through the C ABI.

On twin tables (same shape, tuning and contents) the no-alloc entry point and
lslab_exec_batches run the same batches; their outputs and final contents
must be identical and equal to the oracle's sequential replay (batches of
distinct keys, so exactly one result is valid).  Two calls in a row, the
second smaller, so the second runs on what the first left in the array.
Sizes beyond the array are refused before any device work, and a graph
capture of the call replays correctly.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle.oracle import OP_PUT, Replay

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _h64(t):
    return t.contiguous().cpu().numpy().view(np.uint64)


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _call(lib, name, tbl, offs, ops, keys, vals, out):
    return getattr(lib, name)(tbl.handle, offs.numel() - 1, ctypes.c_void_p(offs.data_ptr()),
                              ctypes.c_void_p(ops.data_ptr()), ctypes.c_void_p(keys.data_ptr()),
                              ctypes.c_void_p(vals.data_ptr()) if vals is not None else None,
                              ctypes.c_void_p(out.data_ptr()), _stream())


def _stream_of(rng, sizes, universe):
    keys = np.concatenate([rng.permutation(universe)[:s] for s in sizes])  # distinct per batch
    n = len(keys)
    ops = rng.choice([0, 1, 2], size=n, p=[0.5, 0.35, 0.15]).astype(np.uint8)
    vals = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return offs, ops, keys, vals


@pytest.mark.parametrize("is_map", [True, False])
def test_noalloc_matches_exec_batches(lslab, oracle_mod, is_map):
    say("noalloc_matches_exec_batches", is_map)
    from lslab_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(70 + is_map)
    n_log2, pool = 9, 1 << 14
    tun = lslab.tuning(lanes_per_op=4, engine=1, max_subbatch=_lib.NOALLOC_MAX_SUBBATCH,
                       combine=1)
    mk = ((lambda: lslab.Map(8, 8, n_log2, pool, tun)) if is_map
          else (lambda: lslab.Set(8, n_log2, pool, tun)))
    a, b = mk(), mk()
    ref = Replay(is_map, 8, 8, n_log2, pool_slabs=pool)
    universe = np.arange(1, 300001, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    pre = universe[:20000]
    for t in (a, b):
        d = torch.from_numpy(pre.view(np.int64)).cuda()
        if is_map:
            t.put(d, d)
        else:
            t.insert(d)
    ref.apply(np.full(len(pre), OP_PUT, np.uint8), pre, pre if is_map else None)
    # first call: small batches sharing runs, one of exactly the sub-batch size,
    # one larger (runs alone in sub-batches), an empty one; second call smaller
    first = np.concatenate([rng.integers(1, 4000, size=40), [1 << 16, 0, 100000, 17]])
    second = rng.integers(1, 2000, size=12)
    for call, sizes in enumerate((first, second)):
        offs, ops, keys, vals = _stream_of(rng, sizes, universe)
        n = len(keys)
        d_offs = torch.from_numpy(offs).cuda()
        d_ops = torch.from_numpy(ops).cuda()
        d_keys = torch.from_numpy(keys.view(np.int64)).cuda()
        d_vals = torch.from_numpy(vals.view(np.int64)).cuda() if is_map else None
        shape = (n, 2) if is_map else (n,)
        dt = torch.int64 if is_map else torch.uint8
        out_a = torch.full(shape, 7, dtype=dt, device="cuda")
        out_b = torch.full(shape, 9, dtype=dt, device="cuda")
        _lib.check(_call(lib, "lslab_exec_batches", a, d_offs, d_ops, d_keys, d_vals, out_a),
                   "exec_batches")
        _lib.check(_call(lib, "lslab_exec_batches_noalloc", b, d_offs, d_ops, d_keys, d_vals,
                         out_b), "exec_batches_noalloc")
        torch.cuda.synchronize()
        assert a.status() == 0 and b.status() == 0
        assert torch.equal(out_a, out_b), call
        if is_map:
            f = out_b.view(torch.uint8).view(n, 16)[:, 0].cpu().numpy()
            v = _h64(out_b[:, 1])
        else:
            f = out_b.cpu().numpy()
            v = np.zeros(n, np.uint64)
        for bi in range(len(sizes)):
            lo, hi = offs[bi], offs[bi + 1]
            rf, rv, fail = ref.apply(ops[lo:hi], keys[lo:hi], vals[lo:hi] if is_map else None)
            assert fail == 0
            np.testing.assert_array_equal(f[lo:hi], rf, err_msg=f"call {call} batch {bi}")
            np.testing.assert_array_equal(v[lo:hi], rv, err_msg=f"call {call} batch {bi}")
    # one dump per table: the order of a dump is unspecified
    rk, rv = ref.dump()
    ro = np.argsort(rk)
    for t in (a, b):
        if is_map:
            k, v = t.dump()
            k, v = _h64(k), _h64(v)
        else:
            k, v = _h64(t.dump()), None
        o = np.argsort(k)
        np.testing.assert_array_equal(k[o], rk[ro])
        if is_map:
            np.testing.assert_array_equal(v[o], rv[ro])
    # Slab counts may differ: after removes, which holes later inserts fill
    # depends on the order concurrent inserts reach a chain, in any engine.
    assert a.stats()["present"] == b.stats()["present"] == ref.present
    a.close()
    b.close()


def test_noalloc_refuses_oversized_scratch(lslab):
    say("noalloc_refuses_oversized_scratch")
    from lslab_amd import _lib
    lib = _lib.load()
    m = lslab.Map(8, 8, 6, 1 << 10, lslab.tuning(engine=1,
                                                  max_subbatch=2 * _lib.NOALLOC_MAX_SUBBATCH))
    offs = torch.tensor([0, 4], dtype=torch.int64, device="cuda")
    ops = torch.ones(4, dtype=torch.uint8, device="cuda")
    keys = torch.arange(4, dtype=torch.int64, device="cuda")
    out = torch.zeros((4, 2), dtype=torch.int64, device="cuda")
    rc = _call(lib, "lslab_exec_batches_noalloc", m, offs, ops, keys, keys, out)
    assert rc == _lib.ERR_INVALID
    torch.cuda.synchronize()
    assert int(out.abs().sum()) == 0 and m.stats()["present"] == 0  # nothing ran
    m.set_tuning(lslab.tuning(engine=1, max_subbatch=_lib.NOALLOC_MAX_SUBBATCH))
    _lib.check(_call(lib, "lslab_exec_batches_noalloc", m, offs, ops, keys, keys, out), "noalloc")
    assert m.status() == 0 and m.stats()["present"] == 4
    m.close()


def test_noalloc_graph_capture(lslab, oracle_mod):
    say("noalloc_graph_capture")
    from lslab_amd import _lib
    lib = _lib.load()
    rng = np.random.default_rng(77)
    m = lslab.Map(8, 8, 7, 1 << 12, lslab.tuning(lanes_per_op=2, engine=1,
                                                  max_subbatch=1 << 14, combine=2))
    ref = Replay(True, 8, 8, 7, pool_slabs=1 << 12)
    universe = np.arange(1, 50001, dtype=np.uint64) * np.uint64(0x2545F4914F6CDD1D)
    offs, ops, keys, vals = _stream_of(rng, rng.integers(1, 3000, size=20), universe)
    n = len(keys)
    d_offs = torch.from_numpy(offs).cuda()
    d_ops = torch.from_numpy(ops).cuda()
    d_keys = torch.from_numpy(keys.view(np.int64)).cuda()
    d_vals = torch.from_numpy(vals.view(np.int64)).cuda()
    out = torch.zeros((n, 2), dtype=torch.int64, device="cuda")
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        _lib.check(_call(lib, "lslab_exec_batches_noalloc", m, d_offs, d_ops, d_keys, d_vals,
                         out), "noalloc")
    for rep in range(3):
        out.zero_()
        g.replay()
        torch.cuda.synchronize()
        assert m.status() == 0
        f = out.view(torch.uint8).view(n, 16)[:, 0].cpu().numpy()
        v = _h64(out[:, 1])
        for bi in range(len(offs) - 1):
            lo, hi = offs[bi], offs[bi + 1]
            rf, rv, _ = ref.apply(ops[lo:hi], keys[lo:hi], vals[lo:hi])
            np.testing.assert_array_equal(f[lo:hi], rf, err_msg=f"replay {rep} batch {bi}")
            np.testing.assert_array_equal(v[lo:hi], rv, err_msg=f"replay {rep} batch {bi}")
        pk = universe[rng.integers(0, len(universe), size=500)]
        ef, ev = m.get(torch.from_numpy(pk.view(np.int64)).cuda())
        rf, rv, _ = ref.apply(np.zeros(len(pk), np.uint8), pk)
        np.testing.assert_array_equal(ef.cpu().numpy(), rf)
        np.testing.assert_array_equal(_h64(ev), rv)
    m.close()
