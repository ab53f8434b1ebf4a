"""Golden fixtures (tests/golden/*.npz, written by make_fixtures.py).

CPU: the oracle still reproduces every fixture (regression of the checker).
GPU: every engine and slab-probe width reproduces them through the C ABI,
op by op, plus the final table contents; the persistent batched form runs all
batches of a fixture in one launch.
"""
import glob
import os

import numpy as np
import pytest

from oracle.oracle import Replay

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURES = sorted(glob.glob(os.path.join(HERE, "golden", "*.npz")))


def load(path):
    d = dict(np.load(path, allow_pickle=False))
    d["name"] = os.path.basename(path)
    return d


@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_oracle_reproduces_fixture(oracle_mod, path):
    d = load(path)
    r = Replay(bool(d["is_map"]), int(d["key_bytes"]), int(d["val_bytes"]), int(d["n_log2"]),
               pool_slabs=int(d["pool"]), slab_keys=int(d["slab_keys"]))
    offs = d["offsets"]
    for b in range(len(offs) - 1):
        sl = slice(offs[b], offs[b + 1])
        f, v, fail = r.apply(d["ops"][sl], d["keys"][sl], d["vals"][sl])
        assert fail == 0
        np.testing.assert_array_equal(f, d["found"][sl])
        np.testing.assert_array_equal(v, d["out"][sl])
    k, v = r.dump()
    o = np.argsort(k)
    np.testing.assert_array_equal(k[o], d["final_keys"])
    np.testing.assert_array_equal(v[o], d["final_vals"])
    assert r.slabs_used == int(d["slabs_used"])


def _dev(a, b):
    import torch
    a = np.asarray(a, np.uint64)
    a = (a & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32) if b == 4 else a.view(np.int64)
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def _host(t, b):
    a = t.contiguous().cpu().numpy()
    return a.view(np.uint32).astype(np.uint64) if b == 4 else a.view(np.uint64)


def _final(tbl, is_map, kb, vb):
    if is_map:
        k, v = tbl.dump()
        k, v = _host(k, kb), _host(v, vb)
    else:
        k = _host(tbl.dump(), kb)
        v = np.zeros(len(k), np.uint64)
    o = np.argsort(k)
    return k[o], v[o]


@pytest.mark.gpu
@pytest.mark.parametrize("engine,combine", [(0, 2), (1, 1), (1, 2)])
@pytest.mark.parametrize("lanes", [1, 2, 4, 8, 16, 32])
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_reproduces_fixture(lslab, path, engine, combine, lanes):
    import torch
    d = load(path)
    print("[case] golden", d["name"], engine, combine, lanes, flush=True)
    is_map, kb, vb = bool(d["is_map"]), int(d["key_bytes"]), int(d["val_bytes"])
    tun = lslab.tuning(lanes_per_op=lanes, engine=engine, max_subbatch=1 << 12, combine=combine)
    tbl = (lslab.Map(kb, vb, int(d["n_log2"]), int(d["pool"]), tun) if is_map
           else lslab.Set(kb, int(d["n_log2"]), int(d["pool"]), tun))
    offs = d["offsets"]
    for b in range(len(offs) - 1):
        sl = slice(offs[b], offs[b + 1])
        ops = torch.from_numpy(d["ops"][sl].copy()).cuda()
        if is_map:
            f, v = tbl.mixed(ops, _dev(d["keys"][sl], kb), _dev(d["vals"][sl], vb))
            np.testing.assert_array_equal(_host(v, vb), d["out"][sl], err_msg=f"batch {b}")
        else:
            f = tbl.mixed(ops, _dev(d["keys"][sl], kb))
        np.testing.assert_array_equal(f.cpu().numpy(), d["found"][sl], err_msg=f"batch {b}")
    assert tbl.status() == 0
    k, v = _final(tbl, is_map, kb, vb)
    np.testing.assert_array_equal(k, d["final_keys"])
    np.testing.assert_array_equal(v, d["final_vals"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", FIXTURES, ids=[os.path.basename(p) for p in FIXTURES])
def test_gpu_batched_form_reproduces_fixture(lslab, path):
    import torch
    d = load(path)
    print("[case] golden batched", d["name"], flush=True)
    is_map, kb, vb = bool(d["is_map"]), int(d["key_bytes"]), int(d["val_bytes"])
    tun = lslab.tuning(lanes_per_op=8, engine=1, max_subbatch=1 << 12)
    tbl = (lslab.Map(kb, vb, int(d["n_log2"]), int(d["pool"]), tun) if is_map
           else lslab.Set(kb, int(d["n_log2"]), int(d["pool"]), tun))
    offs = torch.from_numpy(d["offsets"]).cuda()
    ops = torch.from_numpy(d["ops"]).cuda()
    if is_map:
        f, v = tbl.exec_batches(offs, ops, _dev(d["keys"], kb), _dev(d["vals"], vb))
        np.testing.assert_array_equal(_host(v, vb), d["out"])
    else:
        f = tbl.exec_batches(offs, ops, _dev(d["keys"], kb))
    np.testing.assert_array_equal(f.cpu().numpy(), d["found"])
    assert tbl.status() == 0
    k, v = _final(tbl, is_map, kb, vb)
    np.testing.assert_array_equal(k, d["final_keys"])
    np.testing.assert_array_equal(v, d["final_vals"])
