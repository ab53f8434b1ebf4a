"""Per-table call ordering across streams: a call on one stream runs after
the table's previous call on another stream, with no synchronisation by the
caller.  Each call waits on the table's order event, which the previous
call's last kernel launch records on completion; these tests give the
second call every chance to overtake the first (a large first call, the
second issued at once on an idle stream) and check that it did not."""
import numpy as np
import pytest
import torch

from test_gpu_parity import say

pytestmark = pytest.mark.gpu

N = 1 << 20


@pytest.mark.parametrize("engine", [0, 1])
def test_map_calls_ordered_across_streams(lslab, engine):
    """put on stream A, then get on stream B; remove on B, then get on A;
    exec_batches on A, then get on B."""
    say("map_calls_ordered_across_streams", engine)
    m = lslab.Map(8, 8, 16, 1 << 16, lslab.tuning(lanes_per_op=8, engine=engine))
    keys = torch.arange(1, N + 1, dtype=torch.int64, device="cuda")
    vals = keys * 3 + 1
    half = keys[: N // 2].contiguous()
    ops = torch.ones(N, dtype=torch.uint8, device="cuda")  # puts
    offs = torch.tensor([0, N // 4, N // 2, N], dtype=torch.int64, device="cuda")
    vals2 = keys * 5 + 2
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    m.put(keys, vals, stream=sa)
    f1, v1 = m.get(keys, stream=sb)
    m.remove(half, stream=sb)
    f2, _ = m.get(keys, stream=sa)
    m.exec_batches(offs, ops, keys, vals2, stream=sa)
    f3, v3 = m.get(keys, stream=sb)
    torch.cuda.synchronize()
    assert m.status() == 0
    assert bool((f1 == 1).all()) and torch.equal(v1, vals)
    f2 = f2.cpu().numpy()
    assert (f2[: N // 2] == 0).all() and (f2[N // 2:] == 1).all()
    assert bool((f3 == 1).all()) and torch.equal(v3, vals2)


def test_wide_calls_ordered_across_streams(lslab):
    """Wide-key table: put on stream A, then get on stream B."""
    say("wide_calls_ordered_across_streams")
    n = 1 << 18
    w = lslab.Wide(16, 8, 14, 1 << 17)
    rng = np.random.default_rng(77)
    k = rng.integers(0, 256, size=(n, 16), dtype=np.uint8)
    k[:, :8] = np.arange(n, dtype=np.uint64).view(np.uint8).reshape(n, 8)  # distinct keys
    v = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
    keys = torch.from_numpy(k).cuda()
    vals = torch.from_numpy(v).cuda()
    torch.cuda.synchronize()
    sa, sb = torch.cuda.Stream(), torch.cuda.Stream()
    w.put(keys, vals, stream=sa)
    found, got = w.get(keys, stream=sb)
    torch.cuda.synchronize()
    assert w.status() == 0
    assert bool((found == 1).all())
    assert torch.equal(got, vals)
