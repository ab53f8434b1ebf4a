"""Pins the wide-key oracle (oracle/wide_oracle.py) on the CPU: the
reference's legacy 128-B-key tests, a brute-force dict model, and closed
forms of the slab policy."""
import random
import struct

from oracle.wide_oracle import OP_GET, OP_PUT, OP_REMOVE, SLOTS, WideReplay, wide_hash


def key128(i):
    """The legacy test's Key(unsigned long long i): i in the first 8 bytes,
    zeros after (test/slabunified_test.cu.old:352-355)."""
    return struct.pack("<Q", i) + bytes(120)


def test_put_remove_128b_kat():
    """PutRemoveTest_128B (test/slabunified_test.cu.old:386-474): 100 rounds of
    inserting keys 1..1000 with fresh 8-byte values (each insert reads back
    the absent value) and removing them (each remove returns the value that
    round inserted)."""
    r = WideReplay(128, 8, 10, 10000)
    keys = [key128(j + 1) for j in range(1000)]
    for rep in range(100):
        vals = [struct.pack("<Q", (rep << 32) | (j + 1)) for j in range(1000)]
        f, old, fail = r.apply([OP_PUT] * 1000, keys, vals)
        assert fail == 0 and f == [0] * 1000 and old == [bytes(8)] * 1000
        f, old, fail = r.apply([OP_REMOVE] * 1000, keys)
        assert f == [1] * 1000 and old == vals
    assert r.present() == 0


def test_put_remove_128b_to_128b_kat():
    """PutRemoveTest_128Bto128B (test/slabunified_test.cu.old:476-541): the
    same rounds with 128-B values Key(rep)."""
    r = WideReplay(128, 128, 10, 10000)
    keys = [key128(j + 1) for j in range(1000)]
    for rep in range(100):
        vals = [key128(rep)] * 1000
        f, old, fail = r.apply([OP_PUT] * 1000, keys, vals)
        assert fail == 0 and f == [0] * 1000
        f, old, _ = r.apply([OP_REMOVE] * 1000, keys)
        assert f == [1] * 1000 and old == vals


def test_matches_dict_model():
    """Random get/put/remove streams over keys that share their first 8 bytes
    in small groups (long chains under the identity bucket): every result is
    the dict model's."""
    rng = random.Random(7)
    for kb, vb, mix in ((128, 8, False), (12, 0, False), (4, 4, True), (64, 128, True)):
        r = WideReplay(kb, vb, 2, 2000, mix_hash=mix)
        model = {}
        pool = []
        for i in range(300):
            head = struct.pack("<Q", i % 7)[: min(8, kb)]
            tail = bytes(rng.randrange(256) for _ in range(kb - len(head)))
            pool.append(head + tail if kb > 8 else struct.pack("<I", i * 2654435761 % 2**32))
        pool = list(dict.fromkeys(pool))
        for _ in range(6000):
            k = rng.choice(pool)
            op = rng.choice([OP_GET, OP_PUT, OP_PUT, OP_REMOVE])
            v = bytes(rng.randrange(256) for _ in range(vb))
            f, old, ok = r.op(op, k, v)
            assert ok
            had = k in model
            if op == OP_GET:
                assert f == had and (not vb or old == model.get(k, bytes(vb)))
            elif op == OP_PUT:
                assert f == (had if vb else True)
                if vb:
                    assert old == model.get(k, bytes(vb))
                model[k] = v
            else:
                assert f == had and (not vb or old == model.get(k, bytes(vb)))
                model.pop(k, None)
        assert r.dump() == ({k: v for k, v in model.items()} if vb else
                            {k: b"" for k in model})
        assert r.present() == len(model)


def test_slab_closed_forms():
    """m keys in one bucket take ceil((m - 30) / 30) pool slabs; a hole in a
    non-last slab is not reused, one in the last slab is; inserts past the
    pool fail and count as used."""
    for m in (1, 30, 31, 60, 61, 95):
        r = WideReplay(16, 0, 0, 100)
        keys = [struct.pack("<QQ", 0, i) for i in range(m)]
        r.apply([OP_PUT] * m, keys)
        assert r.slabs_used() == max(0, -(-(m - SLOTS) // SLOTS))
        assert r.max_chain() == 1 + r.slabs_used() and r.present() == m
    r = WideReplay(16, 8, 0, 100)
    keys = [struct.pack("<QQ", 0, i) for i in range(45)]
    r.apply([OP_PUT] * 45, keys, [bytes(8)] * 45)
    r.apply([OP_REMOVE], [keys[3]])              # hole in the head slab
    r.apply([OP_PUT], [b"\x00" * 8 + b"new1" * 2], [bytes(8)])
    assert r.slabs_used() == 1 and r.valid[0] & (1 << 3) == 0  # went to the last slab
    r.apply([OP_REMOVE], [keys[40]])             # hole in the last slab
    r.apply([OP_PUT], [b"\x00" * 8 + b"new2" * 2], [bytes(8)])
    assert r.keys[1 * SLOTS + 40 - SLOTS] == b"\x00" * 8 + b"new2" * 2
    r = WideReplay(4, 0, 0, 1)
    f, _, fail = r.apply([OP_PUT] * 65, [struct.pack("<I", i) for i in range(65)])
    assert fail == 5 and r.present() == 60 and r.slabs_used() == 1


def test_wide_hash_words_and_order():
    """The tag hash depends on every word and on word order (a hash that
    dropped the lane index or a word would collide here)."""
    a = struct.pack("<4I", 1, 2, 3, 4)
    b = struct.pack("<4I", 2, 1, 3, 4)
    c = struct.pack("<4I", 1, 2, 3, 5)
    assert len({wide_hash(a), wide_hash(b), wide_hash(c)}) == 3
    assert wide_hash(bytes(128)) != wide_hash(bytes(124))
