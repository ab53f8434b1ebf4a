"""Oracle for the wide-key tables (test infrastructure only): a sequential
pure-Python replay, for the small cases the tests run.

Only tests/ import this module, as the checker; the product never does.

The table follows the same slab policy the narrow oracle restates from the
reference (lslab_oracle.c; detail/traverse.h:54-55,69-124,
detail/device_allocator.h:62-73): each bucket heads a chain of slabs of
SLOTS (30) keys with a valid bitmap; a lookup scans the valid slots of each
slab in chain order; an insert of an absent key takes the lowest empty slot
of the chain's LAST slab, or a fresh slab from a bump pool linked at the end
(slot 0); a remove clears the valid bit; pool slabs are never reused.  Keys
are byte strings compared in full; a slab also keeps each slot's tag, the
low 32 bits of `wide_hash`, which decides nothing here (the table matches on
the full key) but whose definition the device shares for its buckets under
LSLAB_HASH_MIX.  The reference's wide-key tests are the legacy
test/slabunified_test.cu.old:386-549 (PutRemoveTest_128B,
PutRemoveTest_128Bto128B), restated in tests/test_wide_oracle.py.

Parity pins: those legacy known-answer tests, a brute-force dict model over
random streams, and the closed form of pool slabs for m keys in one bucket
(tests/test_wide_oracle.py).
"""
import struct

OP_GET, OP_PUT, OP_REMOVE = 0, 1, 2
SLOTS = 30
SEED = 0x9E3779B97F4A7C15
M64 = (1 << 64) - 1


def mix64(z):
    """murmur3 64-bit finalizer."""
    z &= M64
    z ^= z >> 33
    z = (z * 0xFF51AFD7ED558CCD) & M64
    z ^= z >> 33
    z = (z * 0xC4CEB9FE1A85EC53) & M64
    z ^= z >> 33
    return z


def wide_hash(key):
    """mix64 of the XOR over 32-bit words l of mix64((l << 32 | word_l) ^ SEED)."""
    words = struct.unpack(f"<{len(key) // 4}I", key)
    m = 0
    for l, w in enumerate(words):
        m ^= mix64(((l << 32) | w) ^ SEED)
    return mix64(m)


class WideReplay:
    def __init__(self, key_bytes, value_bytes, n_log2, pool_slabs, mix_hash=False):
        assert key_bytes % 4 == 0 and 4 <= key_bytes <= 128
        assert value_bytes % 4 == 0 and 0 <= value_bytes <= 128
        self.kb, self.vb = key_bytes, value_bytes
        self.is_map = value_bytes > 0
        self.nb = 1 << n_log2
        self.pool = pool_slabs
        self.mix = mix_hash
        self.used = 0
        nslabs = self.nb + pool_slabs
        self.keys = [None] * (nslabs * SLOTS)
        self.vals = [None] * (nslabs * SLOTS)
        self.valid = [0] * nslabs
        self.next = [0] * nslabs  # 0 = end (slab 0 is a head, never a successor)

    def bucket(self, key):
        if self.mix:
            return (wide_hash(key) >> 32) & (self.nb - 1)
        first = key[:8] if len(key) >= 8 else key + bytes(8 - len(key))
        return struct.unpack("<Q", first)[0] & (self.nb - 1)

    def _find(self, key):
        s = self.bucket(key)
        while True:
            v = self.valid[s]
            for j in range(SLOTS):
                if (v >> j) & 1 and self.keys[s * SLOTS + j] == key:
                    return s, j, s
            if self.next[s] == 0:
                return None, None, s
            s = self.next[s]

    def op(self, op, key, val=None):
        """One op; returns (found, old value bytes or zeros, ok)."""
        zero = bytes(self.vb)
        if op not in (OP_PUT, OP_REMOVE):
            op = OP_GET
        s, j, last = self._find(key)
        found = s is not None
        old = self.vals[s * SLOTS + j] if (found and self.is_map) else zero
        if op == OP_PUT:
            if found:
                if self.is_map:
                    self.vals[s * SLOTS + j] = val
            else:
                empty = ~self.valid[last] & ((1 << SLOTS) - 1)
                if empty:
                    ts, tj = last, (empty & -empty).bit_length() - 1
                else:
                    if self.used >= self.pool:
                        self.used += 1
                        return False, zero, False
                    ts, tj = self.nb + self.used, 0
                    self.used += 1
                    self.next[last] = ts
                self.keys[ts * SLOTS + tj] = key
                self.vals[ts * SLOTS + tj] = val if self.is_map else None
                self.valid[ts] |= 1 << tj
            if not self.is_map:
                found = True  # a set insert reports true for present and new keys alike
        elif op == OP_REMOVE and found:
            self.valid[s] &= ~(1 << j)
        return found, old, True

    def apply(self, ops, keys, vals=None):
        """Replays a batch in order: (found list, old-value list, failed inserts)."""
        f, o, fail = [], [], 0
        for i, (op, k) in enumerate(zip(ops, keys)):
            fd, old, ok = self.op(int(op), k, vals[i] if vals is not None else None)
            f.append(1 if fd else 0)
            o.append(old)
            fail += 0 if ok else 1
        return f, o, fail

    def slabs_used(self):
        return min(self.used, self.pool)

    def present(self):
        return sum(bin(v).count("1") for v in self.valid)

    def max_chain(self):
        best = 0
        for b in range(self.nb):
            n, s = 1, b
            while self.next[s]:
                s = self.next[s]
                n += 1
            best = max(best, n)
        return best

    def dump(self):
        """{key: value bytes} of every present key."""
        out = {}
        for s, v in enumerate(self.valid):
            for j in range(SLOTS):
                if (v >> j) & 1:
                    out[self.keys[s * SLOTS + j]] = self.vals[s * SLOTS + j] if self.is_map else b""
        return out
