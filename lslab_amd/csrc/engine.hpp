// Per-op device logic shared by the plain kernels, the two-phase kernels and
// the persistent batched kernel.
//
// Plain engine (the reference protocol, wave64-native): each op takes its
// bucket's reader/writer lock for the whole traversal -- shared for gets,
// exclusive for put/remove (detail/traverse.h:24-25,57-61,136-143) -- with one
// lane of the op's G-lane group doing the atomic and the group broadcasting
// the outcome.  Lock attempts sit inside the retry loop together with the
// critical section, so a group that holds a lock always finishes and releases
// it in the same iteration and groups of one wave never wait on each other.
//
// Two-phase engine: phase R looks every op up against the table as it was
// before the batch, lock-free, and orders the batch's mutations of each key
// by one atomic exchange on that key's entry in a batch-local combining table:
// the exchange returns the previous mutation of the same key, which fixes the
// op's result (a put leaves present(v), a remove leaves absent, so a
// mutation's result depends only on its predecessor).  Phase W lets the last
// mutation of each key write the key's final state back once.  Per key the
// results are those of the serial order [all gets][mutations in exchange
// order]; batches without same-key conflicts get exactly the sequential
// replay's results.
#pragma once
#include "slab.hpp"

namespace lslab_dev {

// Loads op i of a stream (opcode, key, value).
template <class L>
__device__ __forceinline__ void load_op(const OpsDev& o, uint64_t i, uint8_t& op,
                                        typename L::K& key, typename L::V& val) {
    using K = typename L::K;
    using V = typename L::V;
    op = o.ops ? o.ops[i] : o.uop;
    key = ld_stride<K>(o.keys, o.kstride, i);
    val = (L::IS_MAP && o.vals && op == OP_PUT) ? ld_stride<V>(o.vals, o.vstride, i) : V(0);
}

// Applies one op to the chain of bucket b.  Must be called by all G lanes of
// the group; when LOCKED the caller holds the bucket lock (shared for gets,
// exclusive otherwise).  Allocation of new slabs is aggregated over all groups
// of the wave that reach this point together: one atomicAdd per wave.
// Returns (found, value) with the reference's meaning per op:
//   get    -> (present, value)      put -> (was present, old value)
//   remove -> (was present, value)  set insert -> (true, 0)
// and ok=false if the pool was exhausted.
template <class L, int G, bool COH>
__device__ __forceinline__ void apply_op(const TableDev& t, uint8_t op, typename L::K key,
                                         typename L::V val, uint64_t b, bool& rfound,
                                         typename L::V& rval, bool& ok) {
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    uint64_t s;
    int slot;
    uint32_t valid;
    walk<L, G, COH>(t, b, key, s, slot, valid);
    const bool found = slot >= 0;
    V cur = V(0);
    if constexpr (L::IS_MAP) {
        if (found) cur = COH ? ld_coh(val_slot<L>(t, s, slot)) : *val_slot<L>(t, s, slot);
    }
    ok = true;
    rfound = found;
    rval = cur;
    bool ins = false, need_alloc = false;
    int ins_slot = -1;
    uint64_t ins_s = s;
    if (op == OP_PUT) {
        if (found) {
            if constexpr (L::IS_MAP) {
                if (gl == 0) st_coh(val_slot<L>(t, s, slot), val);
            }
        } else {
            ins = true;
            uint32_t empty = ~valid & L::CAPMASK;
            if (empty) ins_slot = __builtin_ctz(empty);
            else need_alloc = true;
        }
        if constexpr (!L::IS_MAP) rfound = true;
    } else if (op == OP_REMOVE) {
        if (found && gl == 0)
            __hip_atomic_fetch_and(valid_word<L>(t, s), ~(1u << slot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
    // Wave-aggregated bump allocation: one atomic for every group that needs
    // a fresh slab in this step.
    const uint64_t am = __ballot(need_alloc && gl == 0);
    if (am) {
        const int first = __builtin_ctzll(am);
        uint64_t basei = 0;
        if ((int)__lane_id() == first)
            basei = __hip_atomic_fetch_add(t.alloc_ctr, (uint64_t)__builtin_popcountll(am),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        basei = grp_bcast(basei, first, 64);
        if (need_alloc) {
            const int leader = Grp<G>::base();
            const uint64_t idx = basei + __builtin_popcountll(am & ((1ull << leader) - 1ull));
            if (idx < t.pool) {
                ins_s = t.nbuckets + idx;
                ins_slot = 0;
            } else {
                ins = false;
                ok = false;
                rfound = false;
                rval = V(0);
                if (gl == 0) __hip_atomic_fetch_or(t.status, ST_POOL_EXHAUSTED, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if (ins && gl == 0) {
        // Key and value first, then the valid bit, then the link: each step
        // completes before the next is issued.
        st_coh(key_slot<L>(t, ins_s, ins_slot), key);
        if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, ins_s, ins_slot), val);
        wait_vm();
        __hip_atomic_fetch_or(valid_word<L>(t, ins_s), 1u << ins_slot, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        if (ins_s != s) {
            wait_vm();
            st_coh(next_word<L>(t, s), (uint32_t)ins_s);
        }
    }
}

// Runs one op under the bucket lock (or lock-free when !LOCKED) with the
// try-lock retry loop.  All G lanes call it with the same arguments.
template <class L, int G, bool LOCKED>
__device__ __forceinline__ void run_op(const TableDev& t, uint8_t op, typename L::K key,
                                       typename L::V val, bool& rfound, typename L::V& rval,
                                       bool& ok) {
    const uint64_t b = (uint64_t)key & t.mask;  // identity hash, masked (hash.h:10-15)
    if constexpr (!LOCKED) {
        apply_op<L, G, false>(t, op, key, val, b, rfound, rval, ok);
        return;
    } else {
        const int gl = Grp<G>::lane();
        const bool shared = (op == OP_GET);
        uint32_t* lw = t.locks + b;
        bool done = false;
        uint64_t t0 = 0;
        uint32_t backoff = 1;
        // The loop exits only when every lane of the wave is done (a
        // wave-uniform branch).  With a divergent exit the compiler may run the
        // "got the lock" block -- critical section and unlock -- on the loop's
        // exit path, i.e. after all lanes have left the loop, so a group of the
        // same wave spinning on that lock would never see it released.
        for (;;) {
            if (!done) {
                int got = 0;
                if (gl == 0) got = shared ? try_lock_shared(lw) : try_lock_excl(lw);
                got = __shfl(got, 0, G);
                if (got) {
                    // Inside the lock every slab access is agent-coherent, so
                    // the lock needs no cache-wide fence: the critical section's
                    // reads and writes complete (wait_vm) before the word drops.
                    apply_op<L, G, true>(t, op, key, val, b, rfound, rval, ok);
                    wait_vm();
                    if (gl == 0) {
                        if (shared) unlock_shared(lw);
                        else unlock_excl(lw);
                    }
                    done = true;
                } else {
                    const uint64_t now = now_ticks();
                    if (t0 == 0) t0 = now;
                    if (now - t0 > SPIN_LIMIT_TICKS) {
                        if (gl == 0) __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
                        rfound = false;
                        rval = typename L::V(0);
                        ok = false;
                        done = true;
                    }
                }
            }
            if (__ballot(!done) == 0) break;
            // exponential backoff of the whole wave while a group still waits
            for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(4);
            backoff = backoff < 64 ? backoff * 2 : 64;
        }
    }
}

// Finds or claims key's entry in the batch combining table (one lane).
__device__ __forceinline__ uint32_t gt_find_or_insert(const TableDev& t, uint64_t key) {
    if (key == GT_EMPTY) return t.gt_mask + 1u;  // dedicated entry for the all-ones key
    uint32_t h = (uint32_t)mix64(key) & t.gt_mask;
    for (uint32_t probe = 0; probe <= t.gt_mask; ++probe) {
        uint64_t cur = ld_agent(&t.gt_key[h]);
        if (cur == key) return h;
        if (cur == GT_EMPTY) {
            uint64_t e = GT_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&t.gt_key[h], &e, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
                return h;
            if (e == key) return h;
        }
        h = (h + 1u) & t.gt_mask;
    }
    return 0xffffffffu;  // unreachable while a sub-batch holds <= capacity/2 ops
}

// Phase R of op (lo + li) of a sub-batch of m ops, given where its lock-free
// walk of the pre-batch table ended (slab s, slot or -1): result of gets,
// and the per-key exchange for mutations.
// The exchange order is the order in which waves reach the atomic, not op
// index order, so a predecessor may have any index of the sub-batch.
template <class L, int G>
__device__ __forceinline__ void phase_r_op(const TableDev& t, const OpsDev& o, const OutDev& out,
                                           uint64_t lo, uint32_t li, uint64_t m, uint8_t op,
                                           typename L::K key, typename L::V val, uint64_t s,
                                           int slot) {
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    const uint64_t i = lo + li;
    const bool found = slot >= 0;
    V cur = V(0);
    if constexpr (L::IS_MAP) {
        if (found) cur = *val_slot<L>(t, s, slot);
    }
    if (op == OP_GET) {
        if (gl == 0) write_result<L>(out, i, found, cur);
        return;
    }
    // Wave-level combining: the mutation leaders of this wave that share a
    // key are chained in lane order, and only the last of them touches the
    // key's combining entry (one lookup, one exchange per key per wave instead
    // of per op -- a Zipf-hot key's exchanges are what serialise).
    const int lane = (int)__lane_id();
    const bool lead = (gl == 0);
    const uint64_t kk = (uint64_t)key;
    int my_last = lane, my_prev = -1;
    uint64_t todo = __ballot(lead);
    while (todo) {  // wave-uniform: one pass per distinct key among the leaders
        const int l = __builtin_ctzll(todo);
        const uint64_t kl = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(kk >> 32), l) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((uint32_t)kk, l);
        const bool mine = lead && kk == kl;
        const uint64_t same = __ballot(mine);
        todo &= ~same;
        if (mine) {
            my_last = 63 - __builtin_clzll(same);
            const uint64_t below = same & ((1ull << lane) - 1ull);
            my_prev = below ? 63 - __builtin_clzll(below) : -1;
        }
    }
    uint32_t g = 0, pred = 0;
    if (lead && lane == my_last) {
        g = gt_find_or_insert(t, kk);
        if (g != 0xffffffffu) {
            pred = __hip_atomic_exchange(&t.gt_head[g], li + 1u, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    g = __shfl(g, my_last, 64);
    pred = __shfl(pred, my_last, 64);
    const uint32_t prev_li = __shfl(li, my_prev >= 0 ? my_prev : lane, 64);
    if (my_prev >= 0) pred = prev_li + 1u;
    if (lead) {
        bool rf = found;
        V rv = cur;
        if (pred > m) {  // entries only ever hold ops of this sub-batch (any order)
            __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            pred = 0;
        }
        if (pred != 0) {
            const uint64_t pi = lo + pred - 1u;
            const uint8_t pop = o.ops ? o.ops[pi] : o.uop;
            if (pop == OP_PUT) {
                rf = true;
                rv = L::IS_MAP ? ld_stride<V>(o.vals, o.vstride, pi) : V(0);
            } else {
                rf = false;
                rv = V(0);
            }
        }
        if (!L::IS_MAP && op == OP_PUT) rf = true;  // set insert reports true
        write_result<L>(out, i, rf, rv);
        t.op_gidx[li] = g;
        t.op_loc[li] = found ? ((s << 5) | (uint64_t)slot) : NOLOC;
    }
}

// Phase R over ops [lo, lo + m) of a sub-batch, group gid of ng, as a
// three-stage software pipeline per group: while op i finishes (its value
// load, result, exchange), the head-slab keys of op i+1 and the record
// (opcode, key, value) of op i+2 are already in flight, so the three
// dependent round trips of an op overlap with those of its neighbours.
// Chains longer than the head slab continue synchronously (walk_on).
template <class L, int G>
__device__ __forceinline__ void phase_r_range(const TableDev& t, const OpsDev& o,
                                              const OutDev& out, uint64_t lo, uint64_t m,
                                              uint64_t gid, uint64_t ng) {
    using K = typename L::K;
    using V = typename L::V;
    constexpr int S = Grp<G>::S;
    uint64_t li = gid;
    if (li >= m) return;
    uint8_t op;
    K key;
    V val;
    load_op<L>(o, lo + li, op, key, val);
    K blk[S];
    probe_issue<L, G, false>(t, (uint64_t)key & t.mask, blk);
    uint64_t n1 = li + ng;
    uint8_t op1 = 0;
    K key1 = K(0);
    V val1 = V(0);
    if (n1 < m) load_op<L>(o, lo + n1, op1, key1, val1);
    for (;;) {
        const uint64_t n2 = n1 + ng;
        uint8_t op2 = 0;
        K key2 = K(0);
        V val2 = V(0);
        if (n2 < m) load_op<L>(o, lo + n2, op2, key2, val2);
        const uint64_t b = (uint64_t)key & t.mask;
        uint64_t s = b;
        int slot;
        uint32_t valid;
        walk_on<L, G, false>(t, probe_compute<L, G, false>(t, b, key, blk), s, slot, valid, key);
        K blk1[S];
#pragma unroll
        for (int e = 0; e < S; ++e) blk1[e] = K(0);
        if (n1 < m) probe_issue<L, G, false>(t, (uint64_t)key1 & t.mask, blk1);
        phase_r_op<L, G>(t, o, out, lo, (uint32_t)li, m, op, key, val, s, slot);
        if (n1 >= m) break;
        li = n1;
        op = op1;
        key = key1;
        val = val1;
#pragma unroll
        for (int e = 0; e < S; ++e) blk[e] = blk1[e];
        n1 = n2;
        op1 = op2;
        key1 = key2;
        val1 = val2;
    }
}

// Phase W of op (lo + li): the last mutation of each key writes its final
// state and frees the combining-table entry.
template <class L, int G>
__device__ __forceinline__ void phase_w_op(const TableDev& t, const OpsDev& o, uint64_t lo,
                                           uint32_t li) {
    using K = typename L::K;
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    const uint64_t i = lo + li;
    const uint8_t op = o.ops ? o.ops[i] : o.uop;
    if (op == OP_GET) return;
    const uint32_t g = t.op_gidx[li];
    if (g == 0xffffffffu) return;
    if (ld_agent(&t.gt_head[g]) != li + 1u) return;  // a later mutation of this key exists
    const uint64_t loc = t.op_loc[li];
    K key;
    V val;
    uint8_t op2;
    load_op<L>(o, i, op2, key, val);
    if (gl == 0) {
        __hip_atomic_store(&t.gt_head[g], 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (g <= t.gt_mask)
            __hip_atomic_store(&t.gt_key[g], GT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (loc != NOLOC) {
        const uint64_t s = loc >> 5;
        const int slot = (int)(loc & 31u);
        if (gl == 0) {
            if (op == OP_PUT) {
                if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, s, slot), val);
            } else {
                __hip_atomic_fetch_and(valid_word<L>(t, s), ~(1u << slot), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    } else if (op == OP_PUT) {
        bool rf, ok;
        V rv;
        run_op<L, G, true>(t, OP_PUT, key, val, rf, rv, ok);
    }
}

}  // namespace lslab_dev
