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

// Opcode of op i.  A value outside GET/PUT/REMOVE reads as GET, so no caller
// opcode reaches the internal ones.
__device__ __forceinline__ uint8_t ld_opcode(const OpsDev& o, uint64_t i) {
    const uint8_t raw = o.ops ? o.ops[i] : o.uop;
    return raw <= OP_REMOVE ? raw : (uint8_t)OP_GET;
}

// Loads op i of a stream (opcode, key, value).
template <class L>
__device__ __forceinline__ void load_op(const OpsDev& o, uint64_t i, uint8_t& op,
                                        typename L::K& key, typename L::V& val) {
    using K = typename L::K;
    using V = typename L::V;
    op = ld_opcode(o, i);
    key = ld_stride<K>(o.keys, o.kstride, i);
    val = (L::IS_MAP && o.vals && op == OP_PUT) ? ld_stride<V>(o.vals, o.vstride, i) : V(0);
}

// Free list of pool slabs (reclaim policy): a LIFO stack linked through the
// slabs' next words, its head word tagged (tag << 32 | slab) so that a pop
// racing with a pop-and-push of the same slab fails its compare-and-swap.
// One lane each; spins are bounded (a lost race retries, a damaged list
// turns into a status bit).
template <class L>
__device__ __forceinline__ void free_push(const TableDev& t, uint64_t s) {
    for (uint32_t tries = 0; tries < (1u << 20); ++tries) {
        uint64_t old = ld_agent(t.free_head);
        st_coh(next_word<L>(t, s), (uint32_t)old);
        wait_vm();
        const uint64_t nw = (((old >> 32) + 1ull) << 32) | s;
        if (__hip_atomic_compare_exchange_strong(t.free_head, &old, nw, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_fetch_add(t.free_count, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
    }
    __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Pops a free pool slab, or returns 0 (none free).
template <class L>
__device__ __forceinline__ uint64_t free_pop(const TableDev& t) {
    for (uint32_t tries = 0; tries < (1u << 20); ++tries) {
        uint64_t old = ld_agent(t.free_head);
        const uint64_t s = (uint32_t)old;
        if (s == 0) return 0;
        if (s < t.nbuckets || s >= t.nbuckets + t.pool) {
            __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return 0;
        }
        const uint64_t nx = ld_coh(next_word<L>(t, s));
        const uint64_t nw = (((old >> 32) + 1ull) << 32) | nx;
        if (__hip_atomic_compare_exchange_strong(t.free_head, &old, nw, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
            __hip_atomic_fetch_add(t.free_count, ~0ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return s;
        }
    }
    __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return 0;
}

// Unlinks every empty pool slab of bucket b's chain and frees it (reclaim
// policy).  The caller holds b's exclusive lock; all G lanes call it.
template <class L, int G>
__device__ __forceinline__ void compact_chain(const TableDev& t, uint64_t b) {
    const int gl = Grp<G>::lane();
    uint64_t pv = b;
    uint64_t nx = ld_coh(next_word<L>(t, b));
    for (uint64_t hops = 0; nx != 0; ++hops) {
        if (nx < t.nbuckets || nx >= t.nbuckets + t.pool || hops > t.pool) {
            if (gl == 0)
                __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        const uint32_t v = ld_coh(valid_word<L>(t, nx)) & L::CAPMASK;
        const uint64_t nn = ld_coh(next_word<L>(t, nx));
        if (v == 0) {
            if (gl == 0) {
                st_coh(next_word<L>(t, pv), (uint32_t)nn);
                wait_vm();
                free_push<L>(t, nx);
            }
        } else {
            pv = nx;
        }
        nx = nn;
    }
}

// apply_op for a mutation on a table with the reclaim policy, under the
// bucket's exclusive lock: an insert takes the first slab of the chain with
// an empty slot, a freed slab before a fresh one; a remove that leaves a pool
// slab empty unlinks and frees it; OP_COMPACT unlinks every empty pool slab.
// Instantiated only for reclaim tables (RC), so the reference-policy kernels
// carry none of it.
template <class L, int G>
__device__ __forceinline__ void apply_op_reclaim(const TableDev& t, uint8_t op, typename L::K key,
                                                 typename L::V val, uint64_t b, bool& rfound,
                                                 typename L::V& rval, bool& ok) {
    constexpr bool COH = true;
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    ok = true;
    if (op == OP_COMPACT) {
        compact_chain<L, G>(t, b);
        rfound = false;
        rval = V(0);
        return;
    }
    uint64_t s, prev = NOSLAB, room = NOSLAB;
    int slot;
    uint32_t valid, room_valid = 0;
    constexpr bool rc = true;
    walk_full<L, G, COH>(t, b, key, s, slot, valid, prev, room, room_valid);
    const bool found = slot >= 0;
    V cur = V(0);
    if constexpr (L::IS_MAP) {
        if (found) cur = COH ? ld_coh(val_slot<L>(t, s, slot)) : *val_slot<L>(t, s, slot);
    }
    rfound = found;
    rval = cur;
    bool ins = false, need_alloc = false, fresh = false;
    int ins_slot = -1;
    uint64_t ins_s = s;
    if (op == OP_PUT) {
        if (found) {
            if constexpr (L::IS_MAP) {
                if (gl == 0) st_coh(val_slot<L>(t, s, slot), val);
            }
        } else {
            ins = true;
            uint32_t tv = valid;
            if (rc && room != NOSLAB) {
                ins_s = room;
                tv = room_valid;
            }
            const uint32_t empty = ~tv & L::CAPMASK;
            if (empty) ins_slot = __builtin_ctz(empty);
            else need_alloc = true;
        }
        if constexpr (!L::IS_MAP) rfound = true;
    } else if (op == OP_REMOVE) {
        if (found) {
            uint32_t old = 0;
            if (gl == 0)
                old = __hip_atomic_fetch_and(valid_word<L>(t, s), ~(1u << slot), __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
            if (rc) {
                old = grp_bcast(old, 0, G);
                // Under the lock: a pool slab left empty leaves its chain.
                if (s >= t.nbuckets && prev != NOSLAB &&
                    ((old & ~(1u << slot)) & L::CAPMASK) == 0 && gl == 0) {
                    st_coh(next_word<L>(t, prev), ld_coh(next_word<L>(t, s)));
                    wait_vm();
                    free_push<L>(t, s);
                }
            }
        }
    }
    // Reclaim: a freed slab before a fresh one (one pop per group).
    bool need_bump = need_alloc;
    if (rc && need_alloc) {
        uint64_t got = 0;
        if (gl == 0) got = free_pop<L>(t);
        got = grp_bcast(got, 0, G);
        if (got) {
            ins_s = got;
            ins_slot = L::SLOT0;
            fresh = true;
            need_bump = false;
        }
    }
    // Wave-aggregated bump allocation: one atomic for every group that needs
    // a fresh slab in this step.
    const uint64_t am = __ballot(need_bump && gl == 0);
    if (am) {
        const int first = __builtin_ctzll(am);
        uint64_t basei = 0;
        if ((int)__lane_id() == first)
            basei = __hip_atomic_fetch_add(t.alloc_ctr, (uint64_t)__builtin_popcountll(am),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        basei = grp_bcast(basei, first, 64);
        if (need_bump) {
            const int leader = Grp<G>::base();
            const uint64_t idx = basei + __builtin_popcountll(am & ((1ull << leader) - 1ull));
            if (idx < t.pool) {
                ins_s = t.nbuckets + idx;
                ins_slot = L::SLOT0;
                fresh = true;
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
        // A reused slab still carries its free-list link: clear it first.
        // Key and value, then the valid bit, then the link: each step
        // completes before the next is issued.
        if (fresh) st_coh(next_word<L>(t, ins_s), 0u);
        st_coh(key_slot<L>(t, ins_s, ins_slot), key);
        if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, ins_s, ins_slot), val);
        wait_vm();
        __hip_atomic_fetch_or(valid_word<L>(t, ins_s), 1u << ins_slot, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        if (fresh) {
            wait_vm();
            st_coh(next_word<L>(t, s), (uint32_t)ins_s);
        }
    }
}


// Applies one op to bucket b's chain.  With COH the caller holds the bucket
// lock (shared for gets, exclusive otherwise) and every slab access is
// agent-coherent; without, loads are plain (read-only launches).  All G
// lanes of the group call it; lane 0 performs the writes.
// Allocation of new slabs is aggregated over all groups of the wave that
// reach this point together: one atomicAdd per wave.
// Returns (found, value) with the reference's meaning per op:
//   get    -> (present, value)      put -> (was present, old value)
//   remove -> (was present, value)  set insert -> (true, 0)
// and ok=false if the pool was exhausted.
// Placement is the reference's policy (the last slab's lowest empty slot,
// else a new slab linked at the end; traverse.h:54-55,98-124); tables with
// the reclaim policy (RC) hand mutations to apply_op_reclaim.
template <class L, int G, bool COH, bool RC>
__device__ __forceinline__ void apply_op(const TableDev& t, uint8_t op, typename L::K key,
                                         typename L::V val, uint64_t b, bool& rfound,
                                         typename L::V& rval, bool& ok) {
    using V = typename L::V;
    if constexpr (COH && RC) {
        if (op != OP_GET) {
            apply_op_reclaim<L, G>(t, op, key, val, b, rfound, rval, ok);
            return;
        }
    }
    const int gl = Grp<G>::lane();
    ok = true;
    uint64_t s;
    int slot;
    uint32_t valid;
    walk<L, G, COH>(t, b, key, s, slot, valid);
    const bool found = slot >= 0;
    V cur = V(0);
    if constexpr (L::IS_MAP) {
        if (found) cur = COH ? ld_coh(val_slot<L>(t, s, slot)) : *val_slot<L>(t, s, slot);
    }
    rfound = found;
    rval = cur;
    bool ins = false, need_alloc = false, fresh = false;
    int ins_slot = -1;
    uint64_t ins_s = s;
    if (op == OP_PUT) {
        if (found) {
            if constexpr (L::IS_MAP) {
                if (gl == 0) st_coh(val_slot<L>(t, s, slot), val);
            }
        } else {
            ins = true;
            const uint32_t empty = ~valid & L::CAPMASK;
            if (empty) ins_slot = __builtin_ctz(empty);
            else need_alloc = true;
        }
        if constexpr (!L::IS_MAP) rfound = true;
    } else if (op == OP_REMOVE) {
        if (found && gl == 0) {
            if constexpr (COH)
                st_coh(valid_word<L>(t, s), valid & ~(1u << slot));
            else
                __hip_atomic_fetch_and(valid_word<L>(t, s), ~(1u << slot), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    const bool need_bump = need_alloc;
    // Wave-aggregated bump allocation: one atomic for every group that needs
    // a fresh slab in this step.
    const uint64_t am = __ballot(need_bump && gl == 0);
    if (am) {
        const int first = __builtin_ctzll(am);
        uint64_t basei = 0;
        if ((int)__lane_id() == first)
            basei = __hip_atomic_fetch_add(t.alloc_ctr, (uint64_t)__builtin_popcountll(am),
                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        basei = grp_bcast(basei, first, 64);
        if (need_bump) {
            const int leader = Grp<G>::base();
            const uint64_t idx = basei + __builtin_popcountll(am & ((1ull << leader) - 1ull));
            if (idx < t.pool) {
                ins_s = t.nbuckets + idx;
                ins_slot = L::SLOT0;
                fresh = true;
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
        st_coh(key_slot<L>(t, ins_s, ins_slot), key);
        if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, ins_s, ins_slot), val);
        if constexpr (COH) {
            // Under the bucket's exclusive lock: every other op on this chain
            // (gets included) waits for the unlock, before which the caller
            // completes these stores, and only the lock holder writes the
            // chain's valid words.  So the new bitmap is a plain store of the
            // one read in the walk, and the stores need no order among them
            // (an atomic OR issued right behind the key store to the same
            // line costs more than it saves).
            st_coh(valid_word<L>(t, ins_s), (fresh ? 0u : valid) | (1u << ins_slot));
            if (fresh) st_coh(next_word<L>(t, s), (uint32_t)ins_s);
        } else {
            // Key and value, then the valid bit, then the link: each step
            // completes before the next is issued.
            wait_vm();
            __hip_atomic_fetch_or(valid_word<L>(t, ins_s), 1u << ins_slot, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
            if (fresh) {
                wait_vm();
                st_coh(next_word<L>(t, s), (uint32_t)ins_s);
            }
        }
    }
}

// Runs one op under the bucket lock (or lock-free when !LOCKED) with the
// try-lock retry loop.  All G lanes call it with the same arguments.
template <class L, int G, bool LOCKED, bool RC>
__device__ __forceinline__ void run_op(const TableDev& t, uint8_t op, typename L::K key,
                                       typename L::V val, bool& rfound, typename L::V& rval,
                                       bool& ok) {
    const uint64_t b = bucket_of(t, (uint64_t)key);
    if constexpr (!LOCKED) {
        apply_op<L, G, false, RC>(t, op, key, val, b, rfound, rval, ok);
        return;
    } else {
        const int gl = Grp<G>::lane();
        const bool shared = (op == OP_GET);
        uint32_t* lw = t.locks + b;
        bool done = false, first = true;
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
                if (gl == 0)
                    got = shared ? try_lock_shared(lw) : try_lock_excl(lw, first);
                first = false;
                got = __shfl(got, 0, G);
                if (got) {
                    // Inside the lock every slab access is agent-coherent, so
                    // the lock needs no cache-wide fence: the critical section's
                    // reads and writes complete (wait_vm) before the word drops.
                    apply_op<L, G, true, RC>(t, op, key, val, b, rfound, rval, ok);
                    wait_vm();
                    if (gl == 0) {
                        if (shared) unlock_shared(lw);
                        else unlock_excl(lw);
                    }
                    done = true;
                } else {
                    const uint64_t now = now_ticks();
                    if (t0 == 0) t0 = now;
                    if (now - t0 > t.spin_ticks) {
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

// gt_find_or_insert with one compare-and-swap per probe (an empty slot to the
// key), which claims the slot or returns what it holds -- the key itself or
// another key -- in one round trip.  Used with wavefront combining (+2% on
// the headline); block combining keeps the load-first probe, which keeps
// read-modify-writes off the slot of a key that many chains exchange on.
__device__ __forceinline__ uint32_t gt_claim_cas(const TableDev& t, uint64_t key) {
    if (key == GT_EMPTY) return t.gt_mask + 1u;  // dedicated entry for the all-ones key
    uint32_t h = (uint32_t)mix64(key) & t.gt_mask;
    for (uint32_t probe = 0; probe <= t.gt_mask; ++probe) {
        uint64_t e = GT_EMPTY;
        if (__hip_atomic_compare_exchange_strong(&t.gt_key[h], &e, key, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
            e == key)
            return h;
        h = (h + 1u) & t.gt_mask;
    }
    return 0xffffffffu;  // unreachable while a sub-batch holds <= capacity/2 ops
}

// Block-level combining table for one phase-R round.  Two tables alternate
// by round parity: the entries a round uses are cleared in that round, and a
// table is reused two rounds (two block barriers) later.
template <int BT>
struct CombTab {
    static constexpr uint32_t N = 2u * BT;  // power of two; a round has at most BT chains
    uint64_t key[N + 1];                    // entry N is reserved for the all-ones key
    uint32_t head[N + 1];                   // last chain so far: its last op index + 1
    uint32_t first[N + 1];                  // first op of the block's first chain
};

template <int BT>
__device__ __forceinline__ CombTab<BT>* comb_tables() {
    __shared__ CombTab<BT> tabs[2];
    return tabs;
}

template <int BT>
__device__ __forceinline__ void comb_init() {
    CombTab<BT>* tb = comb_tables<BT>();
    constexpr uint32_t E = CombTab<BT>::N + 1;
    for (uint32_t x = threadIdx.x; x < 2 * E; x += blockDim.x) {
        tb[x / E].key[x % E] = GT_EMPTY;
        tb[x / E].head[x % E] = 0;
    }
    __syncthreads();
}

// Entry of key in the round's table (inserted if new; linear probing).
template <int BT>
__device__ __forceinline__ uint32_t comb_slot(CombTab<BT>& c, uint64_t key) {
    constexpr uint32_t N = CombTab<BT>::N;
    if (key == GT_EMPTY) return N;
    uint32_t h = (uint32_t)mix64(key) & (N - 1u);
    for (uint32_t p = 0; p < N; ++p) {
        const uint64_t cur = __hip_atomic_load(&c.key[h], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (cur == key) return h;
        if (cur == GT_EMPTY) {
            uint64_t e = GT_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&c.key[h], &e, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP))
                return h;
            if (e == key) return h;
        }
        h = (h + 1u) & (N - 1u);
    }
    return N;  // unreachable: the table is at most half full
}

// The one global step of a chain: find or insert the key's combining-table
// entry and swap the chain's last op (index + 1) into its head.  pred gets
// the previous head: the last op of the chain exchanged before, or 0.
template <bool CAS_FIRST = false>
__device__ __forceinline__ uint32_t gt_exchange(const TableDev& t, uint64_t key, uint32_t v,
                                                uint32_t& pred) {
    const uint32_t g = CAS_FIRST ? gt_claim_cas(t, key) : gt_find_or_insert(t, key);
    if (g != 0xffffffffu) {
        pred = __hip_atomic_exchange(&t.gt_head[g], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        pred = 0;
        __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    return g;
}

// The global step of a chain whose key the probe found (or not) at slab s,
// slot: a key in the table heads its chain at its own slot when the table
// keeps per-slot heads (no claim), any other key at its combining-table
// entry.  Returns the head's id for phase W (LOC_G | slot index, or the
// entry).
template <bool CAS_FIRST = false>
__device__ __forceinline__ uint32_t chain_exchange(const TableDev& t, uint64_t key, bool found,
                                                   uint64_t s, int slot, uint32_t v,
                                                   uint32_t& pred) {
    if (t.loc_head && found) {
        const uint32_t x = (uint32_t)((s << 5) | (uint64_t)slot);
        pred = __hip_atomic_exchange(&t.loc_head[x], v, __ATOMIC_RELAXED,
                                     __HIP_MEMORY_SCOPE_AGENT);
        return LOC_G | x;
    }
    return gt_exchange<CAS_FIRST>(t, key, v, pred);
}

// Writes the result of mutation (lo + li) with opcode op, whose serial
// predecessor in the sub-batch (or run) is op (lo + pred - 1), or, for
// pred == 0, the key's pre-batch state (found, cur).
template <class L>
__device__ __forceinline__ void mut_result(const TableDev& t, const OpsDev& o, const OutDev& out,
                                           uint64_t lo, uint64_t m, uint32_t li, uint8_t op,
                                           uint32_t pred, bool found, typename L::V cur) {
    using V = typename L::V;
    bool rf = found;
    V rv = cur;
    if (pred > m) {  // entries only ever hold ops of this sub-batch (any order)
        __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        pred = 0;
    }
    if (pred != 0) {
        // The predecessor's opcode and value are loaded together (one round
        // trip); the value counts only for a put.
        const uint64_t pi = lo + pred - 1u;
        const uint8_t pop = ld_opcode(o, pi);
        const V pv = (L::IS_MAP && o.vals) ? ld_stride<V>(o.vals, o.vstride, pi) : V(0);
        rf = pop == OP_PUT;
        rv = rf ? pv : V(0);
    }
    if (!L::IS_MAP && op == OP_PUT) rf = true;  // set insert reports true
    write_result<L>(out, lo + li, rf, rv);
}

// Phase R of op (lo + li) of a sub-batch of m ops, given where its lock-free
// walk of the pre-batch table ended (slab s, slot or -1): the result of gets,
// and for mutations their place in the key's serial order.
//
// Mutations of one key form a chain: within a wavefront in lane order, then
// (block combining) the wavefront chains of the block in the order they
// reach the round's LDS table, then the block chains in the order of their
// one global exchange.  Each op's result comes from its predecessor in the
// chain; the first op of the block's first chain learns its predecessor only
// from the global exchange, so the block's last chain, which does that
// exchange, writes it.  The global exchange order is not op index order, so
// a predecessor may have any index of the sub-batch.
//
// Only the last op of each block chain (wave chain without block combining)
// can be the key's final mutation, so only it records its combining-table
// entry and pre-batch location for phase W; the other mutations record none.
//
// With block combining (BC; ct is the round's table) every thread of the
// block calls this once per round (idle groups with act = false), as it
// contains a block barrier.
template <class L, int G, int BT, bool BC>
__device__ __forceinline__ void phase_r_op(const TableDev& t, const OpsDev& o, const OutDev& out,
                                           uint64_t lo, uint32_t li, uint64_t m, bool act,
                                           uint8_t op, typename L::K key, typename L::V val,
                                           uint64_t s, int slot, CombTab<BT>* ct, RepEnt* rb,
                                           uint32_t* rcnt, RepEnt* lrep, uint32_t lcap) {
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    const bool found = act && slot >= 0;
    V cur = V(0);
    if constexpr (L::IS_MAP) {
        if (found) cur = *val_slot<L>(t, s, slot);
    }
    // The value load stays in flight while the wavefront chains its mutations
    // and issues their exchanges; gets write their results after that.
    const bool wget = act && op == OP_GET && gl == 0;
    const int lane = (int)__lane_id();
    const bool lead = act && op != OP_GET && gl == 0;
    const uint64_t kk = (uint64_t)key;
    uint64_t todo = __ballot(lead);
    if (todo == 0) {  // wave-uniform: no mutation in this wavefront
        if (wget) write_result<L>(out, lo + li, found, cur);
        if constexpr (BC) lds_barrier();
        return;
    }
    // Wave chains: one pass per distinct key among the mutation leaders.
    int my_first = lane, my_last = lane, my_prev = -1;
    while (todo) {
        const int l = __builtin_ctzll(todo);
        const uint64_t kl = ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(kk >> 32), l) << 32) |
                            (uint32_t)__builtin_amdgcn_readlane((uint32_t)kk, l);
        const bool mine = lead && kk == kl;
        const uint64_t same = __ballot(mine);
        todo &= ~same;
        if (mine) {
            if constexpr (BC) my_first = __builtin_ctzll(same);
            my_last = 63 - __builtin_clzll(same);
            const uint64_t below = same & ((1ull << lane) - 1ull);
            my_prev = below ? 63 - __builtin_clzll(below) : -1;
        }
    }
    const bool rep = lead && lane == my_last;  // represents its wave chain
    uint32_t g = 0xffffffffu, fpred = 0;
    bool fdefer = false;  // the chain's first result is written by the block's last chain
    if constexpr (BC) {
        const uint32_t first_li = __shfl(li, my_first, 64);
        uint32_t j = 0;
        if (rep) {
            j = comb_slot<BT>(*ct, kk);
            fpred = __hip_atomic_exchange(&ct->head[j], li + 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_WORKGROUP);
            if (fpred == 0) ct->first[j] = first_li;
        }
        lds_barrier();
        if (rep) {
            fdefer = fpred == 0;
            if (ct->head[j] == li + 1u) {  // the block's last chain of this key
                const uint32_t fl = ct->first[j];
                ct->key[j] = GT_EMPTY;
                ct->head[j] = 0;
                uint32_t P;
                g = chain_exchange(t, kk, found, s, slot, li + 1u, P);
                uint8_t fop = OP_PUT;
                if constexpr (!L::IS_MAP) fop = ld_opcode(o, lo + fl);
                mut_result<L>(t, o, out, lo, m, fl, fop, P, found, cur);
            }
        }
    } else if (rep) {
        g = chain_exchange<true>(t, kk, found, s, slot, li + 1u, fpred);
    }
    if (wget) write_result<L>(out, lo + li, found, cur);
    fpred = __shfl(fpred, my_last, 64);
    if constexpr (BC) fdefer = __shfl((int)fdefer, my_last, 64) != 0;
    const uint32_t prev_li = __shfl(li, my_prev >= 0 ? my_prev : lane, 64);
    if (lead && (my_prev >= 0 || !fdefer))
        mut_result<L>(t, o, out, lo, m, li, op, my_prev >= 0 ? prev_li + 1u : fpred, found, cur);
    // The chain's last op may be the key's final mutation: append it to the
    // block's list (one LDS atomic per wavefront) for phase W.
    const bool cand = rep && g != 0xffffffffu;
    const uint64_t cm = __ballot(cand);
    if (cm) {
        const int l0 = __builtin_ctzll(cm);
        uint32_t wb = 0;
        if (lane == l0)
            wb = __hip_atomic_fetch_add(rcnt, (uint32_t)__builtin_popcountll(cm), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
        wb = __builtin_amdgcn_readlane(wb, l0);
        if (cand) {
            RepEnt e;
            e.li = li;
            e.g = g;
            e.loc = found ? ((s << 5) | (uint64_t)slot) : NOLOC;
            const uint32_t k = wb + (uint32_t)__builtin_popcountll(cm & ((1ull << lane) - 1ull));
            if (k < lcap)
                lrep[k] = e;
            else
                rb[k] = e;
        }
    }
}

// Start of block b's final-candidate list in a launch that runs one
// sub-batch of m ops (k_phase_r / k_phase_w): the number of ops that blocks
// 0..b-1 own when group gid (of ng, gpb per block) owns ops gid, gid + ng,
// ...  Round r gives blocks 0..b-1 the ops [r*ng, r*ng + b*gpb) clipped to
// m; blocks own disjoint ranges, so each block's list fits between its start
// and the next block's, and all lists fit in m entries.
__device__ __forceinline__ uint64_t block_rep_base(uint64_t m, uint64_t ng, uint64_t gpb,
                                                   uint64_t b) {
    const uint64_t A = b * gpb;
    const uint64_t R = (m + ng - 1) / ng;
    uint64_t F = 0;  // rounds in which all of blocks 0..b-1 have ops
    if (m >= A) F = min(R, (m - A) / ng + 1);
    return F * A + (R - F) * m - ng * ((R * (R - 1)) / 2 - (F * (F - 1)) / 2);
}

// Start of block b's final-candidate list in a persistent launch: a fixed
// region of (ops a block can own in a sub-batch of submax ops) entries, the
// same for every sub-batch.  The lists are written with plain stores, which
// stay dirty in the writing XCD's L2 (grid barriers write nothing back);
// with regions that moved with the sub-batch size, another XCD could later
// write back a stale line over a newer entry.  Fixed regions mean each
// entry address is only ever written from one CU.  All regions fit in
// submax + ng entries.
__device__ __forceinline__ uint64_t block_rep_base_fixed(uint64_t submax, uint64_t ng,
                                                         uint64_t gpb, uint64_t b) {
    return b * ((submax + ng - 1) / ng) * gpb;
}

// Final candidates a persistent block keeps in LDS; more spill to its
// global region (same index).
constexpr uint32_t REP_LDS = 1024;

__device__ __forceinline__ RepEnt* rep_lds() {
    __shared__ RepEnt ents[REP_LDS];
    return ents;
}

// The block's LDS counter of final candidates.
__device__ __forceinline__ uint32_t* rep_counter() {
    __shared__ uint32_t cnt;
    return &cnt;
}

// Phase R over ops [lo, lo + m) of a sub-batch, group gid of ng, as a
// three-stage software pipeline per group: while op i finishes (its value
// load, result, exchange), the head-slab keys of op i+1 and the record
// (opcode, key, value) of op i+2 are already in flight, so the three
// dependent round trips of an op overlap with those of its neighbours.
// Chains longer than the head slab continue synchronously (walk_on).  All
// groups of a block run the same number of rounds (idle ones inactive), as
// the block-level combining in phase_r_op requires.
template <class L, int G, int BT, bool BC>
__device__ __forceinline__ void phase_r_range(const TableDev& t, const OpsDev& o,
                                              const OutDev& out, uint64_t lo, uint64_t m,
                                              uint64_t gid, uint64_t ng, RepEnt* rb,
                                              RepEnt* lrep = nullptr, uint32_t lcap = 0) {
    using K = typename L::K;
    using V = typename L::V;
    constexpr int S = PW<L, G>::N;
    uint64_t rounds = 0;  // BC: rounds of the block's first group, the most of any
    CombTab<BT>* tabs = nullptr;
    uint32_t* rcnt = rep_counter();
    if constexpr (BC) {
        const uint64_t gid0 = gid - threadIdx.x / G;  // the block's first group
        if (gid0 >= m) return;  // block-uniform
        rounds = (m - gid0 + ng - 1) / ng;
        tabs = comb_tables<BT>();
    } else {
        if (gid >= m) return;
    }
    uint64_t li = gid;
    uint8_t op = OP_GET;
    K key = K(0);
    V val = V(0);
    if (!BC || li < m) load_op<L>(o, lo + li, op, key, val);
    K blk[S];
#pragma unroll
    for (int e = 0; e < S; ++e) blk[e] = K(0);
    if (!BC || li < m) probe_issue<L, G, false>(t, bucket_of(t, (uint64_t)key), blk);
    uint64_t n1 = li + ng;
    uint8_t op1 = OP_GET;
    K key1 = K(0);
    V val1 = V(0);
    if (n1 < m) load_op<L>(o, lo + n1, op1, key1, val1);
    for (uint64_t r = 0;; ++r) {
        if constexpr (BC) {
            if (r >= rounds) break;
        }
        const uint64_t n2 = n1 + ng;
        uint8_t op2 = OP_GET;
        K key2 = K(0);
        V val2 = V(0);
        if (n2 < m) load_op<L>(o, lo + n2, op2, key2, val2);
        const uint64_t b = bucket_of(t, (uint64_t)key);
        uint64_t s = b;
        int slot = -1;
        uint32_t valid;
        if (!BC || li < m)
            walk_on<L, G, false>(t, probe_compute<L, G, false>(t, b, key, blk), s, slot, valid, key);
        K blk1[S];
#pragma unroll
        for (int e = 0; e < S; ++e) blk1[e] = K(0);
        if (n1 < m) probe_issue<L, G, false>(t, bucket_of(t, (uint64_t)key1), blk1);
        phase_r_op<L, G, BT, BC>(t, o, out, lo, (uint32_t)li, m, !BC || li < m, op, key, val, s,
                                 slot, BC ? tabs + (r & 1) : nullptr, rb, rcnt, lrep, lcap);
        if constexpr (!BC) {
            if (n1 >= m) break;
        }
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

// Unlinks the empty pool slabs of key's chain under the bucket's lock
// (reclaim policy).
template <class L, int G>
__device__ __forceinline__ void compact_locked(const TableDev& t, typename L::K key) {
    bool rf, ok;
    typename L::V rv;
    run_op<L, G, true, true>(t, OP_COMPACT, key, typename L::V(0), rf, rv, ok);
}

// Valid bitmap and next link of slab s, agent-coherent (one 8-byte load for
// u64 keys, where both sit in slot 0).
template <class L>
__device__ __forceinline__ void load_meta(const TableDev& t, uint64_t s, uint32_t& valid,
                                          uint32_t& next) {
    if constexpr (sizeof(typename L::K) == 8) {
        const uint64_t w = ld_coh(reinterpret_cast<const uint64_t*>(valid_word<L>(t, s)));
        valid = (uint32_t)w;
        next = (uint32_t)(w >> 32);
    } else {
        valid = ld_coh(valid_word<L>(t, s));
        next = ld_coh(next_word<L>(t, s));
    }
}

// Phase W insert (reference placement policy) of a key that phase R found
// absent.  No other op of the sub-batch writes this key and no phase-W op
// compares keys, so the slot can be claimed before its key is written: the
// group's lane 0 walks to the chain's last slab and claims its lowest empty
// slot with an atomic OR of the valid word (a lost race retries with the
// bitmap that OR returned).  Only a full last slab takes the bucket lock, to
// link a fresh slab holding the key in its first slot, unless another group
// linked one first.  Concurrent inserts into one bucket thus fill the same
// slots, and use the same number of slabs, as the locked insert in some
// serial order.  The lock is taken and released within one pass of the
// wave-uniform retry loop (as in run_op).
template <class L, int G>
__device__ __forceinline__ void w_insert(const TableDev& t, typename L::K key, typename L::V val) {
    const uint64_t b = bucket_of(t, (uint64_t)key);
    const int gl = Grp<G>::lane();
    uint64_t s = b, hops = 0, t0 = 0;
    uint32_t backoff = 1;
    int done = 0;
    for (;;) {
        int busy = 0;
        if (!done && gl == 0) {
            uint32_t valid, nx;
            load_meta<L>(t, s, valid, nx);
            while (nx != 0 && !done) {
                if (nx < t.nbuckets || nx >= t.nbuckets + t.pool || ++hops > t.pool) {
                    __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    done = 1;
                } else {
                    s = nx;
                    load_meta<L>(t, s, valid, nx);
                }
            }
            for (uint32_t empty = ~valid & L::CAPMASK; empty && !done;) {
                const int slot = __builtin_ctz(empty);
                const uint32_t old = __hip_atomic_fetch_or(valid_word<L>(t, s), 1u << slot,
                                                           __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
                if (old & (1u << slot)) {
                    empty = ~old & L::CAPMASK;
                } else {
                    st_coh(key_slot<L>(t, s, slot), key);
                    if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, s, slot), val);
                    done = 1;
                }
            }
            if (!done) {
                uint32_t* lw = t.locks + b;
                if (try_lock_excl(lw)) {
                    uint32_t v2, n2;
                    load_meta<L>(t, s, v2, n2);
                    if (n2 == 0 && (~v2 & L::CAPMASK) == 0) {
                        const uint64_t idx = __hip_atomic_fetch_add(
                            t.alloc_ctr, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (idx < t.pool) {
                            // A pool slab is all zeros until it is linked.
                            const uint64_t ns = t.nbuckets + idx;
                            st_coh(key_slot<L>(t, ns, L::SLOT0), key);
                            if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, ns, L::SLOT0), val);
                            st_coh(valid_word<L>(t, ns), 1u << L::SLOT0);
                            wait_vm();
                            st_coh(next_word<L>(t, s), (uint32_t)ns);
                        } else {
                            __hip_atomic_fetch_or(t.status, ST_POOL_EXHAUSTED, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
                        }
                        done = 1;
                    }
                    // else a slot was freed or a slab linked: next pass retries
                    wait_vm();
                    unlock_excl(lw);
                } else {
                    busy = 1;
                    const uint64_t now = now_ticks();
                    if (t0 == 0) t0 = now;
                    if (now - t0 > t.spin_ticks) {
                        __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                        done = 1;
                    }
                }
            }
        }
        done = grp_bcast(done, 0, G);
        if (__ballot(!done) == 0) break;
        if (__ballot(busy)) {
            for (uint32_t i = 0; i < backoff; ++i) __builtin_amdgcn_s_sleep(4);
            backoff = backoff < 64 ? backoff * 2 : 64;
        }
    }
}

// Writes key's final state after a sub-batch or run: the last mutation of
// the key was op (put of val, or remove), and loc is where the key sat in the
// table before it (slab << 5 | slot, or NOLOC).  All G lanes call it.
template <class L, int G, bool RC>
__device__ __forceinline__ void w_apply(const TableDev& t, uint8_t op, typename L::K key,
                                        typename L::V val, uint64_t loc) {
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    if (loc != NOLOC) {
        const uint64_t s = loc >> 5;
        const int slot = (int)(loc & 31u);
        uint32_t old = 0;
        if (gl == 0) {
            if (op == OP_PUT) {
                if constexpr (L::IS_MAP) st_coh(val_slot<L>(t, s, slot), val);
            } else {
                old = __hip_atomic_fetch_and(valid_word<L>(t, s), ~(1u << slot),
                                             __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // Reclaim: this remove left a pool slab empty; unlink it (and any
        // other empty pool slab of the chain) under the bucket's lock, which
        // also orders it after any insert into the same chain.
        if (RC && op != OP_PUT && s >= t.nbuckets) {
            old = grp_bcast(old, 0, G);
            if (((old & ~(1u << slot)) & L::CAPMASK) == 0) compact_locked<L, G>(t, key);
        }
    } else if (op == OP_PUT) {
        if constexpr (RC) {
            bool rf, ok;
            V rv;
            run_op<L, G, true, RC>(t, OP_PUT, key, val, rf, rv, ok);
        } else {
            w_insert<L, G>(t, key, val);
        }
    }
}

// Phase W of one final-mutation candidate: if it is still its key's last
// mutation (nothing exchanged after it), it writes the key's final state and
// frees the combining-table entry.
template <class L, int G, bool RC>
__device__ __forceinline__ void phase_w_ent(const TableDev& t, const OpsDev& o, uint64_t lo,
                                            const RepEnt& e) {
    using K = typename L::K;
    using V = typename L::V;
    const int gl = Grp<G>::lane();
    // The value is loaded whatever the opcode (every op has a value slot), so
    // the record's loads and the head check share one round trip.
    const uint64_t i = lo + e.li;
    const uint8_t op = ld_opcode(o, i);
    const K key = ld_stride<K>(o.keys, o.kstride, i);
    const V val = (L::IS_MAP && o.vals) ? ld_stride<V>(o.vals, o.vstride, i) : V(0);
    uint32_t* head = (e.g & LOC_G) ? &t.loc_head[e.g & ~LOC_G] : &t.gt_head[e.g];
    if (ld_agent(head) != e.li + 1u) return;  // a later mutation of this key exists
    if (gl == 0) {
        __hip_atomic_store(head, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (e.g <= t.gt_mask)
            __hip_atomic_store(&t.gt_key[e.g], GT_EMPTY, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    w_apply<L, G, RC>(t, op, key, val, e.loc);
}

// Phase W of a block: the cnt final-mutation candidates of its list (the
// first lcap in LDS at lrep, the rest at rb), one per group.
template <class L, int G, bool RC>
__device__ __forceinline__ void phase_w_block(const TableDev& t, const OpsDev& o, uint64_t lo,
                                              const RepEnt* rb, uint32_t cnt,
                                              const RepEnt* lrep = nullptr, uint32_t lcap = 0) {
    const uint32_t gpb = blockDim.x / G;
    for (uint32_t k = threadIdx.x / G; k < cnt; k += gpb)
        phase_w_ent<L, G, RC>(t, o, lo, k < lcap ? lrep[k] : rb[k]);
}

}  // namespace lslab_dev
