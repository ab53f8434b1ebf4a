// Runs of the persistent batched kernel: up to RUN_MAX_BATCHES consecutive
// batches executed together with three grid barriers in all, instead of one
// phase R and grid barrier per batch.
//
// Batch order is kept per key through two run-scope tables:
//   - chain table: one region per batch of the run; a mutation of key k in
//     batch r exchanges on k's entry of region r, which orders the batch's
//     mutations of k (as the single-batch engine's combining table does);
//   - key table: per key, the bitmap of the run's batches that mutate it.
// Pass A runs the mutations only: one exchange per mutation on its (key,
// batch) entry of the chain table; a mutation with a predecessor gets its
// result at once, and the first mutation of each (key, batch) sets the
// batch's bit in the key's bitmap.  After a
// grid barrier the tables are read-only, so pass B reads them with plain,
// cacheable loads: every get of batch r, and every first mutation of a chain
// of batch r, takes the key's state from the last mutation of the nearest
// earlier batch r' < r that mutates it (the tail of its chain in region r'),
// or from the table as it was before the run when no earlier batch does.
// The first mutation of the key's last chain also appends the key's final
// mutation, with the key's location in the table, to its block's list; phase
// W writes each key's final state once and empties both tables.
//
// Per key the serial order is: batch by batch, the batch's gets, then its
// mutations in exchange order -- the single-batch engine's order within each
// batch, and every batch sees all effects of the batches before it.
#pragma once
#include "engine.hpp"

namespace lslab_dev {

constexpr int RUN_MAX_BATCHES = 128;   // one bit per batch in a key's bitmap
constexpr uint32_t RUN_CF = 0x80000000u;  // run_pred: first mutation of its chain
constexpr uint32_t NOENT = 0xffffffffu;

// The current run's batch boundaries relative to its first op, and the
// window of absolute boundaries it was cut from (LDS).
struct RunOffs {
    uint64_t off[RUN_MAX_BATCHES + 1];
    uint64_t win[RUN_MAX_BATCHES + 1];
    int nb;
};

// Two run windows: the current run's and the next run's (loaded ahead).
__device__ __forceinline__ RunOffs* run_offs() {
    __shared__ RunOffs ro[2];
    return ro;
}

// The block's LDS counters of chain entries to clear, one per table set.
__device__ __forceinline__ uint32_t* clr_counters() {
    __shared__ uint32_t cnt[2];
    return cnt;
}

// Batch of run op li: the largest r with off[r] <= li (an empty batch owns
// no op, so it is never the answer).
__device__ __forceinline__ int run_batch_of(const RunOffs& ro, uint64_t li) {
    int lo = 0, hi = ro.nb - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (ro.off[mid] <= li) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Chain-table region of batch r: 2^k >= 2 * (ops of the batch) entries at
// 4 * (the batch's first op); 2^k <= 4 * ops, so regions never overlap and
// the run's regions fit in 4 * (run ops) entries.
__device__ __forceinline__ void run_region(const RunOffs& ro, int r, uint64_t& base,
                                           uint32_t& cmask) {
    const uint64_t sz = ro.off[r + 1] - ro.off[r];
    const uint64_t cap = sz <= 1 ? 2ull : (1ull << (64 - __builtin_clzll(2 * sz - 1)));
    base = 4 * ro.off[r];
    cmask = (uint32_t)(cap - 1);
}

// Claims or finds key's entry in region r: a coherent load of each probed
// slot, and a compare-and-swap only on an empty one (the mutations of a hot
// key in one batch find its entry without a read-modify-write); won: this
// call claimed it (exactly one call per (key, batch) does).  The
// all-ones key has a dedicated entry per batch after the regions, whose
// claim is the exchange that finds its head empty.
__device__ __forceinline__ uint32_t rc_claim(const TableDev& t, const RunTabs& rt, uint64_t key, uint64_t base,
                                             uint32_t cmask, int r, bool& won) {
    won = false;
    if (key == GT_EMPTY) return (uint32_t)(t.rc_cap + r);
    uint32_t h = (uint32_t)mix64(key) & cmask;
    for (uint32_t p = 0; p <= cmask; ++p) {
        const uint64_t cur = ld_agent(&rt.rc[base + h].key);
        if (cur == key) return (uint32_t)(base + h);
        if (cur == GT_EMPTY) {
            uint64_t e = GT_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&rt.rc[base + h].key, &e, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) {
                won = true;
                return (uint32_t)(base + h);
            }
            if (e == key) return (uint32_t)(base + h);
        }
        h = (h + 1u) & cmask;
    }
    return NOENT;  // unreachable: a region is at most half full
}

// Tail (last mutation + 1) of key's chain in region r, read with plain loads
// (pass B: the table is read-only); 0 if the batch has none.
__device__ __forceinline__ uint32_t rc_tail(const TableDev& t, const RunTabs& rt, uint64_t key, uint64_t base,
                                            uint32_t cmask, int r) {
    if (key == GT_EMPTY) return rt.rc[t.rc_cap + r].head;
    uint32_t h = (uint32_t)mix64(key) & cmask;
    for (uint32_t p = 0; p <= cmask; ++p) {
        const uint4 w = *reinterpret_cast<const uint4*>(&rt.rc[base + h]);
        const uint64_t k = ((uint64_t)w.y << 32) | w.x;
        if (k == key) return w.z;
        if (k == GT_EMPTY) return 0;
        h = (h + 1u) & cmask;
    }
    return 0;
}

// A key's 128-bit bitmap of the run's batches that mutate it.
struct BatchMask {
    uint64_t lo, hi;
};

// The bits of m below bit r.
__device__ __forceinline__ BatchMask bm_below(BatchMask m, int r) {
    if (r < 64) return BatchMask{m.lo & ((1ull << r) - 1ull), 0};
    return BatchMask{m.lo, m.hi & ((1ull << (r - 64)) - 1ull)};
}

// Highest set bit of m, or -1.
__device__ __forceinline__ int bm_top(BatchMask m) {
    if (m.hi) return 127 - __builtin_clzll(m.hi);
    return m.lo ? 63 - __builtin_clzll(m.lo) : -1;
}

// Whether no bit of m above bit r is set.
__device__ __forceinline__ bool bm_none_above(BatchMask m, int r) {
    if (r >= 64) return r == 127 || (m.hi >> (r - 63)) == 0;
    return m.hi == 0 && (r == 63 || (m.lo >> (r + 1)) == 0);
}

// Claims or finds key's entry in the key table (the all-ones key: the last).
__device__ __forceinline__ uint32_t rk_claim(const TableDev& t, const RunTabs& rt, uint64_t key) {
    if (key == GT_EMPTY) return t.rk_mask + 1u;
    uint32_t h = (uint32_t)mix64(key) & t.rk_mask;
    for (uint32_t p = 0; p <= t.rk_mask; ++p) {
        // Keys are only ever added during a run: a coherent load that finds
        // the key skips the compare-and-swap (a Zipf-hot key is registered
        // by its first chain in every batch of the run).
        const uint64_t cur = ld_agent(&rt.rk[h].key);
        if (cur == key) return h;
        if (cur == GT_EMPTY) {
            uint64_t e = GT_EMPTY;
            if (__hip_atomic_compare_exchange_strong(&rt.rk[h].key, &e, key, __ATOMIC_RELAXED,
                                                     __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ||
                e == key)
                return h;
        }
        h = (h + 1u) & t.rk_mask;
    }
    return NOENT;
}

// Key's entry and batch bitmap in the key table, plain loads (pass B).  The
// first slot's entry (key and low mask word in w0, high mask word in w1) was
// loaded ahead; NOENT when the run does not mutate the key.
__device__ __forceinline__ uint32_t rk_lookup(const TableDev& t, const RunTabs& rt, uint64_t key, uint4 w0,
                                              uint64_t w1, BatchMask& mask) {
    mask = BatchMask{0, 0};
    if (key == GT_EMPTY) {
        const RunKey& e = rt.rk[t.rk_mask + 1u];
        mask = BatchMask{e.mask[0], e.mask[1]};
        return (mask.lo | mask.hi) ? t.rk_mask + 1u : NOENT;
    }
    uint32_t h = (uint32_t)mix64(key) & t.rk_mask;
    uint4 w = w0;
    uint64_t wh = w1;
    for (uint32_t p = 0; p <= t.rk_mask; ++p) {
        const uint64_t k = ((uint64_t)w.y << 32) | w.x;
        if (k == key) {
            mask = BatchMask{((uint64_t)w.w << 32) | w.z, wh};
            return h;
        }
        if (k == GT_EMPTY) return NOENT;
        h = (h + 1u) & t.rk_mask;
        w = *reinterpret_cast<const uint4*>(&rt.rk[h]);
        wh = rt.rk[h].mask[1];
    }
    return NOENT;
}

__device__ __forceinline__ void rk_first(const TableDev& t, const RunTabs& rt, uint64_t key,
                                         uint4& w0, uint64_t& w1) {
    if (key == GT_EMPTY) {
        w0 = make_uint4(0xffffffffu, 0xffffffffu, 0u, 0u);
        w1 = 0;
        return;
    }
    const RunKey* e = &rt.rk[(uint32_t)mix64(key) & t.rk_mask];
    w0 = *reinterpret_cast<const uint4*>(e);
    w1 = e->mask[1];
}

__device__ __forceinline__ void corrupt(const TableDev& t) {
    __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Ops of a run whose mutations one wavefront of pass A gathers per step:
// 64 * RUN_A_CHUNKS at most, and fewer when the run is small, so that every
// wavefront gets about one step.
constexpr int RUN_A_CHUNKS = 16;

// Chunks per pass-A step for a run of m ops over nw wavefronts.
__host__ __device__ constexpr inline uint64_t run_a_chunks(uint64_t m, uint64_t nw) {
    uint64_t p = (m + 64 * nw - 1) / (64 * nw);
    return p > RUN_A_CHUNKS ? RUN_A_CHUNKS : (p ? p : 1);
}

// Ops pass A hands one block of nb (1024 threads each) for runs of at most
// run_cap ops, and so the length of each block's clear list: a wavefront
// takes ceil(steps / nw) steps of 64 * P ops.  It does not fall for a smaller
// run (P shrinks only while every wavefront takes at most one step).
__host__ __device__ constexpr inline uint64_t run_clr_stride(uint64_t run_cap, uint64_t nb) {
    const uint64_t nw = nb * 16;
    const uint64_t p = run_a_chunks(run_cap, nw);
    const uint64_t steps = (run_cap + 64 * p - 1) / (64 * p);
    return 1024 * p * ((steps + nw - 1) / nw);
}

// The wavefront's buffer of mutations gathered by pass A (op index | opcode
// << 30), two rounds deep.
__device__ __forceinline__ uint32_t* run_a_buf() {
    __shared__ uint32_t buf[1024 / 64][128];
    return buf[threadIdx.x / 64];
}

// One round of pass A: the wavefront's c <= 64 gathered mutations, one per
// lane.  Each mutation exchanges itself on its (key, batch) entry of the
// chain table, so the exchanges order the batch's mutations of the key.  A
// mutation whose exchange returns a predecessor writes its result now; the
// first of a (key, batch) is left to pass B (run_pred = RUN_CF | entry).  The
// op that claims a (key, batch) entry registers the batch in the key's bitmap
// and lists the entry for clearing (clr, counted by ccnt in LDS), in parallel
// with its exchange.  (Chaining a wavefront's mutations of one key in lane
// order first, one exchange per chain, cost more in the ballot loop that
// finds the chains than the exchanges it saved on this workload.)
template <class L>
__device__ __forceinline__ void run_a_round(const TableDev& t, const RunTabs& rt, const OpsDev& o,
                                            const OutDev& out, uint64_t lo, uint64_t m,
                                            const RunOffs& ro, const uint32_t* buf, uint32_t c,
                                            uint32_t* clr, uint32_t* ccnt) {
    using K = typename L::K;
    using V = typename L::V;
    const int lane = (int)__lane_id();
    const bool lead = (uint32_t)lane < c;
    uint32_t li = 0;
    uint8_t op = OP_GET;
    K key = K(0);
    int r = 0;
    if (lead) {
        const uint32_t e = buf[lane];
        li = e & 0x3fffffffu;
        op = (uint8_t)(e >> 30);
        key = ld_stride<K>(o.keys, o.kstride, lo + li);
        r = run_batch_of(ro, li);
    }
    const uint64_t kk = (uint64_t)key;
    uint32_t ce = NOENT, fpred = 0;
    bool won = false;
    if (lead) {
        uint64_t base;
        uint32_t cmask;
        run_region(ro, r, base, cmask);
        ce = rc_claim(t, rt, kk, base, cmask, r, won);
        if (ce != NOENT) {
            fpred = __hip_atomic_exchange(&rt.rc[ce].head, li + 1u, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
            auto enroll = [&]() {  // the batch into the key's bitmap
                const uint32_t gk = rk_claim(t, rt, kk);
                if (gk != NOENT)
                    __hip_atomic_fetch_or(&rt.rk[gk].mask[r >> 6], 1ull << (r & 63),
                                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                else
                    corrupt(t);
            };
            // The claim alone says whether an ordinary key's op registers
            // the batch, so its registration is issued while the exchange is
            // in flight.  The reserved key has a dedicated entry per batch
            // and no claim: its first exchange registers.
            if (kk != GT_EMPTY) {
                if (won) enroll();
            } else {
                won = fpred == 0;
                if (won) enroll();
            }
        } else {
            corrupt(t);
        }
    }
    if (lead) {
        if (fpred != 0) {
            st_coh(&t.run_pred[li], fpred);
            mut_result<L>(t, o, out, lo, m, li, op, fpred, false, V(0));
        } else {
            st_coh(&t.run_pred[li], RUN_CF | ce);
        }
    }
    const uint64_t fm = __ballot(won);
    if (fm) {
        const int l0 = __builtin_ctzll(fm);
        uint32_t wb = 0;
        if (lane == l0)
            wb = __hip_atomic_fetch_add(ccnt, (uint32_t)__builtin_popcountll(fm), __ATOMIC_RELAXED,
                                        __HIP_MEMORY_SCOPE_WORKGROUP);
        wb = __builtin_amdgcn_readlane(wb, l0);
        if (won)
            st_coh(&clr[wb + (uint32_t)__builtin_popcountll(fm & ((1ull << lane) - 1ull))], ce);
    }
}

// Pass A over the run's ops [lo, lo + m): the mutations.  Each wavefront
// takes steps of 64 * P consecutive ops, reads them one 64-op chunk at a
// time (one opcode per lane) and gathers the chunk's mutations into its LDS
// buffer, so that a round of the chain-and-exchange work runs 64 mutations
// at once however sparse they are among the gets.  P is chosen so that every
// wavefront gets about one step.
template <class L>
__device__ __forceinline__ void run_pass_a(const TableDev& t, const RunTabs& rt, const OpsDev& o,
                                           const OutDev& out, uint64_t lo, uint64_t m,
                                           const RunOffs& ro, uint32_t* clr, uint32_t* ccnt) {
    const int lane = (int)__lane_id();
    const uint64_t nw = (uint64_t)gridDim.x * (blockDim.x / 64);
    const uint64_t wid = (uint64_t)blockIdx.x * (blockDim.x / 64) + threadIdx.x / 64;
    const uint64_t P = run_a_chunks(m, nw);
    uint32_t* buf = run_a_buf();
    uint32_t cnt = 0;  // gathered, wave-uniform
    for (uint64_t sb = wid * 64 * P; sb < m; sb += nw * 64 * P) {
        // the step's opcodes, all loads in flight together
        uint8_t ops[RUN_A_CHUNKS];
#pragma unroll
        for (int k = 0; k < RUN_A_CHUNKS; ++k) {
            const uint64_t li = sb + 64 * k + lane;
            ops[k] = (k < (int)P && li < m) ? ld_opcode(o, lo + li) : (uint8_t)OP_GET;
        }
#pragma unroll
        for (int k = 0; k < RUN_A_CHUNKS; ++k) {
            if (k >= (int)P) break;
            const uint64_t li = sb + 64 * k + lane;
            const uint8_t op = ops[k];
            const bool mut = op != OP_GET;
            const uint64_t bal = __ballot(mut);
            if (mut)
                buf[cnt + (uint32_t)__builtin_popcountll(bal & ((1ull << lane) - 1ull))] =
                    (uint32_t)li | ((uint32_t)op << 30);
            cnt += (uint32_t)__builtin_popcountll(bal);
            if (cnt >= 64) {
                __builtin_amdgcn_wave_barrier();
                run_a_round<L>(t, rt, o, out, lo, m, ro, buf, 64, clr, ccnt);
                // keep the overflow for the next round
                const uint32_t carry = lane + 64 < (int)cnt ? buf[lane + 64] : 0u;
                __builtin_amdgcn_wave_barrier();
                if (lane + 64 < (int)cnt) buf[lane] = carry;
                __builtin_amdgcn_wave_barrier();
                cnt -= 64;
            }
        }
    }
    if (cnt) {
        __builtin_amdgcn_wave_barrier();
        run_a_round<L>(t, rt, o, out, lo, m, ro, buf, cnt, clr, ccnt);
    }
}

// State a mutation leaves: present with its value (put) or absent (remove).
template <class L>
__device__ __forceinline__ void op_effect(const OpsDev& o, uint64_t i, bool& found,
                                          typename L::V& cur) {
    using V = typename L::V;
    const uint8_t pop = ld_opcode(o, i);
    const V pv = (L::IS_MAP && o.vals) ? ld_stride<V>(o.vals, o.vstride, i) : V(0);
    found = pop == OP_PUT;
    cur = found ? pv : V(0);
}

// One op of pass B (see run_pass_b), its record (opcode, key, value and
// run_pred) already loaded.
template <class L, int G>
__device__ __forceinline__ void run_b_op(const TableDev& t, const RunTabs& rt, const OpsDev& o,
                                         const OutDev& out, uint64_t lo, uint64_t m,
                                         const RunOffs& ro, uint64_t li,
                                         uint8_t op, typename L::K key, uint32_t pred, RepEnt* rb,
                                         RepEnt* lrep, uint32_t lcap, uint32_t* rcnt) {
    using K = typename L::K;
    using V = typename L::V;
    constexpr int S = PW<L, G>::N;
    const int gl = Grp<G>::lane();
    if (op != OP_GET && !(pred & RUN_CF)) return;  // group-uniform: written in pass A
    const uint64_t kk = (uint64_t)key;
    const uint64_t b = bucket_of(t, kk);
    K blk[S];
    uint4 w0 = make_uint4(0, 0, 0, 0);
    uint64_t w1 = 0;
    if (gl == 0) rk_first(t, rt, kk, w0, w1);
    const int r = run_batch_of(ro, li);
    BatchMask mask{0, 0};
    uint32_t gk = NOENT;
    if (gl == 0) gk = rk_lookup(t, rt, kk, w0, w1, mask);
    mask.lo = grp_bcast(mask.lo, 0, G);
    mask.hi = grp_bcast(mask.hi, 0, G);
    gk = grp_bcast(gk, 0, G);
    const int r2 = bm_top(bm_below(mask, r));  // nearest earlier batch that mutates the key
    const bool earlier = r2 >= 0;
    const bool last = op != OP_GET && bm_none_above(mask, r);
    V cur = V(0);
    uint64_t s = b;
    int slot = -1;
    if (!earlier || last) {
        uint32_t valid;
        probe_issue<L, G, false>(t, b, blk);
        walk_on<L, G, false>(t, probe_compute<L, G, false>(t, b, key, blk), s, slot, valid, key);
        if constexpr (L::IS_MAP) {
            if (slot >= 0 && !earlier && gl == 0) cur = *val_slot<L>(t, s, slot);
        }
    }
    if (gl != 0) return;
    const bool in_table = slot >= 0;
    bool found = in_table;
    uint32_t p = 0;  // predecessor: the tail of the nearest earlier batch's chain
    if (earlier) {
        uint64_t base;
        uint32_t cmask;
        run_region(ro, r2, base, cmask);
        p = rc_tail(t, rt, kk, base, cmask, r2);
        if (p == 0 || p > m) {
            corrupt(t);
            p = 0;
        }
    }
    if (op == OP_GET) {
        if (p) op_effect<L>(o, lo + p - 1u, found, cur);
        write_result<L>(out, lo + li, found, cur);
        return;
    }
    mut_result<L>(t, o, out, lo, m, (uint32_t)li, op, p, found, cur);
    if (!last) return;
    const uint32_t c = pred & ~RUN_CF;
    const uint32_t fin = c < t.rc_cap + RUN_MAX_BATCHES ? rt.rc[c].head : 0u;
    if (gk == NOENT || fin == 0 || fin > m) {
        corrupt(t);
        return;
    }
    RepEnt e;
    e.li = fin - 1u;
    e.g = gk;
    e.loc = in_table ? ((s << 5) | (uint64_t)slot) : NOLOC;
    const uint32_t k = __hip_atomic_fetch_add(rcnt, 1u, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_WORKGROUP);
    if (k < lcap) lrep[k] = e;
    else rb[k] = e;
}

// Pass B over the run's ops: every get and every first mutation of a (key,
// batch) chain.  The key table is read first; the table is probed only when
// the key has no earlier mutation in the run (the op reads the table's
// value) or this is the key's last chain (the probe gives the key's location
// for phase W).  Under skew most ops of a hot key need neither, so probing
// after the key-table lookup, not beside it, saves more bucket-line reads
// than the serialisation costs.  A record's opcode, key and run_pred are
// loaded together, one op ahead.
template <class L, int G>
__device__ __forceinline__ void run_pass_b(const TableDev& t, const RunTabs& rt, const OpsDev& o,
                                           const OutDev& out, uint64_t lo, uint64_t m,
                                           const RunOffs& ro,
                                           uint64_t gid, uint64_t ng, RepEnt* rb, RepEnt* lrep,
                                           uint32_t lcap, uint32_t* rcnt) {
    using K = typename L::K;
    if (gid >= m) return;  // group-uniform
    auto load = [&](uint64_t li, uint8_t& op, K& key, uint32_t& pred) {
        op = ld_opcode(o, lo + li);
        key = ld_stride<K>(o.keys, o.kstride, lo + li);
        pred = t.run_pred[li];  // meaningful for mutations only
    };
    uint8_t op;
    K key;
    uint32_t pred;
    load(gid, op, key, pred);
    for (uint64_t li = gid;; li += ng) {
        const uint64_t ln = li + ng;
        uint8_t op_n = OP_GET;
        K key_n = K(0);
        uint32_t pred_n = 0;
        if (ln < m) load(ln, op_n, key_n, pred_n);
        run_b_op<L, G>(t, rt, o, out, lo, m, ro, li, op, key, pred, rb, lrep, lcap, rcnt);
        if (ln >= m) break;
        op = op_n;
        key = key_n;
        pred = pred_n;
    }
}

// Phase W of a run, per block: each listed key's final mutation writes the
// key's final state and empties its key-table entry; each listed chain entry
// is emptied.
template <class L, int G, bool RC>
__device__ __forceinline__ void run_phase_w(const TableDev& t, const RunTabs& rt, const OpsDev& o,
                                            uint64_t lo,
                                            const RepEnt* rb, uint32_t cnt, const RepEnt* lrep,
                                            uint32_t lcap, const uint32_t* clr, uint32_t ccnt) {
    using K = typename L::K;
    using V = typename L::V;
    const uint32_t gpb = blockDim.x / G;
    const int gl = Grp<G>::lane();
    for (uint32_t k = threadIdx.x / G; k < cnt; k += gpb) {
        const RepEnt e = k < lcap ? lrep[k] : rb[k];
        const uint64_t i = lo + e.li;
        const uint8_t op = ld_opcode(o, i);
        const K key = ld_stride<K>(o.keys, o.kstride, i);
        const V val = (L::IS_MAP && o.vals) ? ld_stride<V>(o.vals, o.vstride, i) : V(0);
        if (gl == 0) {  // the 32-B entry: key and mask[0], then mask[1] and pad
            st_coh16(&rt.rk[e.g].key, GT_EMPTY, 0);
            st_coh16(&rt.rk[e.g].mask[1], 0, 0);
        }
        w_apply<L, G, RC>(t, op, key, val, e.loc);
    }
    for (uint32_t k = threadIdx.x; k < ccnt; k += blockDim.x) {
        const uint32_t c = ld_coh(&clr[k]);
        if (c < t.rc_cap + RUN_MAX_BATCHES)  // the 16-B entry: key, head and pad
            st_coh16(&rt.rc[c], GT_EMPTY, 0);
    }
}

}  // namespace lslab_dev
