// Slab layout and wave64 device building blocks for the LSlab map and set on
// gfx950.
//
// Slab layout (one slab = one key block, then one value block for maps):
//   key block: 32 slots of K.  31 slots hold keys, one holds meta.
//     K = u64        : slot 0 = valid bitmap (low 32 bits) | next slab (high 32),
//                      keys in slots 1..31
//     K = u32        : keys in slots 0..30, slot 31 = valid bitmap; the next
//                      link is value slot 31 (maps) or, for sets, which have
//                      no value block, entry s of a u32 link array that
//                      follows the slab array
//   value block (maps): 32 slots of V; value slot i belongs to key slot i.
//   Valid bit i belongs to key slot i.
// Every layout keeps the reference's 31 keys per slab and its valid-bit
// semantics (detail/slab_node.h:13-35) with a 32-bit slab index in place of
// the 8-byte pointer, so a u64 key block is exactly 256 B (two 128-B lines)
// and a u32 key block one line.  The u64 meta word sits in the first line
// together with the first 15 keys: keys fill the lowest free slot first, so a
// probe of a slab holding at most 15 keys, or of a key among its first 15,
// reads one line (see PW).  A u32 probe reads the key line; its next link is
// read only when the key is not in the slab.  Slabs 0..2^n-1 are the bucket
// heads; pool slabs follow them in the same array, so next == 0 means end of
// chain.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace lslab_dev {

enum : uint8_t { OP_GET = 0, OP_PUT = 1, OP_REMOVE = 2 };
// ST_BAD_INPUT: batch or request offsets that decrease (the call did no work
// past them).
enum : uint32_t { ST_POOL_EXHAUSTED = 1u, ST_TIMEOUT = 2u, ST_CORRUPT = 4u, ST_BAD_INPUT = 8u };

constexpr uint32_t LOCK_WRITER = 0x80000000u;
constexpr uint64_t GT_EMPTY = ~0ull;
constexpr uint64_t NOLOC = ~0ull;
constexpr uint64_t NOSLAB = ~0ull;
// Internal opcode (never in a caller's op stream): unlink every empty pool
// slab of the key's chain (reclaim policy), under the bucket's lock.
constexpr uint8_t OP_COMPACT = 3;
// 10 s at the 100 MHz constant clock: the default bound on every spin, so
// that a lost lock holder or barrier arrival turns into a status code instead
// of a hung device (a table's tuning may set another, TableDev::spin_ticks).
constexpr uint64_t SPIN_LIMIT_TICKS = 1000000000ull;
constexpr uint64_t TICKS_PER_MS = 100000ull;

template <typename K_, typename V_, bool MAP>
struct Layout {
    using K = K_;
    using V = V_;
    static constexpr bool IS_MAP = MAP;
    static constexpr uint32_t KB = 32u * sizeof(K);
    static constexpr uint32_t VB = MAP ? 32u * sizeof(V) : 0u;
    static constexpr uint32_t SLAB_BYTES = KB + VB;
    static constexpr int CAP = 31;
    static constexpr int SLOT0 = (sizeof(K) == 8) ? 1 : 0;  // first key slot
    static constexpr uint32_t CAPMASK = ((1u << CAP) - 1u) << SLOT0;  // key slots
    // u32 sets keep their next links in an array after the slab array.
    static constexpr bool SIDE_NEXT = sizeof(K) == 4 && !MAP;
    // Byte offsets inside a slab.
    static constexpr uint32_t VALID_OFF = (sizeof(K) == 8) ? 0u : 31u * 4u;
    static constexpr uint32_t NEXT_OFF = (sizeof(K) == 8) ? 4u : (MAP ? KB + 31u * sizeof(V) : 0u);
    // Device bytes per slab, its share of the link array included.
    static constexpr uint32_t FOOTPRINT = SLAB_BYTES + (SIDE_NEXT ? 4u : 0u);
};

// Device view of one table (passed by value into kernels).
// A two-phase mutation that may be its key's final one in the sub-batch (the
// last op of a combining chain): its op index, its key's combining-table
// entry and the key's pre-batch location (slab << 5 | slot, or NOLOC).
struct RepEnt {
    uint32_t li;
    uint32_t g;
    uint64_t loc;
};

// Scratch of a run of batches in the persistent batched kernel.  The key
// table holds, per key the run mutates, the bitmap of the run's batches that
// mutate it; the chain table holds, per (key, batch), the last mutation of
// that batch's exchange chain.  Key and payload share one entry (32 B
// aligned, so one cache sector holds it).
struct alignas(32) RunKey {
    uint64_t key;      // GT_EMPTY = free
    uint64_t mask[2];  // bit r (word r / 64): batch r of the run mutates the key
    uint64_t pad;
};
struct RunChain {
    uint64_t key;   // GT_EMPTY = free
    uint32_t head;  // last exchanged mutation (run op index + 1)
    uint32_t pad;
};
// Phase W clears an entry with 16-B stores: key table as (key, mask[0]) and
// (mask[1], pad), chain table as (key, head | pad).
static_assert(sizeof(RunKey) == 32 && offsetof(RunKey, mask) == 8, "RunKey layout");
static_assert(sizeof(RunChain) == 16 && offsetof(RunChain, head) == 8, "RunChain layout");

// One set of run tables (the batched kernel alternates two, so that one
// run's write-back overlaps the next run's first pass).
struct RunTabs {
    RunKey* rk;     // key table: rk_mask + 2 entries (the last: the all-ones key)
    RunChain* rc;   // chain tables: rc_cap entries, then one per batch of a run
                    // for the all-ones key
    uint32_t* clr;  // chain entries to clear, one list per block (as rep)
};

struct TableDev {
    char* slabs;           // (nbuckets + pool) * SLAB_BYTES, 256-B aligned, then
                           // (u32 sets) one u32 next link per slab
    uint32_t* locks;       // one RW lock word per bucket
    uint64_t* alloc_ctr;   // bump pointer into the pool (slabs handed out)
    uint32_t* status;      // sticky ST_* bits
    uint64_t* gt_key;      // batch combining table: key per entry (GT_EMPTY = free)
    uint32_t* gt_head;     // last mutation (sub-batch op index + 1) per entry
    RepEnt* rep;           // final-mutation candidates, one list per block (see RepEnt)
    uint32_t* rep_cnt;     // per block of the two-phase launch pair: its list length
    uint64_t nbuckets;
    uint64_t pool;
    uint32_t mask;         // nbuckets - 1
    uint32_t gt_mask;      // gt capacity - 1
    uint32_t hash_mix;     // bucket by mix64(key) instead of the key itself
    uint64_t* free_head;   // reclaim: free pool slabs, (tag << 32 | slab), 0 = empty
    uint64_t* free_count;  // reclaim: slabs on the free list
    RunTabs run[2];        // run tables, used by runs of alternating parity
    uint32_t* run_pred;    // per op of a run: chain predecessor (op index + 1) or
                           // RUN_CF | chain entry for the first mutation of a chain
    uint64_t rc_cap;       // 4 * (ops a run may hold)
    uint32_t rk_mask;      // key-table capacity - 1 (at least 2 * ops a run may hold)
    uint64_t spin_ticks;   // bound on every device spin (100 MHz ticks)
    uint32_t* loc_head;    // per key slot of the table (slab << 5 | slot): the last
                           // mutation (op index + 1) of the key found there in the
                           // current sub-batch; null when not allocated
};

// RepEnt.g of a final-mutation candidate whose chain head is loc_head[g & ~LOC_G].
constexpr uint32_t LOC_G = 0x80000000u;

// Operation stream: per-op opcode (or a uniform one), keys and values with a
// byte stride so that both SoA arrays and the reference's AoS pair<K,V> work.
struct OpsDev {
    const uint8_t* ops;    // null -> every op is `uop`
    const char* keys;
    const char* vals;      // may be null (gets, set ops)
    uint32_t kstride, vstride;
    uint8_t uop;
};

// Result sinks (any may be null): pair<bool,V> as the reference lays it out
// (bool at 0, V at alignof(V)), a plain V array (reference map put) or a bool
// array (reference set ops).
struct OutDev {
    char* pairs;
    char* vals;
    uint8_t* flags;
};

__device__ __forceinline__ uint64_t now_ticks() { return __builtin_amdgcn_s_memrealtime(); }

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

// Bucket of a key: the reference's identity hash masked to 2^n buckets
// (include/lslab/hash.h:10-15, map.h:95-96), or the murmur3 finalizer of
// the key when the table was created with LSLAB_HASH_MIX.
__device__ __forceinline__ uint64_t bucket_of(const TableDev& t, uint64_t key) {
    return (t.hash_mix ? mix64(key) : key) & t.mask;
}

template <int G>
struct Grp {
    static_assert(G == 1 || G == 2 || G == 4 || G == 8 || G == 16 || G == 32, "lanes per op");
    static constexpr int S = 32 / G;  // key slots per lane
    __device__ static int lane() { return (int)(__lane_id() & (G - 1)); }
    __device__ static int base() { return (int)(__lane_id() & ~(G - 1)); }
    __device__ static uint64_t bits(uint64_t ballot) {
        return (ballot >> base()) & ((1ull << G) - 1ull);
    }
};

// Key slots one probe load covers, per G-lane group: the whole key block, or
// for u64 keys (G <= 16) its first 128-B line, slots 0..15 (meta word and the
// first 15 keys).  The second line is loaded only when the first cannot
// settle the lookup: no match there and a valid bit set in slots 16..31.
template <class L, int G>
struct PW {
    static constexpr bool SPLIT = sizeof(typename L::K) == 8 && G <= 16;
    static constexpr int SPAN = SPLIT ? 16 : 32;  // slots per load
    static constexpr int N = SPAN / G;            // slots per lane per load
};

template <typename T>
__device__ __forceinline__ T grp_bcast(T v, int src, int width) {
    return __shfl(v, src, width);
}
__device__ __forceinline__ uint64_t grp_bcast(uint64_t v, int src, int width) {
    uint32_t lo = __shfl((uint32_t)v, src, width);
    uint32_t hi = __shfl((uint32_t)(v >> 32), src, width);
    return ((uint64_t)hi << 32) | lo;
}

// Loads `S` consecutive key slots starting at p (p is S*sizeof(K)-aligned).
template <typename K, int S>
__device__ __forceinline__ void load_slots(const K* p, K (&out)[S]) {
    constexpr int BYTES = S * (int)sizeof(K);
    if constexpr (BYTES >= 16) {
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int j = 0; j < BYTES / 16; ++j) {
            uint4 w = q[j];
            uint32_t tmp[4] = {w.x, w.y, w.z, w.w};
            __builtin_memcpy(reinterpret_cast<char*>(out) + 16 * j, tmp, 16);
        }
    } else if constexpr (BYTES == 8) {
        uint2 w = *reinterpret_cast<const uint2*>(p);
        uint32_t tmp[2] = {w.x, w.y};
        __builtin_memcpy(out, tmp, 8);
    } else {
        uint32_t w = *reinterpret_cast<const uint32_t*>(p);
        __builtin_memcpy(out, &w, 4);
    }
}

// Agent-coherent variant: slab words that other waves of the same launch may
// be writing are read with agent-scope atomic loads (global_load ... sc1), so
// no stale cache line can answer and no cache-wide acquire fence is needed.
template <typename K, int S>
__device__ __forceinline__ void load_slots_coherent(const K* p, K (&out)[S]) {
    constexpr int BYTES = S * (int)sizeof(K);
    if constexpr (BYTES >= 8) {
        const uint64_t* q = reinterpret_cast<const uint64_t*>(p);
#pragma unroll
        for (int j = 0; j < BYTES / 8; ++j) {
            uint64_t w = __hip_atomic_load(q + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __builtin_memcpy(reinterpret_cast<char*>(out) + 8 * j, &w, 8);
        }
    } else {
        uint32_t w = __hip_atomic_load(reinterpret_cast<const uint32_t*>(p), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
        __builtin_memcpy(out, &w, 4);
    }
}

template <typename T>
__device__ __forceinline__ T ld_coh(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Write-through store of a shared slab word (global_store ... sc1).
template <typename T>
__device__ __forceinline__ void st_coh(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Write-through store of 16 bytes (lo at p, hi at p + 8): the encoding of an
// agent-scope store (sc1), four words wide.  For clearing a run-table entry
// in one store where each word only has to be visible after the next grid
// barrier (which waits for it), not as one atomic unit.  p: 8-B aligned.
__device__ __forceinline__ void st_coh16(void* p, uint64_t lo, uint64_t hi) {
    typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
    const u32x4 w = {(unsigned)lo, (unsigned)(lo >> 32), (unsigned)hi, (unsigned)(hi >> 32)};
    __asm__ volatile("global_store_dwordx4 %0, %1, off sc1" ::"v"(p), "v"(w) : "memory");
}
// Waits until this wave's earlier memory operations have completed; with
// every shared write issued as an agent-scope store or atomic, this is all a
// lock release or barrier arrival needs (no cache-wide write-back).
__device__ __forceinline__ void wait_vm() { __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory"); }

// Block barrier that orders LDS accesses only: unlike __syncthreads it does
// not drain the wave's outstanding global loads, so prefetches issued before
// it stay in flight across it.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

template <class L>
__device__ __forceinline__ char* slab_ptr(const TableDev& t, uint64_t s) {
    return t.slabs + s * (uint64_t)L::SLAB_BYTES;
}
template <class L>
__device__ __forceinline__ uint32_t* valid_word(const TableDev& t, uint64_t s) {
    return reinterpret_cast<uint32_t*>(slab_ptr<L>(t, s) + L::VALID_OFF);
}
template <class L>
__device__ __forceinline__ uint32_t* next_word(const TableDev& t, uint64_t s) {
    if constexpr (L::SIDE_NEXT)
        return reinterpret_cast<uint32_t*>(t.slabs + (t.nbuckets + t.pool) * L::SLAB_BYTES) + s;
    else
        return reinterpret_cast<uint32_t*>(slab_ptr<L>(t, s) + L::NEXT_OFF);
}
template <class L>
__device__ __forceinline__ typename L::K* key_slot(const TableDev& t, uint64_t s, int i) {
    return reinterpret_cast<typename L::K*>(slab_ptr<L>(t, s)) + i;
}
template <class L>
__device__ __forceinline__ typename L::V* val_slot(const TableDev& t, uint64_t s, int i) {
    return reinterpret_cast<typename L::V*>(slab_ptr<L>(t, s) + L::KB) + i;
}

// Result of probing one slab with one group of G lanes.  All fields are
// uniform across the group.
struct Probe {
    int slot;        // matching slot, or -1
    uint32_t valid;  // valid bitmap of the slab
    uint32_t next;   // next slab index (0 = last); only filled on a miss
};

// Group-cooperative probe of slab s for key.  Lane l of the group loads key
// slots [l*N, l*N+N) of the first load span in one coalesced access (PW).
// Plain loads when the caller guarantees no concurrent writer (read-only
// launches, phase R); coherent loads (COH) under a bucket lock.
// probe_issue starts the load; probe_compute finishes the probe (loading the
// second line of a split u64 block when needed).
template <class L, int G, bool COH>
__device__ __forceinline__ void probe_issue(const TableDev& t, uint64_t s,
                                            typename L::K (&mine)[PW<L, G>::N]) {
    using K = typename L::K;
    constexpr int N = PW<L, G>::N;
    const K* kb = reinterpret_cast<const K*>(slab_ptr<L>(t, s));
    if constexpr (COH) load_slots_coherent<K, N>(kb + Grp<G>::lane() * N, mine);
    else load_slots<K, N>(kb + Grp<G>::lane() * N, mine);
}

// Slot of key among the loaded slots [base + gl*N, base + gl*N + N), or -1
// (group-uniform).
template <class L, int G>
__device__ __forceinline__ int probe_match(uint32_t valid, typename L::K key, int base,
                                           const typename L::K (&mine)[PW<L, G>::N]) {
    constexpr int N = PW<L, G>::N;
    const uint32_t v = valid & L::CAPMASK;
    int hit = -1;
#pragma unroll
    for (int e = 0; e < N; ++e) {
        const int slot = base + Grp<G>::lane() * N + e;
        if (((v >> slot) & 1u) && mine[e] == key) hit = slot;
    }
    const uint64_t b = Grp<G>::bits(__ballot(hit >= 0));
    return b ? grp_bcast(hit, __builtin_ctzll(b), G) : -1;
}

template <class L, int G, bool COH>
__device__ __forceinline__ Probe probe_compute(const TableDev& t, uint64_t s, typename L::K key,
                                               const typename L::K (&mine)[PW<L, G>::N]) {
    using K = typename L::K;
    constexpr int N = PW<L, G>::N;
    Probe p;
    if constexpr (sizeof(K) == 8) {
        const uint64_t meta = grp_bcast((uint64_t)mine[0], 0, G);  // slot 0, lane 0
        p.valid = (uint32_t)meta;
        p.next = (uint32_t)(meta >> 32);
    } else {  // u32 keys: the valid word is slot 31, the last lane's last slot
        p.valid = grp_bcast((uint32_t)mine[N - 1], G - 1, G);
        p.next = 0;
    }
    p.slot = probe_match<L, G>(p.valid, key, 0, mine);
    if constexpr (PW<L, G>::SPLIT) {
        if (p.slot < 0 && (p.valid & L::CAPMASK & 0xffff0000u)) {
            K hi[N];
            const K* kb = reinterpret_cast<const K*>(slab_ptr<L>(t, s)) + 16;
            if constexpr (COH) load_slots_coherent<K, N>(kb + Grp<G>::lane() * N, hi);
            else load_slots<K, N>(kb + Grp<G>::lane() * N, hi);
            p.slot = probe_match<L, G>(p.valid, key, 16, hi);
        }
    }
    if (p.slot < 0) {
        if constexpr (sizeof(K) == 4)
            p.next = COH ? ld_coh(next_word<L>(t, s)) : *next_word<L>(t, s);
    }
    return p;
}

template <class L, int G, bool COH>
__device__ __forceinline__ Probe probe_slab(const TableDev& t, uint64_t s, typename L::K key) {
    typename L::K mine[PW<L, G>::N];
    probe_issue<L, G, COH>(t, s, mine);
    return probe_compute<L, G, COH>(t, s, key, mine);
}

// Continues a chain walk from probe p of slab s (see walk).
template <class L, int G, bool COH>
__device__ __forceinline__ void walk_on(const TableDev& t, Probe p, uint64_t& s, int& slot,
                                        uint32_t& valid, typename L::K key) {
    // A chain has at most pool + 1 slabs and every link points into the slab
    // array; both are checked so that a damaged chain ends the walk with a
    // status bit instead of a runaway loop or a wild access.
    for (uint64_t hops = 0;; ++hops) {
        if (p.slot >= 0) {
            slot = p.slot;
            valid = p.valid;
            return;
        }
        if (p.next == 0) {
            slot = -1;
            valid = p.valid;
            return;
        }
        if (p.next < t.nbuckets || p.next >= t.nbuckets + t.pool || hops > t.pool) {
            if (Grp<G>::lane() == 0)
                __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            slot = -1;
            valid = L::CAPMASK;  // report the last slab as full: no insert lands here
            return;
        }
        s = p.next;
        p = probe_slab<L, G, COH>(t, s, key);
    }
}

// Walk of bucket b's chain for a mutation under the reclaim policy: as walk,
// and also the slab before s (prev, NOSLAB when s is the head) and the first
// slab up to s with an empty key slot (room, NOSLAB if none) with its bitmap.
template <class L, int G, bool COH>
__device__ __forceinline__ void walk_full(const TableDev& t, uint64_t b, typename L::K key,
                                          uint64_t& s, int& slot, uint32_t& valid,
                                          uint64_t& prev, uint64_t& room, uint32_t& room_valid) {
    s = b;
    prev = NOSLAB;
    room = NOSLAB;
    room_valid = 0;
    Probe p = probe_slab<L, G, COH>(t, s, key);
    for (uint64_t hops = 0;; ++hops) {
        if (room == NOSLAB && (~p.valid & L::CAPMASK)) {
            room = s;
            room_valid = p.valid;
        }
        if (p.slot >= 0) {
            slot = p.slot;
            valid = p.valid;
            return;
        }
        if (p.next == 0) {
            slot = -1;
            valid = p.valid;
            return;
        }
        if (p.next < t.nbuckets || p.next >= t.nbuckets + t.pool || hops > t.pool) {
            if (Grp<G>::lane() == 0)
                __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            slot = -1;
            valid = L::CAPMASK;
            room = NOSLAB;
            return;
        }
        prev = s;
        s = p.next;
        p = probe_slab<L, G, COH>(t, s, key);
    }
}

// Walks the chain of bucket b: returns the slab/slot holding key, or slot -1
// with `s` = last slab of the chain and `valid` its bitmap.  COH selects the
// agent-coherent loads (inside locked sections, where other waves write).
template <class L, int G, bool COH>
__device__ __forceinline__ void walk(const TableDev& t, uint64_t b, typename L::K key,
                                     uint64_t& s, int& slot, uint32_t& valid) {
    s = b;
    walk_on<L, G, COH>(t, probe_slab<L, G, COH>(t, b, key), s, slot, valid, key);
}

template <typename T>
__device__ __forceinline__ T ld_stride(const char* base, uint32_t stride, uint64_t i) {
    return *reinterpret_cast<const T*>(base + i * (uint64_t)stride);
}

template <class L>
__device__ __forceinline__ void write_result(const OutDev& o, uint64_t i, bool found,
                                             typename L::V v) {
    using V = typename L::V;
    if (o.pairs) {
        char* p = o.pairs + i * 2ull * sizeof(V);
        *reinterpret_cast<uint8_t*>(p) = found ? 1 : 0;
        *reinterpret_cast<V*>(p + sizeof(V)) = v;
    }
    if (o.vals) reinterpret_cast<V*>(o.vals)[i] = v;
    if (o.flags) o.flags[i] = found ? 1 : 0;
}

__device__ __forceinline__ void acquire_agent() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
__device__ __forceinline__ void release_agent() { __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent"); }

__device__ __forceinline__ uint32_t ld_agent(const uint32_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ uint64_t ld_agent(const uint64_t* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave lock primitives on a bucket's 32-bit word (one lane calls them).
// bit 31 = held exclusively, bit 30 = a writer is waiting, bits 0..29 = number
// of readers.  A writer's first attempt issues its compare-and-swap at once
// (an uncontended lock costs one round trip, and a failed CAS changes
// nothing); its retries, and every reader attempt, read the word first and
// only then issue the RMW (test-and-test-and-set): a reader's RMW on a held
// lock would have to be undone.  A writer that fails raises the waiting bit,
// which turns new readers away until it gets in -- so a hot bucket's stream
// of readers cannot starve its writers.
constexpr uint32_t LOCK_WAITING = 0x40000000u;

__device__ __forceinline__ bool try_lock_excl(uint32_t* w, bool first = false) {
    uint32_t v = first ? 0u : ld_coh(w);
    if (v == 0 || v == LOCK_WAITING) {
        if (__hip_atomic_compare_exchange_strong(w, &v, LOCK_WRITER, __ATOMIC_RELAXED,
                                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            return true;
        if (!first) return false;
    }
    if (!(v & LOCK_WAITING))
        __hip_atomic_fetch_or(w, LOCK_WAITING, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return false;
}
__device__ __forceinline__ bool try_lock_shared(uint32_t* w) {
    if (ld_coh(w) & (LOCK_WRITER | LOCK_WAITING)) return false;
    const uint32_t old = __hip_atomic_fetch_add(w, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old & (LOCK_WRITER | LOCK_WAITING)) {
        __hip_atomic_fetch_add(w, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return false;
    }
    return true;
}
__device__ __forceinline__ void unlock_excl(uint32_t* w) {
    __hip_atomic_fetch_and(w, ~LOCK_WRITER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void unlock_shared(uint32_t* w) {
    __hip_atomic_fetch_add(w, 0xffffffffu, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace lslab_dev
