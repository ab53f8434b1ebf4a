// lslab wide-key tables: map / set whose keys are 4..128 bytes (values
// 0..128 bytes), the reference's legacy 128-B-key tests
// (test/slabunified_test.cu.old:344-553: Key of 128 bytes -> int* and
// Key -> Key).
//
// Layout.  A slab is one 128-B line of 32 words: words 0..29 hold the 32-bit
// tags (hash fingerprints) of its 30 key slots, word 30 the next slab of the
// chain (0 = none; slab 0 is a bucket head, never a successor) and word 31 the
// valid bitmap.  The keys and values themselves live in two arenas parallel
// to the slots: slot j of slab s owns key words [(30 s + j) kw, +kw) and value
// words [(30 s + j) vw, +vw).  A slot keeps its arena entry for life, so
// removes and re-inserts need no allocator beyond the slab pool, which works
// as the narrow tables' does (bump pointer, insert into the last slab's lowest
// empty slot, no reuse: detail/traverse.h:54-55,98-124,
// detail/device_allocator.h:62-73).
//
// Execution.  One op per 8-lane group (eight per wavefront): lane l holds
// words l, l+8, l+16, l+24 of the key and of the value.  A probe loads the
// slab line (four words per lane), ORs together the slots whose tag matches,
// and confirms each candidate by comparing the key arena entry four words per
// lane (coalesced reads, with the value read beside it).  Mutations hold the
// bucket's reader/writer lock exclusively (the reference protocol); gets take
// no locks: get calls, and the first pass of a mixed call, which runs the
// call's gets against the table as it was before the call; its second pass
// then runs the mutations.
//
// Hash.  h = mix64(XOR over key words l of mix64((l << 32 | word_l) ^ SEED));
// tag = low 32 bits of h.  Bucket = the key's first 8 bytes (identity, as the
// narrow tables) or h >> 32 (LSLAB_HASH_MIX), masked to 2^n buckets.
#include <hip/hip_runtime.h>
#include <hip/hip_ext.h>

#include <algorithm>
#include <new>

#include "lslab_c.h"
#include "slab.hpp"

namespace lslab_dev {

constexpr int WSLOTS = 30;
constexpr int WNEXT = 30;
constexpr int WVALID = 31;
constexpr uint32_t WMASK = (1u << WSLOTS) - 1u;
constexpr uint64_t WSEED = 0x9E3779B97F4A7C15ull;
constexpr int WG = 8;    // lanes per op (eight ops per wavefront)
constexpr int WPL = 4;   // words of a 32-word line / key / value per lane

struct WideDev {
    uint32_t* slabs;  // (nbuckets + pool) * 32 words
    uint32_t* keys;   // (nbuckets + pool) * 30 * kw words
    uint32_t* vals;   // (nbuckets + pool) * 30 * vw words (null for sets)
    uint32_t* locks;  // one per bucket
    uint64_t* alloc_ctr;
    uint32_t* status;
    uint64_t nbuckets;
    uint64_t pool;
    uint32_t mask;
    int kw, vw;
    int hash_mix;
    int is_map;
};

struct WideOps {
    const uint8_t* ops;  // null: every op is uop
    uint8_t uop;
    const uint32_t* keys;  // n * kw words
    const uint32_t* vals;  // n * vw words (puts), may be null
    uint8_t* found;        // n, may be null
    uint32_t* out_vals;    // n * vw words, may be null
};

struct WideStats {
    unsigned long long present;
    unsigned long long max_chain;
};

__device__ __forceinline__ int wlane() { return (int)(__lane_id() & (WG - 1)); }
// This group's bits of a wave ballot.
__device__ __forceinline__ uint32_t wballot(bool p) {
    return (uint32_t)(__ballot(p) >> (__lane_id() & ~(uint32_t)(WG - 1))) & ((1u << WG) - 1u);
}
// OR of v over the group's lanes.
__device__ __forceinline__ uint32_t grp_or(uint32_t v) {
#pragma unroll
    for (int d = WG / 2; d >= 1; d >>= 1) v |= __shfl_xor(v, d, WG);
    return v;
}

template <bool COH>
__device__ __forceinline__ uint32_t wld(const uint32_t* p) {
    if constexpr (COH) return ld_coh(p);
    return *p;
}

// A group holds a record of up to 32 words strided over its lanes: lane l
// holds words l, l + 8, l + 16, l + 24, so each of the four accesses covers 32
// contiguous bytes (one sector) -- also for agent-scope loads and
// write-through stores, which are issued one word per lane.
__device__ __forceinline__ int word_of(int k) { return WG * k + wlane(); }

// This lane's words of a `cnt`-word record at p (zeros past cnt).
template <bool COH>
__device__ __forceinline__ void ld_words(const uint32_t* p, int cnt, uint32_t (&w)[WPL]) {
#pragma unroll
    for (int k = 0; k < WPL; ++k) w[k] = word_of(k) < cnt ? wld<COH>(p + word_of(k)) : 0u;
}
__device__ __forceinline__ void st_words(uint32_t* p, int cnt, const uint32_t (&w)[WPL]) {
#pragma unroll
    for (int k = 0; k < WPL; ++k)
        if (word_of(k) < cnt) st_coh(p + word_of(k), w[k]);
}

// Finds the key (this lane's words `kw_` of it) in bucket b's chain.  On
// return slot >= 0: the key sits in slab s, slot `slot`, and `vfound` holds
// this lane's words of its value (loaded with the key words, so a hit costs
// no third dependent load); slot < 0: absent, `last` is the chain's last slab
// and `lastvalid` its valid bitmap.
template <bool COH>
__device__ __forceinline__ void wide_find(const WideDev& t, uint64_t b, uint32_t tag,
                                          const uint32_t (&kw_)[WPL], uint64_t& s, int& slot,
                                          uint64_t& last, uint32_t& lastvalid,
                                          uint32_t (&vfound)[WPL], bool& ok) {
    const int gl = wlane();
    s = b;
    slot = -1;
    last = b;
    lastvalid = 0;
    ok = true;
    const uint64_t nslabs = t.nbuckets + t.pool;
    for (uint64_t steps = 0;; ++steps) {
        uint32_t w[WPL];
        ld_words<COH>(t.slabs + s * 32, 32, w);  // word x sits in w[x / WG] of lane x % WG
        const uint32_t next = __shfl(w[WNEXT / WG], WNEXT % WG, WG);
        const uint32_t valid = __shfl(w[WVALID / WG], WVALID % WG, WG);
        uint32_t mine = 0;
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
            const int sl = word_of(k);
            if (sl < WSLOTS && ((valid >> sl) & 1u) && w[k] == tag) mine |= 1u << sl;
        }
        uint32_t cand = grp_or(mine);
        while (cand) {
            const int j = __builtin_ctz(cand);
            cand &= cand - 1u;
            const uint64_t e = s * WSLOTS + (uint64_t)j;
            uint32_t kk[WPL], vv[WPL];
            ld_words<COH>(t.keys + e * (uint64_t)t.kw, t.kw, kk);
            ld_words<COH>(t.vals + e * (uint64_t)t.vw, t.vw, vv);
            bool diff = false;
#pragma unroll
            for (int k = 0; k < WPL; ++k) diff |= kk[k] != kw_[k];
            if (wballot(diff) == 0u) {
                slot = j;
#pragma unroll
                for (int k = 0; k < WPL; ++k) vfound[k] = vv[k];
                break;
            }
        }
        if (slot >= 0) return;
        last = s;
        lastvalid = valid;
        if (next == 0) return;
        if (next < t.nbuckets || next >= nslabs || steps > t.pool) {
            if (gl == 0)
                __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT);
            ok = false;
            return;
        }
        s = next;
    }
}

// One op under its bucket's lock (COH: agent-coherent accesses) or, for a
// read-only call, without it.  Results: found, this lane's words of the old
// value, and ok = false when the op could not complete (pool exhausted,
// damaged chain).
template <bool COH>
__device__ __forceinline__ void wide_apply(const WideDev& t, uint8_t op, uint64_t b, uint32_t tag,
                                           const uint32_t (&kw_)[WPL],
                                           const uint32_t (&vw_)[WPL], bool& found,
                                           uint32_t (&old)[WPL], bool& ok) {
    const int gl = wlane();
    uint64_t s, last;
    int slot;
    uint32_t lastvalid;
    uint32_t vfound[WPL] = {0u, 0u, 0u, 0u};
    wide_find<COH>(t, b, tag, kw_, s, slot, last, lastvalid, vfound, ok);
    found = slot >= 0;
#pragma unroll
    for (int k = 0; k < WPL; ++k) old[k] = found ? vfound[k] : 0u;
    if (!ok) return;
    if (op == OP_PUT) {
        if (found) {
            st_words(t.vals + (s * WSLOTS + (uint64_t)slot) * (uint64_t)t.vw, t.vw, vw_);
        } else {
            const uint32_t empty = ~lastvalid & WMASK;
            uint64_t ts = last;
            int tj = 0;
            bool fresh = false;
            if (empty) {
                tj = __builtin_ctz(empty);
            } else {
                uint64_t ns = 0;
                if (gl == 0)
                    ns = __hip_atomic_fetch_add(t.alloc_ctr, 1ull, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT);
                ns = grp_bcast(ns, 0, WG);
                if (ns >= t.pool) {
                    if (gl == 0)
                        __hip_atomic_fetch_or(t.status, ST_POOL_EXHAUSTED, __ATOMIC_RELAXED,
                                              __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    return;
                }
                ts = t.nbuckets + ns;
                fresh = true;
            }
            const uint64_t e = ts * WSLOTS + (uint64_t)tj;
            st_words(t.keys + e * (uint64_t)t.kw, t.kw, kw_);
            st_words(t.vals + e * (uint64_t)t.vw, t.vw, vw_);
            if (gl == 0) st_coh(t.slabs + ts * 32 + tj, tag);
            if (fresh) {
                if (gl == 0) st_coh(t.slabs + ts * 32 + WVALID, 1u);
                wait_vm();
                if (gl == 0) st_coh(t.slabs + last * 32 + WNEXT, (uint32_t)ts);
            } else {
                wait_vm();
                if (gl == 0)
                    __hip_atomic_fetch_or(t.slabs + ts * 32 + WVALID, 1u << tj, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    } else if (op == OP_REMOVE) {
        if (found && gl == 0)
            __hip_atomic_fetch_and(t.slabs + s * 32 + WVALID, ~(1u << slot), __ATOMIC_RELAXED,
                                   __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Which ops of the stream a launch runs: all, only the gets, or only the
// mutations (the two passes of a mixed call).
enum : int { SEL_ALL = 0, SEL_GETS = 1, SEL_MUTS = 2 };

// Op i of the stream on one 8-lane group (lane l holds words l, l+8, l+16,
// l+24 of the key, the value and a probed slab line); i >= n or an op that
// sel leaves out makes the group idle.  LOCKED: every op holds its bucket's
// lock (shared for gets); otherwise (calls or passes without mutations) no
// locks.  Every group of the wave calls it together (the lock loop is
// wave-uniform).
template <bool LOCKED>
__device__ __forceinline__ void wide_op(const WideDev& t, const WideOps& o, uint64_t i,
                                        uint64_t n, int sel) {
    const int gl = wlane();
    uint8_t op = OP_GET;
    if (i < n) {
        op = o.ops ? o.ops[i] : o.uop;
        if (op > OP_REMOVE) op = OP_GET;
    }
    const bool act = i < n && (sel == SEL_ALL || (sel == SEL_GETS) == (op == OP_GET));
    uint32_t kw_[WPL] = {0u, 0u, 0u, 0u}, vw_[WPL] = {0u, 0u, 0u, 0u};
    if (act) ld_words<false>(o.keys + i * (uint64_t)t.kw, t.kw, kw_);
    if (act && op == OP_PUT && o.vals) ld_words<false>(o.vals + i * (uint64_t)t.vw, t.vw, vw_);
    uint64_t m = 0;
#pragma unroll
    for (int k = 0; k < WPL; ++k) {
        const int w = word_of(k);
        if (w < t.kw) m ^= mix64((((uint64_t)w << 32) | kw_[k]) ^ WSEED);
    }
#pragma unroll
    for (int d = WG / 2; d >= 1; d >>= 1) {
        const uint32_t lo = __shfl_xor((uint32_t)m, d, WG);
        const uint32_t hi = __shfl_xor((uint32_t)(m >> 32), d, WG);
        m ^= ((uint64_t)hi << 32) | lo;
    }
    const uint64_t h = mix64(m);
    const uint32_t tag = (uint32_t)h;
    const uint32_t w0 = __shfl(kw_[0], 0, WG);
    const uint32_t w1 = __shfl(kw_[0], 1, WG);  // 0 when the key is one word
    const uint64_t b = (t.hash_mix ? (h >> 32) : ((uint64_t)w1 << 32 | w0)) & t.mask;

    bool found = false, ok = false;
    uint32_t old[WPL] = {0u, 0u, 0u, 0u};
    if constexpr (!LOCKED) {
        if (act) wide_apply<false>(t, op, b, tag, kw_, vw_, found, old, ok);
    } else {
        const bool shared = op == OP_GET;
        uint32_t* lw = t.locks + b;
        bool done = !act, first = true;
        uint64_t t0 = 0;
        uint32_t backoff = 1;
        // Wave-uniform exit: with a divergent exit the compiler could move the
        // critical section and unlock past the loop, where a group of the same
        // wave spinning on that lock would never see it released.
        for (;;) {
            if (!done) {
                int got = 0;
                if (gl == 0)
                    got = shared ? try_lock_shared(lw) : try_lock_excl(lw, first);
                first = false;
                got = __shfl(got, 0, WG);
                if (got) {
                    wide_apply<true>(t, op, b, tag, kw_, vw_, found, old, ok);
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
                        if (gl == 0)
                            __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT);
                        done = true;
                    }
                }
            }
            if (__ballot(!done) == 0) break;
            for (uint32_t k = 0; k < backoff; ++k) __builtin_amdgcn_s_sleep(4);
            backoff = backoff < 64 ? backoff * 2 : 64;
        }
    }
    if (!act) return;
    // a set insert reports true for present and new keys alike (not when it failed)
    const bool rf = (!t.is_map && op == OP_PUT) ? ok : found;
    if (o.found && gl == 0) o.found[i] = rf ? 1 : 0;
    if (o.out_vals) {
#pragma unroll
        for (int k = 0; k < WPL; ++k)
            if (word_of(k) < t.vw) o.out_vals[i * (uint64_t)t.vw + word_of(k)] = old[k];
    }
}

// One op per group.
template <bool LOCKED>
__global__ __launch_bounds__(256) void k_wide(WideDev t, WideOps o, uint64_t n, int sel) {
    wide_op<LOCKED>(t, o, (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WG, n, sel);
}

// The mutation pass of a mixed call: each wave reads the opcodes of 64
// consecutive ops, and its 8 groups take the mutations among them in rounds
// of 8, so a stream of mostly gets does not start one idle group per get.
__global__ __launch_bounds__(256) void k_wide_muts(WideDev t, WideOps o, uint64_t n) {
    const uint64_t base = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / 64u * 64u;
    const int lane = (int)__lane_id();
    uint8_t op = OP_GET;
    if (base + lane < n) op = o.ops[base + lane];
    uint64_t mask = __ballot(op == OP_PUT || op == OP_REMOVE);  // wave-uniform
    const int grp = lane / WG;
    while (mask) {
        uint64_t mine = mask;
        for (int k = 0; k < grp && mine; ++k) mine &= mine - 1u;
        wide_op<true>(t, o, mine ? base + (uint64_t)__builtin_ctzll(mine) : n, n, SEL_MUTS);
        for (int k = 0; k < 64 / WG && mask; ++k) mask &= mask - 1u;
    }
}

// Present keys and longest chain: one thread per bucket.
__global__ __launch_bounds__(256) void k_wide_stats(WideDev t, WideStats* st) {
    const uint64_t b = blockIdx.x * (uint64_t)blockDim.x + threadIdx.x;
    if (b >= t.nbuckets) return;
    uint64_t s = b, present = 0, len = 0;
    const uint64_t nslabs = t.nbuckets + t.pool;
    for (;;) {
        ++len;
        present += __builtin_popcount(t.slabs[s * 32 + WVALID] & WMASK);
        const uint32_t next = t.slabs[s * 32 + WNEXT];
        if (next == 0 || next < t.nbuckets || next >= nslabs || len > t.pool) break;
        s = next;
    }
    atomicAdd(&st->present, (unsigned long long)present);
    atomicMax(&st->max_chain, (unsigned long long)len);
}

// Copies every present key (and value) out: one 8-lane group per slab.
__global__ __launch_bounds__(256) void k_wide_dump(WideDev t, uint64_t nslabs, uint32_t* keys,
                                                   uint32_t* vals, uint64_t cap,
                                                   unsigned long long* counter) {
    const int gl = wlane();
    const uint64_t s = (blockIdx.x * (uint64_t)blockDim.x + threadIdx.x) / WG;
    if (s >= nslabs) return;
    uint32_t valid = t.slabs[s * 32 + WVALID] & WMASK;
    while (valid) {
        const int j = __builtin_ctz(valid);
        valid &= valid - 1u;
        unsigned long long idx = 0;
        if (gl == 0) idx = atomicAdd(counter, 1ull);
        idx = grp_bcast((uint64_t)idx, 0, WG);
        if (idx >= cap) continue;
        const uint64_t e = s * WSLOTS + (uint64_t)j;
#pragma unroll
        for (int k = 0; k < WPL; ++k) {
            const int w = word_of(k);
            if (keys && w < t.kw) keys[idx * t.kw + w] = t.keys[e * t.kw + w];
            if (vals && w < t.vw) vals[idx * t.vw + w] = t.vals[e * t.vw + w];
        }
    }
}

}  // namespace lslab_dev

using namespace lslab_dev;

struct lslab_wide {
    WideDev dev;
    uint64_t* state;  // [0] alloc counter, [1] status, [2..3] stats, [4] dump counter
    hipEvent_t order_ev;
    int key_bytes, value_bytes;
    int device;
    uint64_t device_bytes;
};

namespace {

#define WCHK(x)                                     \
    do {                                            \
        if ((x) != hipSuccess) return LSLAB_ERR_HIP; \
    } while (0)

void free_wide(lslab_wide* w) {
    (void)hipFree(w->dev.slabs);
    (void)hipFree(w->dev.keys);
    (void)hipFree(w->dev.vals);
    (void)hipFree(w->dev.locks);
    (void)hipFree(w->state);
    if (w->order_ev) (void)hipEventDestroy(w->order_ev);
    delete w;
}

bool capturing(hipStream_t s) {
    hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &st) == hipSuccess && st != hipStreamCaptureStatusNone;
}

// Key and value records are read and written as 32-bit words.
bool aligned4(const void* p) { return ((uintptr_t)p & 3u) == 0; }

// locked: the call has mutations.  A mixed call (per-op opcodes) runs in two
// passes: its gets lock-free against the table as it was before the call,
// then its mutations under the bucket locks.  Per key the results are those
// of the serial order [gets][mutations in lock order], which a batch allows;
// a batch without same-key conflicts gets exactly the sequential replay's.
int run(lslab_wide* w, const WideOps& o, uint64_t n, bool locked, hipStream_t s) {
    if (n == 0) return LSLAB_OK;
    if (n > (1ull << 31) || !aligned4(o.keys) || !aligned4(o.vals) || !aligned4(o.out_vals))
        return LSLAB_ERR_INVALID;
    // The call's last kernel launch records the order event on completion
    // (no separate marker in the stream); none while s is being captured.
    hipEvent_t done = nullptr;
    if (!capturing(s)) {
        WCHK(hipStreamWaitEvent(s, w->order_ev, 0));
        done = w->order_ev;
    }
    const uint64_t blocks = (n * WG + 255) / 256;
    if (locked && o.ops) {
        hipLaunchKernelGGL(k_wide<false>, dim3((unsigned)blocks), dim3(256), 0, s, w->dev, o, n,
                           (int)SEL_GETS);
        hipExtLaunchKernelGGL(k_wide_muts, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                              nullptr, done, 0, w->dev, o, n);
    } else if (locked) {
        hipExtLaunchKernelGGL(k_wide<true>, dim3((unsigned)blocks), dim3(256), 0, s, nullptr, done,
                              0, w->dev, o, n, (int)SEL_ALL);
    } else {
        hipExtLaunchKernelGGL(k_wide<false>, dim3((unsigned)blocks), dim3(256), 0, s, nullptr,
                              done, 0, w->dev, o, n, (int)SEL_ALL);
    }
    WCHK(hipGetLastError());
    return LSLAB_OK;
}

WideOps make(const uint8_t* ops, uint8_t uop, const void* keys, const void* vals, uint8_t* found,
             void* out_vals) {
    WideOps o;
    o.ops = ops;
    o.uop = uop;
    o.keys = static_cast<const uint32_t*>(keys);
    o.vals = static_cast<const uint32_t*>(vals);
    o.found = found;
    o.out_vals = static_cast<uint32_t*>(out_vals);
    return o;
}

}  // namespace

extern "C" uint64_t lslab_wide_slab_count(unsigned n_log2, uint64_t pool_slabs) {
    if (n_log2 > 30) return 0;
    const uint64_t ns = (1ull << n_log2) + pool_slabs;
    return ns >= (1ull << 32) ? 0 : ns;
}

extern "C" int lslab_wide_create(lslab_wide** out, int key_bytes, int value_bytes,
                                 unsigned n_log2, uint64_t pool_slabs, int hash) {
    if (!out) return LSLAB_ERR_INVALID;
    *out = nullptr;
    if (key_bytes < 4 || key_bytes > 128 || key_bytes % 4 || value_bytes < 0 ||
        value_bytes > 128 || value_bytes % 4 || n_log2 > 30 ||
        (hash != 0 && hash != LSLAB_HASH_IDENTITY && hash != LSLAB_HASH_MIX))
        return LSLAB_ERR_INVALID;
    const uint64_t nslabs = lslab_wide_slab_count(n_log2, pool_slabs);
    if (nslabs == 0) return LSLAB_ERR_INVALID;
    lslab_wide* w = new (std::nothrow) lslab_wide{};
    if (!w) return LSLAB_ERR_OOM;
    w->key_bytes = key_bytes;
    w->value_bytes = value_bytes;
    WideDev& d = w->dev;
    d.nbuckets = 1ull << n_log2;
    d.pool = pool_slabs;
    d.mask = (uint32_t)(d.nbuckets - 1);
    d.kw = key_bytes / 4;
    d.vw = value_bytes / 4;
    d.hash_mix = hash == LSLAB_HASH_MIX;
    d.is_map = value_bytes > 0;
    const uint64_t sb = nslabs * 128, kb = nslabs * WSLOTS * (uint64_t)key_bytes,
                   vb = nslabs * WSLOTS * (uint64_t)value_bytes, lb = d.nbuckets * 4;
    if (hipGetDevice(&w->device) != hipSuccess ||
        hipMalloc(&d.slabs, sb) != hipSuccess || hipMalloc(&d.keys, kb) != hipSuccess ||
        (vb && hipMalloc(&d.vals, vb) != hipSuccess) || hipMalloc(&d.locks, lb) != hipSuccess ||
        hipMalloc(&w->state, 8 * sizeof(uint64_t)) != hipSuccess) {
        free_wide(w);
        return LSLAB_ERR_OOM;
    }
    if (hipMemset(d.slabs, 0, sb) != hipSuccess || hipMemset(d.locks, 0, lb) != hipSuccess ||
        hipMemset(w->state, 0, 8 * sizeof(uint64_t)) != hipSuccess ||
        hipEventCreateWithFlags(&w->order_ev, hipEventDisableTiming) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        free_wide(w);
        return LSLAB_ERR_HIP;
    }
    d.alloc_ctr = w->state;
    d.status = reinterpret_cast<uint32_t*>(w->state + 1);
    w->device_bytes = sb + kb + vb + lb + 8 * sizeof(uint64_t);
    *out = w;
    return LSLAB_OK;
}

extern "C" int lslab_wide_destroy(lslab_wide* w) {
    if (!w) return LSLAB_ERR_INVALID;
    (void)hipDeviceSynchronize();
    free_wide(w);
    return LSLAB_OK;
}

extern "C" int lslab_wide_put(lslab_wide* w, const void* keys, const void* values,
                              uint8_t* out_found, void* out_old, size_t n, hipStream_t stream) {
    if (!w || (n && (!keys || (w->dev.is_map && !values)))) return LSLAB_ERR_INVALID;
    return run(w, make(nullptr, OP_PUT, keys, values, out_found, out_old), n, true, stream);
}

extern "C" int lslab_wide_get(lslab_wide* w, const void* keys, uint8_t* out_found,
                              void* out_values, size_t n, hipStream_t stream) {
    if (!w || (n && !keys)) return LSLAB_ERR_INVALID;
    return run(w, make(nullptr, OP_GET, keys, nullptr, out_found, out_values), n, false, stream);
}

extern "C" int lslab_wide_remove(lslab_wide* w, const void* keys, uint8_t* out_found,
                                 void* out_old, size_t n, hipStream_t stream) {
    if (!w || (n && !keys)) return LSLAB_ERR_INVALID;
    return run(w, make(nullptr, OP_REMOVE, keys, nullptr, out_found, out_old), n, true, stream);
}

extern "C" int lslab_wide_mixed(lslab_wide* w, const uint8_t* ops, const void* keys,
                                const void* values, uint8_t* out_found, void* out_values, size_t n,
                                hipStream_t stream) {
    if (!w || (n && (!ops || !keys || (w->dev.is_map && !values)))) return LSLAB_ERR_INVALID;
    return run(w, make(ops, 0, keys, values, out_found, out_values), n, true, stream);
}

extern "C" int lslab_wide_status(lslab_wide* w, hipStream_t stream) {
    if (!w) return LSLAB_ERR_INVALID;
    WCHK(hipStreamWaitEvent(stream, w->order_ev, 0));
    WCHK(hipStreamSynchronize(stream));
    uint32_t st = 0;
    WCHK(hipMemcpy(&st, w->dev.status, sizeof(st), hipMemcpyDeviceToHost));
    if (st) WCHK(hipMemset(w->dev.status, 0, sizeof(uint32_t)));
    if (st & ST_CORRUPT) return LSLAB_ERR_CORRUPT;
    if (st & ST_TIMEOUT) return LSLAB_ERR_TIMEOUT;
    if (st & ST_POOL_EXHAUSTED) return LSLAB_ERR_POOL_EXHAUSTED;
    return LSLAB_OK;
}

extern "C" int lslab_wide_stats(lslab_wide* w, lslab_stats* out, hipStream_t stream) {
    if (!w || !out) return LSLAB_ERR_INVALID;
    WCHK(hipStreamWaitEvent(stream, w->order_ev, 0));
    WideStats* st = reinterpret_cast<WideStats*>(w->state + 2);
    WCHK(hipMemsetAsync(st, 0, sizeof(WideStats), stream));
    hipLaunchKernelGGL(k_wide_stats, dim3((unsigned)((w->dev.nbuckets + 255) / 256)), dim3(256), 0,
                       stream, w->dev, st);
    WCHK(hipGetLastError());
    uint64_t h[3];
    WCHK(hipMemcpyAsync(h, w->state, 3 * sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    WCHK(hipStreamSynchronize(stream));
    WideStats hs;
    WCHK(hipMemcpy(&hs, st, sizeof(hs), hipMemcpyDeviceToHost));
    out->buckets = w->dev.nbuckets;
    out->pool_slabs = w->dev.pool;
    out->slabs_used = std::min<uint64_t>(h[0], w->dev.pool);
    out->present = hs.present;
    out->max_chain = hs.max_chain;
    out->slab_bytes = 128 + WSLOTS * (uint64_t)(w->key_bytes + w->value_bytes);
    out->device_bytes = w->device_bytes;
    WCHK(hipEventRecord(w->order_ev, stream));
    return LSLAB_OK;
}

extern "C" int lslab_wide_dump(lslab_wide* w, void* keys_out, void* vals_out, uint64_t cap,
                               uint64_t* count, hipStream_t stream) {
    if (!w || !count || (cap && !keys_out)) return LSLAB_ERR_INVALID;
    WCHK(hipStreamWaitEvent(stream, w->order_ev, 0));
    uint64_t used = 0;
    WCHK(hipMemcpyAsync(&used, w->state, sizeof(used), hipMemcpyDeviceToHost, stream));
    WCHK(hipStreamSynchronize(stream));
    const uint64_t nslabs = w->dev.nbuckets + std::min<uint64_t>(used, w->dev.pool);
    unsigned long long* ctr = reinterpret_cast<unsigned long long*>(w->state + 4);
    WCHK(hipMemsetAsync(ctr, 0, sizeof(uint64_t), stream));
    hipLaunchKernelGGL(k_wide_dump, dim3((unsigned)((nslabs * WG + 255) / 256)), dim3(256), 0,
                       stream, w->dev, nslabs, static_cast<uint32_t*>(keys_out),
                       static_cast<uint32_t*>(vals_out), cap, ctr);
    WCHK(hipGetLastError());
    WCHK(hipMemcpyAsync(count, ctr, sizeof(uint64_t), hipMemcpyDeviceToHost, stream));
    WCHK(hipStreamSynchronize(stream));
    WCHK(hipEventRecord(w->order_ev, stream));
    return LSLAB_OK;
}
