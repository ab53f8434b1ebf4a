// lslab: HIP kernels for gfx950 and the C ABI declared in include/lslab_c.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <new>

#include "../../include/lslab_c.h"
#include "engine.hpp"

using namespace lslab_dev;

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// Grid-wide barrier state for the persistent batched kernel.
// Persistent kernels run one block of 1024 threads (16 wavefronts) per CU:
// 16 waves per CU keep random slab reads at full rate, and one block per CU
// keeps grid barriers cheap (256 arrivals, one cache invalidate per CU).
constexpr int PERSIST_THREADS = 1024;

// Grid-wide barrier state for the persistent kernels.  Blocks arrive on one
// of BAR_GROUPS group counters (block id mod BAR_GROUPS), the last block of a
// group arrives on the top counter, and the last group releases every group
// by bumping the group's generation word.  Each word sits on its own 128-B
// line, so arrivals and polls of different groups do not share a line.
constexpr uint32_t BAR_GROUPS = 16;
struct alignas(128) BarWord {
    uint32_t count;
    uint32_t gen;
    uint32_t pad[30];
};
struct BarState {
    BarWord top;
    BarWord group[BAR_GROUPS];
    uint32_t abort;
    uint32_t pad[31];
};

// Plain per-batch kernel: one G-lane group per op, grid-stride.
template <class L, int G, bool LOCKED>
__global__ __launch_bounds__(256) void k_plain(TableDev t, OpsDev o, OutDev out, uint64_t n) {
    using K = typename L::K;
    using V = typename L::V;
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    for (uint64_t i = gid; i < n; i += ng) {
        uint8_t op;
        K key;
        V val;
        load_op<L>(o, i, op, key, val);
        bool rf, ok;
        V rv;
        run_op<L, G, LOCKED>(t, op, key, val, rf, rv, ok);
        if (Grp<G>::lane() == 0) write_result<L>(out, i, rf, rv);
    }
}

template <class L, int G, bool BC>
__global__ __launch_bounds__(256) void k_phase_r(TableDev t, OpsDev o, OutDev out, uint64_t lo,
                                                 uint64_t m) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if constexpr (BC) comb_init<256>();
    phase_r_range<L, G, 256, BC>(t, o, out, lo, m, gid, ng);
}

template <class L, int G>
__global__ __launch_bounds__(256) void k_phase_w(TableDev t, OpsDev o, uint64_t lo, uint64_t m) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    for (uint64_t li = gid; li < m; li += ng) phase_w_op<L, G>(t, o, lo, (uint32_t)li);
}

// Grid barrier for a grid that is fully resident.  Every wave waits for its
// own writes to complete (shared table writes are agent-scope stores or
// atomics, so nothing else is needed to publish them), one lane per block
// arrives, the last arrival releases everyone, and every wave acquires (drops
// stale cached lines) before the next phase reads the table with plain
// loads.  No cache-wide write-back is issued.  Spins are bounded; on timeout
// or abort the barrier returns false and the caller leaves the kernel.
__device__ __forceinline__ bool grid_sync(BarState* bar, uint32_t nblocks) {
    wait_vm();
    __syncthreads();
    __shared__ uint32_t ok_s;
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        const uint32_t x = blockIdx.x % BAR_GROUPS;
        const uint32_t ngroups = nblocks < BAR_GROUPS ? nblocks : BAR_GROUPS;
        const uint32_t members = (nblocks - x + BAR_GROUPS - 1) / BAR_GROUPS;
        BarWord* gw = &bar->group[x];
        const uint32_t g = ld_agent(&gw->gen);
        const uint32_t arrived =
            __hip_atomic_fetch_add(&gw->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        if (arrived == members - 1u) {
            st_coh(&gw->count, 0u);
            wait_vm();
            const uint32_t top = __hip_atomic_fetch_add(&bar->top.count, 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if (top == ngroups - 1u) {
                st_coh(&bar->top.count, 0u);
                wait_vm();
                for (uint32_t y = 0; y < ngroups; ++y)
                    __hip_atomic_fetch_add(&bar->group[y].gen, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        const uint64_t t0 = now_ticks();
        while (ld_agent(&gw->gen) == g) {
            if (ld_agent(&bar->abort) || now_ticks() - t0 > SPIN_LIMIT_TICKS) {
                st_coh(&bar->abort, 1u);
                ok = 0;
                break;
            }
            __builtin_amdgcn_s_sleep(2);
        }
        ok_s = ok;
    }
    // One acquire per block: wave 0 invalidates the CU's L1 and the XCD's
    // L2 and waits for it; the block barrier then releases the other waves.
    // (An acquire in every wave costs 4-8x more per barrier: each is a
    // cache-wide invalidate.)
    if (threadIdx.x < 64) {
        acquire_agent();
        wait_vm();
    }
    __syncthreads();
    return ok_s != 0;
}

// Persistent batched kernel: batches in order, each as two-phase sub-batches
// separated by grid barriers.
template <class L, int G, bool BC>
__global__ __launch_bounds__(PERSIST_THREADS) void k_batches(TableDev t, OpsDev o, OutDev out,
                                                 const uint64_t* offsets, uint64_t nbatches,
                                                 uint64_t submax, BarState* bar) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    if constexpr (BC) comb_init<PERSIST_THREADS>();
    for (uint64_t b = 0; b < nbatches; ++b) {
        const uint64_t lo = offsets[b], hi = offsets[b + 1];
        for (uint64_t slo = lo; slo < hi; slo += submax) {
            const uint64_t m = min(submax, hi - slo);
            phase_r_range<L, G, PERSIST_THREADS, BC>(t, o, out, slo, m, gid, ng);
            if (!grid_sync(bar, gridDim.x)) {
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
            for (uint64_t li = gid; li < m; li += ng) phase_w_op<L, G>(t, o, slo, (uint32_t)li);
            if (!grid_sync(bar, gridDim.x)) {
                if (threadIdx.x == 0)
                    __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                return;
            }
        }
    }
}

__device__ __forceinline__ void reset_bar(BarState* bar) {
    if (threadIdx.x <= BAR_GROUPS && blockIdx.x == 0) {
        BarWord* w = threadIdx.x == 0 ? &bar->top : &bar->group[threadIdx.x - 1];
        w->count = 0;
        if (threadIdx.x == 0) bar->abort = 0;
    }
}

__global__ void k_reset_bar(BarState* bar) { reset_bar(bar); }

// ---------------------------------------------------------------------------
// Served form: a device-resident request loop.
//
// Requests r = 0..nreq-1 own the contiguous op ranges [off[r], off[r+1]) and
// arrive at arrival[r] (ticks of the 100 MHz constant clock after the loop
// starts, nondecreasing).  Each tick, one scheduler lane admits, in arrival
// order, every request that has arrived, up to max_tick_ops ops and
// max_tick_reqs requests, provided the slab pool keeps headroom for it: a
// request reserves one slab per put, the most its puts can allocate (an
// insert allocates at most one slab).  A request whose reservation can never
// fit (the pool only shrinks) is rejected.  All blocks then run the admitted
// ops as one two-phase batch; the scheduler stamps each request's completion
// time.  Requests arriving while a tick runs join the next tick.
// ---------------------------------------------------------------------------

struct ServeDev {
    uint64_t nreq;
    const uint64_t* off;
    const uint64_t* arrival;
    const uint32_t* reserve;
    uint64_t* complete;
    uint32_t* tick_of;
    uint64_t max_tick_ops;
    uint32_t max_tick_reqs;
};

struct ServeState {
    uint64_t t0;
    uint64_t next_req;
    uint64_t r_lo, r_hi;
    uint64_t ticks;
    uint64_t rejected;
    uint64_t served_ops;
    uint64_t end;
    uint32_t done;
    uint32_t pad;
};

constexpr uint32_t SERVE_REJECTED = 0xffffffffu;

__device__ void serve_schedule(const TableDev& t, const ServeDev& sv, ServeState* ss) {
    uint64_t r = ss->next_req;
    const uint64_t used = ld_agent(t.alloc_ctr);
    uint64_t reserved = 0, ops = 0, n = 0, lo = r;
    while (r < sv.nreq) {
        const uint64_t el = now_ticks() - ss->t0;
        if (sv.arrival[r] > el) {
            if (n > 0) break;
            // Idle: wait for the next arrival (the gap is bounded by the caller).
            const uint64_t start = now_ticks();
            while (now_ticks() - ss->t0 < sv.arrival[r] && now_ticks() - start < SPIN_LIMIT_TICKS / 2)
                __builtin_amdgcn_s_sleep(2);
            continue;
        }
        const uint64_t sz = sv.off[r + 1] - sv.off[r];
        if (n > 0 && (ops + sz > sv.max_tick_ops || n >= sv.max_tick_reqs)) break;
        const uint64_t need = sv.reserve ? sv.reserve[r] : sz;
        if (used + reserved + need > t.pool) {
            if (n > 0) break;
            sv.tick_of[r] = SERVE_REJECTED;
            sv.complete[r] = now_ticks() - ss->t0;
            ss->rejected++;
            r++;
            lo = r;
            continue;
        }
        reserved += need;
        ops += sz;
        n++;
        r++;
    }
    // Read by every block after the barrier: agent-scope stores.
    st_coh(&ss->r_lo, lo);
    st_coh(&ss->r_hi, r);
    ss->next_req = r;
    ss->served_ops += ops;
    if (n == 0) st_coh(&ss->done, 1u);
}

template <class L, int G, bool BC>
__global__ __launch_bounds__(PERSIST_THREADS) void k_serve(TableDev t, OpsDev o, OutDev out, ServeDev sv,
                                               ServeState* ss, BarState* bar, uint64_t submax) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    __shared__ uint64_t stamp_s;
    if constexpr (BC) comb_init<PERSIST_THREADS>();
    if (blockIdx.x == 0 && threadIdx.x == 0) ss->t0 = now_ticks();
    for (uint64_t tick = 0;; ++tick) {
        if (blockIdx.x == 0 && threadIdx.x == 0) serve_schedule(t, sv, ss);
        if (!grid_sync(bar, gridDim.x)) goto fail;
        {
            if (ld_agent(&ss->done)) break;
            const uint64_t r_lo = ld_coh(&ss->r_lo), r_hi = ld_coh(&ss->r_hi);
            const uint64_t lo = sv.off[r_lo], hi = sv.off[r_hi];
            for (uint64_t slo = lo; slo < hi; slo += submax) {
                const uint64_t m = min(submax, hi - slo);
                phase_r_range<L, G, PERSIST_THREADS, BC>(t, o, out, slo, m, gid, ng);
                if (!grid_sync(bar, gridDim.x)) goto fail;
                for (uint64_t li = gid; li < m; li += ng) phase_w_op<L, G>(t, o, slo, (uint32_t)li);
                if (!grid_sync(bar, gridDim.x)) goto fail;
            }
            if (blockIdx.x == 0) {
                if (threadIdx.x == 0) stamp_s = now_ticks() - ss->t0;
                __syncthreads();
                for (uint64_t r = r_lo + threadIdx.x; r < r_hi; r += blockDim.x) {
                    sv.complete[r] = stamp_s;
                    sv.tick_of[r] = (uint32_t)tick;
                }
                __syncthreads();
                if (threadIdx.x == 0) ss->ticks = tick + 1;
            }
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ss->end = now_ticks() - ss->t0;
    return;
fail:
    if (threadIdx.x == 0)
        __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__global__ void k_reset_serve(ServeState* ss, BarState* bar) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        ServeState z{};
        *ss = z;
    }
    reset_bar(bar);
}

// Introspection: one thread per bucket walks its chain.
struct StatsDev {
    unsigned long long present;
    unsigned long long max_chain;
};

template <class L>
__global__ void k_stats(TableDev t, StatsDev* st) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned long long present = 0, len = 0;
    if (b < t.nbuckets) {
        uint64_t s = b;
        for (;;) {
            present += __builtin_popcount(*valid_word<L>(t, s) & L::CAPMASK);
            len++;
            uint32_t nx = *next_word<L>(t, s);
            if (!nx) break;
            if (nx < t.nbuckets || nx >= t.nbuckets + t.pool || len > t.pool + 1) {
                __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            s = nx;
        }
    }
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        present += __shfl_down(present, off, 64);
        len = max(len, (unsigned long long)__shfl_down(len, off, 64));
    }
    if (__lane_id() == 0) {
        atomicAdd(&st->present, present);
        atomicMax(&st->max_chain, len);
    }
}

template <class L>
__global__ void k_dump(TableDev t, typename L::K* keys, typename L::V* vals, uint64_t cap,
                       unsigned long long* counter) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= t.nbuckets) return;
    uint64_t s = b, hops = 0;
    for (;;) {
        const uint32_t v = *valid_word<L>(t, s) & L::CAPMASK;
        for (int i = 0; i < L::CAP; ++i) {
            if ((v >> i) & 1u) {
                unsigned long long idx = atomicAdd(counter, 1ull);
                if (idx < cap) {
                    keys[idx] = *key_slot<L>(t, s, i);
                    if constexpr (L::IS_MAP) {
                        if (vals) vals[idx] = *val_slot<L>(t, s, i);
                    }
                }
            }
        }
        const uint32_t nx = *next_word<L>(t, s);
        if (!nx) break;
        if (nx < t.nbuckets || nx >= t.nbuckets + t.pool || ++hops > t.pool) {
            __hip_atomic_fetch_or(t.status, ST_CORRUPT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            break;
        }
        s = nx;
    }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------

struct DevState {
    uint64_t alloc_ctr;
    uint32_t status;
    uint32_t pad;
    BarState bar;
    StatsDev stats;
    unsigned long long dump_counter;
    ServeState serve;
};

enum TableType { T_MAP_32_32, T_MAP_32_64, T_MAP_64_32, T_MAP_64_64, T_SET_32, T_SET_64 };

struct lslab_table {
    TableType type;
    bool is_map;
    int key_bytes, val_bytes;
    unsigned n_log2;
    uint64_t nbuckets, pool, slab_bytes, device_bytes;
    lslab_tuning tun;
    TableDev dev;
    DevState* state;
    uint64_t gt_cap, submax;
    hipEvent_t order_ev;
    int device;
    int num_cus;
};

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return LSLAB_ERR_HIP; \
    } while (0)

using L_M3232 = Layout<uint32_t, uint32_t, true>;
using L_M3264 = Layout<uint32_t, uint64_t, true>;
using L_M6432 = Layout<uint64_t, uint32_t, true>;
using L_M6464 = Layout<uint64_t, uint64_t, true>;
using L_S32 = Layout<uint32_t, uint32_t, false>;
using L_S64 = Layout<uint64_t, uint32_t, false>;

template <typename F>
static int dispatch(const lslab_table* t, F&& f) {
    switch (t->type) {
        case T_MAP_32_32: return f(L_M3232{});
        case T_MAP_32_64: return f(L_M3264{});
        case T_MAP_64_32: return f(L_M6432{});
        case T_MAP_64_64: return f(L_M6464{});
        case T_SET_32: return f(L_S32{});
        case T_SET_64: return f(L_S64{});
    }
    return LSLAB_ERR_INVALID;
}

template <typename F>
static int dispatch_g(int g, F&& f) {
    switch (g) {
        case 1: return f(std::integral_constant<int, 1>{});
        case 2: return f(std::integral_constant<int, 2>{});
        case 4: return f(std::integral_constant<int, 4>{});
        case 8: return f(std::integral_constant<int, 8>{});
        case 16: return f(std::integral_constant<int, 16>{});
        case 32: return f(std::integral_constant<int, 32>{});
    }
    return LSLAB_ERR_INVALID;
}

// Combining scope as a compile-time flag: true = block, false = wavefront.
template <class F>
static int dispatch_c(int combine, F&& f) {
    if (combine == LSLAB_COMBINE_BLOCK) return f(std::true_type{});
    return f(std::false_type{});
}

static void normalise(lslab_tuning& tu) {
    if (tu.lanes_per_op == 0) tu.lanes_per_op = 4;
    if (tu.block_threads == 0) tu.block_threads = 256;
    if (tu.max_subbatch == 0) tu.max_subbatch = 1ull << 20;
    if (tu.combine == 0) tu.combine = LSLAB_COMBINE_BLOCK;
}

static bool tuning_ok(const lslab_tuning& tu) {
    const int g = tu.lanes_per_op;
    if (g != 1 && g != 2 && g != 4 && g != 8 && g != 16 && g != 32) return false;
    if (tu.engine != LSLAB_ENGINE_PLAIN && tu.engine != LSLAB_ENGINE_TWO_PHASE) return false;
    if (tu.block_threads < 64 || tu.block_threads > 256 || (tu.block_threads % 64)) return false;
    if (tu.max_blocks < 0) return false;
    if (tu.max_subbatch < 64 || tu.max_subbatch > (1ull << 30)) return false;
    if (tu.combine != LSLAB_COMBINE_WAVE && tu.combine != LSLAB_COMBINE_BLOCK) return false;
    return true;
}

static uint64_t grid_for(const lslab_table* t, uint64_t n) {
    const uint64_t groups_per_block = (uint64_t)t->tun.block_threads / t->tun.lanes_per_op;
    uint64_t blocks = (n + groups_per_block - 1) / groups_per_block;
    if (blocks == 0) blocks = 1;
    if (t->tun.max_blocks > 0) blocks = std::min<uint64_t>(blocks, (uint64_t)t->tun.max_blocks);
    return std::min<uint64_t>(blocks, 1ull << 30);
}

static void free_table(lslab_table* t) {
    if (!t) return;
    (void)hipFree(t->dev.slabs);
    (void)hipFree(t->dev.locks);
    (void)hipFree(t->state);
    (void)hipFree(t->dev.gt_key);
    (void)hipFree(t->dev.gt_head);
    (void)hipFree(t->dev.op_gidx);
    (void)hipFree(t->dev.op_loc);
    if (t->order_ev) (void)hipEventDestroy(t->order_ev);
    delete t;
}

static int alloc_scratch(lslab_table* t) {
    uint64_t cap = 1;
    while (cap < 2 * t->tun.max_subbatch) cap <<= 1;
    (void)hipFree(t->dev.gt_key);
    (void)hipFree(t->dev.gt_head);
    (void)hipFree(t->dev.op_gidx);
    (void)hipFree(t->dev.op_loc);
    t->dev.gt_key = nullptr;
    t->dev.gt_head = nullptr;
    t->dev.op_gidx = nullptr;
    t->dev.op_loc = nullptr;
    if (hipMalloc(&t->dev.gt_key, (cap + 1) * sizeof(uint64_t)) != hipSuccess ||
        hipMalloc(&t->dev.gt_head, (cap + 1) * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&t->dev.op_gidx, t->tun.max_subbatch * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&t->dev.op_loc, t->tun.max_subbatch * sizeof(uint64_t)) != hipSuccess)
        return LSLAB_ERR_OOM;
    HIPCHK(hipMemset(t->dev.gt_key, 0xff, (cap + 1) * sizeof(uint64_t)));
    HIPCHK(hipMemset(t->dev.gt_head, 0, (cap + 1) * sizeof(uint32_t)));
    t->gt_cap = cap;
    t->submax = t->tun.max_subbatch;
    t->dev.gt_mask = (uint32_t)(cap - 1);
    return LSLAB_OK;
}

static int create_common(lslab_table** out, TableType type, bool is_map, int kb, int vb,
                         unsigned n_log2, uint64_t pool, const lslab_tuning* tuning) {
    if (!out) return LSLAB_ERR_INVALID;
    *out = nullptr;
    if (n_log2 > 30) return LSLAB_ERR_INVALID;
    lslab_tuning tu{};
    if (tuning) tu = *tuning;
    normalise(tu);
    if (!tuning_ok(tu)) return LSLAB_ERR_INVALID;
    lslab_table* t = new (std::nothrow) lslab_table();
    if (!t) return LSLAB_ERR_OOM;
    t->type = type;
    t->is_map = is_map;
    t->key_bytes = kb;
    t->val_bytes = is_map ? vb : 0;
    t->n_log2 = n_log2;
    t->nbuckets = 1ull << n_log2;
    t->pool = pool;
    t->tun = tu;
    dispatch(t, [&](auto l) {
        t->slab_bytes = decltype(l)::SLAB_BYTES;
        return 0;
    });
    const uint64_t nslabs = t->nbuckets + pool;
    if (nslabs >= (1ull << 32)) {
        delete t;
        return LSLAB_ERR_INVALID;
    }
    if (hipGetDevice(&t->device) != hipSuccess ||
        hipDeviceGetAttribute(&t->num_cus, hipDeviceAttributeMultiprocessorCount, t->device) !=
            hipSuccess) {
        delete t;
        return LSLAB_ERR_HIP;
    }
    int rc = LSLAB_OK;
    const uint64_t slab_total = nslabs * t->slab_bytes;
    if (hipMalloc(&t->dev.slabs, slab_total) != hipSuccess ||
        hipMalloc(&t->dev.locks, t->nbuckets * sizeof(uint32_t)) != hipSuccess ||
        hipMalloc(&t->state, sizeof(DevState)) != hipSuccess) {
        free_table(t);
        return LSLAB_ERR_OOM;
    }
    if (hipMemset(t->dev.slabs, 0, slab_total) != hipSuccess ||
        hipMemset(t->dev.locks, 0, t->nbuckets * sizeof(uint32_t)) != hipSuccess ||
        hipMemset(t->state, 0, sizeof(DevState)) != hipSuccess) {
        free_table(t);
        return LSLAB_ERR_HIP;
    }
    rc = alloc_scratch(t);
    if (rc != LSLAB_OK) {
        free_table(t);
        return rc;
    }
    if (hipEventCreateWithFlags(&t->order_ev, hipEventDisableTiming) != hipSuccess ||
        hipDeviceSynchronize() != hipSuccess) {
        free_table(t);
        return LSLAB_ERR_HIP;
    }
    t->dev.alloc_ctr = &t->state->alloc_ctr;
    t->dev.status = &t->state->status;
    t->dev.nbuckets = t->nbuckets;
    t->dev.pool = pool;
    t->dev.mask = (uint32_t)(t->nbuckets - 1);
    t->device_bytes = slab_total + t->nbuckets * 4 + sizeof(DevState) + (t->gt_cap + 1) * 12 +
                      t->submax * 12;
    *out = t;
    return LSLAB_OK;
}

extern "C" int lslab_map_create(lslab_table** out, int key_bytes, int value_bytes, unsigned n_log2,
                                uint64_t pool_slabs, const lslab_tuning* tuning) {
    TableType ty;
    if (key_bytes == 4 && value_bytes == 4) ty = T_MAP_32_32;
    else if (key_bytes == 4 && value_bytes == 8) ty = T_MAP_32_64;
    else if (key_bytes == 8 && value_bytes == 4) ty = T_MAP_64_32;
    else if (key_bytes == 8 && value_bytes == 8) ty = T_MAP_64_64;
    else return LSLAB_ERR_INVALID;
    return create_common(out, ty, true, key_bytes, value_bytes, n_log2, pool_slabs, tuning);
}

extern "C" int lslab_set_create(lslab_table** out, int key_bytes, unsigned n_log2,
                                uint64_t pool_slabs, const lslab_tuning* tuning) {
    TableType ty;
    if (key_bytes == 4) ty = T_SET_32;
    else if (key_bytes == 8) ty = T_SET_64;
    else return LSLAB_ERR_INVALID;
    return create_common(out, ty, false, key_bytes, 0, n_log2, pool_slabs, tuning);
}

extern "C" int lslab_destroy(lslab_table* t) {
    if (!t) return LSLAB_ERR_INVALID;
    (void)hipDeviceSynchronize();
    free_table(t);
    return LSLAB_OK;
}

extern "C" int lslab_set_tuning(lslab_table* t, const lslab_tuning* tuning) {
    if (!t || !tuning) return LSLAB_ERR_INVALID;
    lslab_tuning tu = *tuning;
    normalise(tu);
    if (!tuning_ok(tu)) return LSLAB_ERR_INVALID;
    const bool resize = tu.max_subbatch != t->tun.max_subbatch;
    t->tun = tu;
    if (resize) {
        // The scratch may still be in use by queued calls of this table.
        HIPCHK(hipDeviceSynchronize());
        return alloc_scratch(t);
    }
    return LSLAB_OK;
}

extern "C" int lslab_get_tuning(const lslab_table* t, lslab_tuning* out) {
    if (!t || !out) return LSLAB_ERR_INVALID;
    *out = t->tun;
    return LSLAB_OK;
}

// Orders this call after the table's previous call, whatever its stream.
// While the stream is being captured into a graph the graph's own edges order
// the calls, so the table's event is neither waited on nor recorded.
static bool capturing(hipStream_t s) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    return hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone;
}
static int begin_call(lslab_table* t, hipStream_t s) {
    if (!capturing(s)) HIPCHK(hipStreamWaitEvent(s, t->order_ev, 0));
    return LSLAB_OK;
}
static int end_call(lslab_table* t, hipStream_t s) {
    HIPCHK(hipGetLastError());
    if (!capturing(s)) HIPCHK(hipEventRecord(t->order_ev, s));
    return LSLAB_OK;
}

// Runs n ops through the plain kernel (locked or lock-free) or the two-phase
// kernels, depending on `engine`.
static int run_ops(lslab_table* t, const OpsDev& o, const OutDev& out, uint64_t n, int engine,
                   bool read_only, hipStream_t s) {
    if (n == 0) return LSLAB_OK;
    int rc = begin_call(t, s);
    if (rc) return rc;
    const int bt = t->tun.block_threads;
    rc = dispatch(t, [&](auto l) {
        using L = decltype(l);
        return dispatch_g(t->tun.lanes_per_op, [&](auto gc) {
            constexpr int G = decltype(gc)::value;
            if (read_only || engine == LSLAB_ENGINE_PLAIN) {
                const uint64_t grid = grid_for(t, n);
                if (read_only)
                    hipLaunchKernelGGL((k_plain<L, G, false>), dim3((unsigned)grid), dim3(bt), 0, s,
                                       t->dev, o, out, n);
                else
                    hipLaunchKernelGGL((k_plain<L, G, true>), dim3((unsigned)grid), dim3(bt), 0, s,
                                       t->dev, o, out, n);
            } else {
                for (uint64_t lo = 0; lo < n; lo += t->submax) {
                    const uint64_t m = std::min<uint64_t>(t->submax, n - lo);
                    const uint64_t grid = grid_for(t, m);
                    if (t->tun.combine == LSLAB_COMBINE_BLOCK)
                        hipLaunchKernelGGL((k_phase_r<L, G, true>), dim3((unsigned)grid), dim3(bt),
                                           0, s, t->dev, o, out, lo, m);
                    else
                        hipLaunchKernelGGL((k_phase_r<L, G, false>), dim3((unsigned)grid),
                                           dim3(bt), 0, s, t->dev, o, out, lo, m);
                    hipLaunchKernelGGL((k_phase_w<L, G>), dim3((unsigned)grid), dim3(bt), 0, s,
                                       t->dev, o, lo, m);
                }
            }
            return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
        });
    });
    if (rc) return rc;
    return end_call(t, s);
}

static OpsDev make_ops(const uint8_t* ops, uint8_t uop, const void* keys, uint32_t kstride,
                       const void* vals, uint32_t vstride) {
    OpsDev o;
    o.ops = ops;
    o.uop = uop;
    o.keys = static_cast<const char*>(keys);
    o.kstride = kstride;
    o.vals = static_cast<const char*>(vals);
    o.vstride = vstride;
    return o;
}

static OutDev make_out(void* pairs, void* vals, uint8_t* flags) {
    OutDev d;
    d.pairs = static_cast<char*>(pairs);
    d.vals = static_cast<char*>(vals);
    d.flags = flags;
    return d;
}

extern "C" int lslab_map_put_pairs(lslab_table* t, const void* pairs, void* out_old, size_t n,
                                   hipStream_t stream) {
    if (!t || !t->is_map || (n && (!pairs || !out_old))) return LSLAB_ERR_INVALID;
    // cuda::std::pair<K,V>: first at 0, second at alignof(pair) = max(sizeof K, sizeof V).
    const uint32_t al = (uint32_t)std::max(t->key_bytes, t->val_bytes);
    const uint32_t stride = 2 * al;
    OpsDev o = make_ops(nullptr, OP_PUT, pairs, stride, static_cast<const char*>(pairs) + al, stride);
    return run_ops(t, o, make_out(nullptr, out_old, nullptr), n, t->tun.engine, false, stream);
}

extern "C" int lslab_map_put(lslab_table* t, const void* keys, const void* values, void* out_old,
                             size_t n, hipStream_t stream) {
    if (!t || !t->is_map || (n && (!keys || !values || !out_old))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_PUT, keys, t->key_bytes, values, t->val_bytes);
    return run_ops(t, o, make_out(nullptr, out_old, nullptr), n, t->tun.engine, false, stream);
}

extern "C" int lslab_map_get(lslab_table* t, const void* keys, void* out_pairs, size_t n,
                             hipStream_t stream) {
    if (!t || !t->is_map || (n && (!keys || !out_pairs))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_GET, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(out_pairs, nullptr, nullptr), n, t->tun.engine, true, stream);
}

extern "C" int lslab_map_remove(lslab_table* t, const void* keys, void* out_pairs, size_t n,
                                hipStream_t stream) {
    if (!t || !t->is_map || (n && (!keys || !out_pairs))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_REMOVE, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(out_pairs, nullptr, nullptr), n, t->tun.engine, false, stream);
}

extern "C" int lslab_map_mixed(lslab_table* t, const uint8_t* ops, const void* keys,
                               const void* values, void* out_pairs, size_t n, hipStream_t stream) {
    if (!t || !t->is_map || (n && (!ops || !keys || !values || !out_pairs))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(ops, 0, keys, t->key_bytes, values, t->val_bytes);
    return run_ops(t, o, make_out(out_pairs, nullptr, nullptr), n, t->tun.engine, false, stream);
}

extern "C" int lslab_set_insert(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                                hipStream_t stream) {
    if (!t || t->is_map || (n && (!keys || !out))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_PUT, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(nullptr, nullptr, out), n, t->tun.engine, false, stream);
}

extern "C" int lslab_set_contains(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                                  hipStream_t stream) {
    if (!t || t->is_map || (n && (!keys || !out))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_GET, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(nullptr, nullptr, out), n, t->tun.engine, true, stream);
}

extern "C" int lslab_set_remove(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                                hipStream_t stream) {
    if (!t || t->is_map || (n && (!keys || !out))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(nullptr, OP_REMOVE, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(nullptr, nullptr, out), n, t->tun.engine, false, stream);
}

extern "C" int lslab_set_mixed(lslab_table* t, const uint8_t* ops, const void* keys, uint8_t* out,
                               size_t n, hipStream_t stream) {
    if (!t || t->is_map || (n && (!ops || !keys || !out))) return LSLAB_ERR_INVALID;
    OpsDev o = make_ops(ops, 0, keys, t->key_bytes, nullptr, 0);
    return run_ops(t, o, make_out(nullptr, nullptr, out), n, t->tun.engine, false, stream);
}

// Grid of a persistent kernel whose blocks wait on each other at grid
// barriers: every block must be resident at once.  One block of
// PERSIST_THREADS (4 waves per SIMD) per CU: the kernels are built for that
// (launch bounds 1024 cap them at 128 VGPRs; their ~106 SGPRs allow 7 waves
// per SIMD), and the occupancy query must confirm at least one block per CU.
// max_blocks (tuning) can shrink the grid for small batches.
static int persistent_grid(const lslab_table* t, const void* kernel, uint64_t* grid) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, PERSIST_THREADS, 0) !=
            hipSuccess ||
        per_cu < 1)
        return LSLAB_ERR_HIP;
    uint64_t g = (uint64_t)t->num_cus;
    if (t->tun.max_blocks > 0) g = std::min<uint64_t>(g, (uint64_t)t->tun.max_blocks);
    *grid = g;
    return LSLAB_OK;
}

extern "C" int lslab_exec_batches(lslab_table* t, size_t nbatches, const uint64_t* offsets,
                                  const uint8_t* ops, const void* keys, const void* values,
                                  void* out, hipStream_t stream) {
    if (!t || (nbatches && (!offsets || !ops || !keys || !out))) return LSLAB_ERR_INVALID;
    if (t->is_map && nbatches && !values) return LSLAB_ERR_INVALID;
    if (nbatches == 0) return LSLAB_OK;
    int rc = begin_call(t, stream);
    if (rc) return rc;
    OpsDev o = make_ops(ops, 0, keys, t->key_bytes, t->is_map ? values : nullptr, t->val_bytes);
    OutDev d = t->is_map ? make_out(out, nullptr, nullptr)
                         : make_out(nullptr, nullptr, static_cast<uint8_t*>(out));
    BarState* bar = &t->state->bar;
    rc = dispatch(t, [&](auto l) {
        using L = decltype(l);
        return dispatch_g(t->tun.lanes_per_op, [&](auto gc) {
            constexpr int G = decltype(gc)::value;
            return dispatch_c(t->tun.combine, [&](auto cc) {
                constexpr bool BC = decltype(cc)::value;
                uint64_t grid = 0;
                const int rc2 =
                    persistent_grid(t, reinterpret_cast<const void*>(k_batches<L, G, BC>), &grid);
                if (rc2) return rc2;
                if (t->tun.max_blocks > 0)
                    grid = std::min<uint64_t>(grid, (uint64_t)t->tun.max_blocks);
                hipLaunchKernelGGL(k_reset_bar, dim3(1), dim3(64), 0, stream, bar);
                hipLaunchKernelGGL((k_batches<L, G, BC>), dim3((unsigned)grid),
                                   dim3(PERSIST_THREADS), 0, stream, t->dev, o, d, offsets,
                                   (uint64_t)nbatches, t->submax, bar);
                return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
            });
        });
    });
    if (rc) return rc;
    return end_call(t, stream);
}

extern "C" int lslab_serve(lslab_table* t, const lslab_serve_params* p, const uint8_t* ops,
                           const void* keys, const void* values, void* out, hipStream_t stream) {
    if (!t || !p || (p->nreq && (!p->req_offsets || !p->arrival_ticks || !p->complete_ticks ||
                                 !p->tick_of || !ops || !keys || !out)))
        return LSLAB_ERR_INVALID;
    if (t->is_map && p->nreq && !values) return LSLAB_ERR_INVALID;
    int rc = begin_call(t, stream);
    if (rc) return rc;
    ServeDev sv;
    sv.nreq = p->nreq;
    sv.off = p->req_offsets;
    sv.arrival = p->arrival_ticks;
    sv.reserve = p->req_reserve;
    sv.complete = p->complete_ticks;
    sv.tick_of = p->tick_of;
    sv.max_tick_ops = p->max_tick_ops ? p->max_tick_ops : t->submax;
    sv.max_tick_reqs = p->max_tick_reqs ? p->max_tick_reqs : 0xffffffffu;
    OpsDev o = make_ops(ops, 0, keys, t->key_bytes, t->is_map ? values : nullptr, t->val_bytes);
    OutDev d = t->is_map ? make_out(out, nullptr, nullptr)
                         : make_out(nullptr, nullptr, static_cast<uint8_t*>(out));
    BarState* bar = &t->state->bar;
    ServeState* ss = &t->state->serve;
    rc = dispatch(t, [&](auto l) {
        using L = decltype(l);
        return dispatch_g(t->tun.lanes_per_op, [&](auto gc) {
            constexpr int G = decltype(gc)::value;
            return dispatch_c(t->tun.combine, [&](auto cc) {
                constexpr bool BC = decltype(cc)::value;
                uint64_t grid = 0;
                const int rc2 =
                    persistent_grid(t, reinterpret_cast<const void*>(k_serve<L, G, BC>), &grid);
                if (rc2) return rc2;
                if (t->tun.max_blocks > 0)
                    grid = std::min<uint64_t>(grid, (uint64_t)t->tun.max_blocks);
                hipLaunchKernelGGL(k_reset_serve, dim3(1), dim3(64), 0, stream, ss, bar);
                hipLaunchKernelGGL((k_serve<L, G, BC>), dim3((unsigned)grid),
                                   dim3(PERSIST_THREADS), 0, stream, t->dev, o, d, sv, ss, bar,
                                   t->submax);
                return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
            });
        });
    });
    if (rc) return rc;
    return end_call(t, stream);
}

extern "C" int lslab_serve_result(lslab_table* t, lslab_serve_stats* out, hipStream_t stream) {
    if (!t || !out) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipStreamSynchronize(stream));
    ServeState ss;
    HIPCHK(hipMemcpy(&ss, &t->state->serve, sizeof(ss), hipMemcpyDeviceToHost));
    out->ticks = ss.ticks;
    out->rejected = ss.rejected;
    out->served_ops = ss.served_ops;
    out->elapsed_ticks = ss.end;
    return LSLAB_OK;
}

extern "C" int lslab_status(lslab_table* t, hipStream_t stream) {
    if (!t) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipStreamSynchronize(stream));
    uint32_t st = 0;
    HIPCHK(hipMemcpy(&st, &t->state->status, sizeof(st), hipMemcpyDeviceToHost));
    if (st) HIPCHK(hipMemset(&t->state->status, 0, sizeof(uint32_t)));
    if (st & ST_CORRUPT) return LSLAB_ERR_CORRUPT;
    if (st & ST_TIMEOUT) return LSLAB_ERR_TIMEOUT;
    if (st & ST_POOL_EXHAUSTED) return LSLAB_ERR_POOL_EXHAUSTED;
    return LSLAB_OK;
}

extern "C" int lslab_stats_get(lslab_table* t, lslab_stats* out, hipStream_t stream) {
    if (!t || !out) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipMemsetAsync(&t->state->stats, 0, sizeof(StatsDev), stream));
    const uint64_t grid = (t->nbuckets + 255) / 256;
    int rc = dispatch(t, [&](auto l) {
        using L = decltype(l);
        hipLaunchKernelGGL(k_stats<L>, dim3((unsigned)grid), dim3(256), 0, stream, t->dev,
                           &t->state->stats);
        return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
    });
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    DevState hs;
    HIPCHK(hipMemcpy(&hs, t->state, sizeof(DevState), hipMemcpyDeviceToHost));
    out->buckets = t->nbuckets;
    out->pool_slabs = t->pool;
    out->slabs_used = std::min<uint64_t>(hs.alloc_ctr, t->pool);
    out->present = hs.stats.present;
    out->max_chain = hs.stats.max_chain;
    out->slab_bytes = t->slab_bytes;
    out->device_bytes = t->device_bytes;
    HIPCHK(hipEventRecord(t->order_ev, stream));
    return LSLAB_OK;
}

extern "C" int lslab_dump(lslab_table* t, void* keys_out, void* vals_out, uint64_t cap,
                          uint64_t* count, hipStream_t stream) {
    if (!t || !count || (cap && !keys_out)) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipMemsetAsync(&t->state->dump_counter, 0, sizeof(unsigned long long), stream));
    const uint64_t grid = (t->nbuckets + 255) / 256;
    int rc = dispatch(t, [&](auto l) {
        using L = decltype(l);
        hipLaunchKernelGGL(k_dump<L>, dim3((unsigned)grid), dim3(256), 0, stream, t->dev,
                           static_cast<typename L::K*>(keys_out),
                           static_cast<typename L::V*>(vals_out), cap, &t->state->dump_counter);
        return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
    });
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    unsigned long long c = 0;
    HIPCHK(hipMemcpy(&c, &t->state->dump_counter, sizeof(c), hipMemcpyDeviceToHost));
    *count = c;
    HIPCHK(hipEventRecord(t->order_ev, stream));
    return LSLAB_OK;
}

extern "C" const char* lslab_error_string(int status) {
    switch (status) {
        case LSLAB_OK: return "ok";
        case LSLAB_ERR_INVALID: return "invalid argument";
        case LSLAB_ERR_HIP: return "HIP runtime error";
        case LSLAB_ERR_OOM: return "device out of memory";
        case LSLAB_ERR_POOL_EXHAUSTED: return "slab pool exhausted";
        case LSLAB_ERR_TIMEOUT: return "device spin time bound exceeded";
        case LSLAB_ERR_CORRUPT: return "table structure check failed";
    }
    return "unknown status";
}
