// lslab: the gfx950 kernels of every form (plain, two-phase, persistent
// batched, served, introspection) and the per-layout host launchers that
// start them.  launch.hip instantiates the launchers for one table layout
// per translation unit; lslab.hip (the C ABI) only calls them.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <type_traits>

#include "../../include/lslab_c.h"
#include "runs.hpp"

using namespace lslab_dev;

// Persistent kernels run one block of 1024 threads (16 wavefronts) per CU:
// 16 waves per CU keep random slab reads at full rate, and one block per CU
// keeps grid barriers cheap (256 arrivals, one cache invalidate per CU).
constexpr int PERSIST_THREADS = 1024;

// A wave-uniform 64-bit value moved to scalar registers (loop bounds loaded
// from LDS or with atomic loads would otherwise hold vector registers across
// the phases of a batch, run or tick).
__device__ __forceinline__ uint64_t uniform64(uint64_t v) {
    return ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(v >> 32)) << 32) |
           __builtin_amdgcn_readfirstlane((uint32_t)v);
}

// Grid-wide barrier state for the persistent kernels.  Blocks arrive on one
// of BAR_GROUPS group counters (block id mod BAR_GROUPS), the last block of a
// group arrives on the top counter, and the last group releases every group
// by bumping the group's generation word.  Counters only grow within a
// launch (k_reset_bar zeroes them before it): a block's arrival count tells
// it which barrier instance it is in, so nothing is reset on the critical
// path and no generation is read before arriving.  Each word sits on its own
// 128-B line, so arrivals and polls of different groups do not share a line.
constexpr uint32_t BAR_GROUPS = 16;
struct alignas(128) BarWord {
    uint32_t count;  // arrivals so far in this launch
    uint32_t gen;    // completed barrier instances (group words only)
    uint32_t pad[30];
};
struct BarState {
    BarWord top;
    BarWord group[BAR_GROUPS];
    uint32_t abort;
    uint32_t beat;  // heartbeat of a live wait elsewhere (the served loop's idle scheduler)
    uint32_t runs;  // batched launch: 1 if two consecutive batches fit one run (k_reset_batches)
    uint32_t bad;   // batched launch: the batch offsets decrease somewhere (k_reset_batches)
    uint32_t pad[28];
};

// Plain per-batch kernel: one G-lane group per op, grid-stride.  RC: the
// table's reclaim policy (only mutations, under locks, differ).
template <class L, int G, bool LOCKED, bool RC>
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
        run_op<L, G, LOCKED, RC>(t, op, key, val, rf, rv, ok);
        if (Grp<G>::lane() == 0) write_result<L>(out, i, rf, rv);
    }
}

template <class L, int G, bool BC>
__global__ __launch_bounds__(256) void k_phase_r(TableDev t, OpsDev o, OutDev out, uint64_t lo,
                                                 uint64_t m) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    uint32_t* rc = rep_counter();
    if (threadIdx.x == 0) *rc = 0;
    if constexpr (BC)
        comb_init<256>();
    else
        __syncthreads();
    phase_r_range<L, G, 256, BC>(t, o, out, lo, m, gid, ng,
                                 t.rep + block_rep_base(m, ng, blockDim.x / G, blockIdx.x));
    __syncthreads();
    if (threadIdx.x == 0) t.rep_cnt[blockIdx.x] = *rc;
}

// Same grid as the k_phase_r launch before it: block b runs block b's list.
template <class L, int G, bool RC>
__global__ __launch_bounds__(256) void k_phase_w(TableDev t, OpsDev o, uint64_t lo, uint64_t m) {
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    phase_w_block<L, G, RC>(t, o, lo, t.rep + block_rep_base(m, ng, blockDim.x / G, blockIdx.x),
                        t.rep_cnt[blockIdx.x]);
}

// Grid barrier for a grid that is fully resident.  Every wave waits for its
// own writes to complete (shared table writes are agent-scope stores or
// atomics, so nothing else is needed to publish them), one lane per block
// arrives, and the last arrival releases everyone.  With `acquire` the block
// then drops stale cached lines (one L2 + L1 invalidate by wave 0) so that
// the next phase may read the table with plain loads; a phase that reads
// shared state only through agent-scope loads, or data its own CU wrote,
// passes acquire = false.  No cache-wide write-back is issued.  Spins are
// bounded; on timeout or abort the barrier returns false and the caller
// leaves the kernel (limit: ticks without progress, see the heartbeat
// below).  reset, reset2 (optional): LDS words that thread 0 zeroes while the
// whole block is inside the barrier.
__device__ __forceinline__ bool grid_sync(BarState* bar, uint32_t nblocks, uint64_t limit,
                                          bool acquire, uint32_t* reset = nullptr,
                                          uint32_t* reset2 = nullptr) {
    wait_vm();
    __syncthreads();
    if (reset && threadIdx.x == 0) *reset = 0;
    if (reset2 && threadIdx.x == 0) *reset2 = 0;
    __shared__ uint32_t ok_s;
    if (threadIdx.x == 0) {
        uint32_t ok = 1;
        const uint32_t x = blockIdx.x % BAR_GROUPS;
        const uint32_t ngroups = nblocks < BAR_GROUPS ? nblocks : BAR_GROUPS;
        const uint32_t members = (nblocks - x + BAR_GROUPS - 1) / BAR_GROUPS;
        BarWord* gw = &bar->group[x];
        const uint32_t arrived =
            __hip_atomic_fetch_add(&gw->count, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t inst = arrived / members;  // this barrier's instance for the group
        if ((arrived + 1u) % members == 0u) {
            const uint32_t top = __hip_atomic_fetch_add(&bar->top.count, 1u, __ATOMIC_RELAXED,
                                                        __HIP_MEMORY_SCOPE_AGENT);
            if ((top + 1u) % ngroups == 0u) {
                for (uint32_t y = 0; y < ngroups; ++y)
                    __hip_atomic_fetch_add(&bar->group[y].gen, 1u, __ATOMIC_RELAXED,
                                           __HIP_MEMORY_SCOPE_AGENT);
            }
        }
        // The bound restarts whenever the heartbeat moves: a block that waits
        // here while the served loop's scheduler waits for an arrival is not
        // stuck, however long the gap.
        uint64_t t0 = now_ticks();
        uint32_t seen = ld_agent(&bar->beat);
        while ((int32_t)(ld_agent(&gw->gen) - inst) <= 0) {
            const uint64_t now = now_ticks();
            const uint32_t beat = ld_agent(&bar->beat);
            if (beat != seen) {
                seen = beat;
                t0 = now;
            }
            if (ld_agent(&bar->abort) || now - t0 > limit) {
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
    if (acquire && threadIdx.x < 64) {
        acquire_agent();
        wait_vm();
    }
    __syncthreads();
    return ok_s != 0;
}

// Persistent batched kernel: batches in order.  Consecutive batches that
// together hold at most run_cap ops (and at most RUN_MAX_BATCHES of them) form
// a run (runs.hpp): pass A orders the run's mutations per (key, batch), pass
// B resolves every get and every first mutation of a chain against the
// nearest earlier batch, phase W writes each key's final state once.  Runs
// alternate between two sets of run tables, so phase W of a run shares its
// step with pass A of the next run (pass A reads nothing of the table and
// writes only its own set): two grid barriers per run,
//     [W(k-1) | A(k)]  barrier+acquire  B(k)  barrier  [W(k) | A(k+1)] ...
// A batch that forms no run with its successor runs alone after any pending
// phase W: phase R, barrier, phase W, barrier, in sub-batches of submax ops
// when it is larger.
//
// RUNS = false is the same kernel without the run path: every batch runs
// alone.  A stream in which no two consecutive batches fit one run (the
// launcher's k_reset_batches finds out, bar->runs = 0) takes that kernel,
// whose phase R loop keeps all of its state in registers (the run path's
// code raises the whole kernel's register use past the 128 a 1024-thread
// block allows, and the compiler then spills).  Both kernels are launched;
// the one that does not match bar->runs returns at once.
template <class L, int G, bool BC, bool RC, bool RUNS>
__global__ __launch_bounds__(PERSIST_THREADS) void k_batches(TableDev t, OpsDev o, OutDev out,
                                                 const uint64_t* offsets, uint64_t nbatches,
                                                 uint64_t submax, uint64_t run_cap,
                                                 BarState* bar) {
    if (ld_agent(&bar->bad)) return;                   // grid-uniform: refused input
    if ((ld_agent(&bar->runs) != 0) != RUNS) return;  // grid-uniform
    const uint64_t ng = (uint64_t)gridDim.x * blockDim.x / G;
    const uint64_t gid = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const uint64_t gpb = blockDim.x / G;
    uint32_t* rc = rep_counter();
    uint32_t* cc = clr_counters();
    RepEnt* rb = t.rep + block_rep_base_fixed(submax, ng, gpb, blockIdx.x);
    const uint64_t clr_base = run_clr_stride(run_cap, gridDim.x) * blockIdx.x;
    if (threadIdx.x == 0) {
        *rc = 0;
        cc[0] = 0;
        cc[1] = 0;
    }
    if constexpr (BC)
        comb_init<PERSIST_THREADS>();
    else
        __syncthreads();
    auto fail = [&]() {
        if (threadIdx.x == 0)
            __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED,
                                  __HIP_MEMORY_SCOPE_AGENT);
    };
    int par = 0;         // table set of the next run
    bool pend = false;   // the previous run's phase W is still to do (set 1 - par)
    uint64_t plo = 0;    // that run's first op
    // phase W of the pending run, in the other table set
    auto flush_w = [&]() {
        const RunTabs pt = par ? t.run[0] : t.run[1];
        run_phase_w<L, G, RC>(t, pt, o, plo, rb, *rc, rep_lds(), REP_LDS, pt.clr + clr_base,
                              cc[par ^ 1]);
    };
    if constexpr (!RUNS) {
        // Every batch alone.  The end of the next batch is loaded while this
        // batch runs, so a batch starts right after the previous barrier.
        if (nbatches == 0) return;
        uint64_t lo = offsets[0], hi = offsets[1];
        for (uint64_t b = 0; b < nbatches; ++b) {
            const uint64_t nxt = b + 2 <= nbatches ? offsets[b + 2] : hi;
            for (uint64_t slo = lo; slo < hi; slo += submax) {
                const uint64_t m = min(submax, hi - slo);
                phase_r_range<L, G, PERSIST_THREADS, BC>(t, o, out, slo, m, gid, ng, rb,
                                                         rep_lds(), REP_LDS);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, false)) return fail();
                phase_w_block<L, G, RC>(t, o, slo, rb, *rc, rep_lds(), REP_LDS);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc)) return fail();
            }
            lo = hi;
            hi = nxt;
        }
        return;
    }
    // The window offsets[b .. b + RUN_MAX_BATCHES] of the run starting at
    // batch b sits in LDS, ros[cur]; one thread per boundary loads it.  The
    // next run's window is loaded before this run's first grid barrier and
    // stored into ros[cur ^ 1] after it, so no run waits for its boundaries.
    RunOffs* ros = run_offs();
    int cur = 0;
    if (nbatches == 0) return;
    {
        const uint64_t wn = min((uint64_t)RUN_MAX_BATCHES, nbatches);
        if (threadIdx.x <= wn) ros[0].win[threadIdx.x] = offsets[threadIdx.x];
        if (threadIdx.x == 0) ros[0].nb = (int)wn;
        __syncthreads();
    }
    uint64_t b = 0;
    while (b < nbatches) {
        RunOffs* ro = &ros[cur];
        // Batches in the run: the first e < wn whose end lies more than
        // run_cap ops past the window's start (nb was preset to wn).
        const uint64_t wn = min((uint64_t)RUN_MAX_BATCHES, nbatches - b);
        if (threadIdx.x < wn && ro->win[threadIdx.x + 1] - ro->win[0] > run_cap)
            atomicMin(&ro->nb, (int)threadIdx.x);
        __syncthreads();
        const uint64_t lo = uniform64(ro->win[0]);
        const uint64_t nr = uniform64((uint64_t)ro->nb);
        const uint64_t b2 = b + (nr <= 1 ? 1 : nr);  // the next run's first batch
        const uint64_t wn2 = b2 < nbatches ? min((uint64_t)RUN_MAX_BATCHES, nbatches - b2) : 0;
        const bool pfl = b2 < nbatches && threadIdx.x <= wn2;
        uint64_t pf = 0;
        auto pf_store = [&]() {
            RunOffs* nx = &ros[cur ^ 1];
            if (pfl) nx->win[threadIdx.x] = pf;
            if (threadIdx.x == 0) nx->nb = (int)wn2;
        };
        if (nr <= 1) {  // batch b alone, in sub-batches of at most submax ops
            const uint64_t hi = uniform64(ro->win[1]);
            if (pend) {
                flush_w();
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc, &cc[par ^ 1])) return fail();
                pend = false;
            }
            if (pfl) pf = offsets[b2 + threadIdx.x];
            if (hi == lo) pf_store();
            for (uint64_t slo = lo; slo < hi; slo += submax) {
                const uint64_t m = min(submax, hi - slo);
                phase_r_range<L, G, PERSIST_THREADS, BC>(t, o, out, slo, m, gid, ng, rb,
                                                         rep_lds(), REP_LDS);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, false)) return fail();
                if (slo == lo) pf_store();
                phase_w_block<L, G, RC>(t, o, slo, rb, *rc, rep_lds(), REP_LDS);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc)) return fail();
            }
            __syncthreads();  // the next window is in LDS
            cur ^= 1;
            b = b2;
            continue;
        }
        const uint64_t m = uniform64(ro->win[nr] - lo);
        if (threadIdx.x <= nr) ro->off[threadIdx.x] = ro->win[threadIdx.x] - lo;
        __syncthreads();
        const RunTabs rt = par ? t.run[1] : t.run[0];
        if (pend) flush_w();  // the previous run's write-back, in the other table set
        run_pass_a<L>(t, rt, o, out, lo, m, *ro, rt.clr + clr_base, &cc[par]);
        if (pfl) pf = offsets[b2 + threadIdx.x];
        // pass B reads the table (after that write-back) and the run tables
        // with plain loads: drop stale lines
        if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc, &cc[par ^ 1])) return fail();
        pf_store();
        run_pass_b<L, G>(t, rt, o, out, lo, m, *ro, gid, ng, rb, rep_lds(), REP_LDS, rc);
        if (!grid_sync(bar, gridDim.x, t.spin_ticks, false)) return fail();
        pend = true;
        plo = lo;
        par ^= 1;
        cur ^= 1;
        b = b2;
    }
    if (pend) {
        flush_w();
        if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc, &cc[par ^ 1])) return fail();
    }
}

__device__ __forceinline__ void reset_bar(BarState* bar) {
    if (threadIdx.x <= BAR_GROUPS && blockIdx.x == 0) {
        BarWord* w = threadIdx.x == 0 ? &bar->top : &bar->group[threadIdx.x - 1];
        w->count = 0;
        w->gen = 0;
        if (threadIdx.x == 0) bar->abort = 0;
    }
}

// Templated on the layout only so that each layout's unit has its own copy.
template <class L>
__global__ void k_reset_bar(BarState* bar) { reset_bar(bar); }

// Before a batched launch (one block): resets the barrier and sets
// bar->runs.  With mode LSLAB_RUNS_ON or _OFF that is 1 or 0; with 0 (auto)
// it is whether any two consecutive batches together hold at most run_cap
// ops, i.e. whether k_batches would form a run.  Sets bar->bad (and the
// table's ST_BAD_INPUT) when the offsets decrease.
template <class L>
__global__ void k_reset_batches(BarState* bar, const uint64_t* offsets, uint64_t nbatches,
                                uint64_t run_cap, int mode, uint32_t* status) {
    reset_bar(bar);
    __shared__ uint32_t any, bad;
    if (threadIdx.x == 0) {
        any = mode == LSLAB_RUNS_ON ? 1u : 0u;
        bad = 0;
    }
    __syncthreads();
    // Decreasing offsets would give a batch a negative (wrapped) size: the
    // launch is refused before any op is read.
    bool fits = false, dec = false;
    for (uint64_t i = threadIdx.x; i < nbatches; i += blockDim.x) {
        const uint64_t a = offsets[i];
        dec |= offsets[i + 1] < a;
        if (mode == 0 && i + 2 <= nbatches) fits |= offsets[i + 2] - a <= run_cap;
    }
    if (fits) any = 1;
    if (dec) bad = 1;
    __syncthreads();
    if (threadIdx.x == 0) {
        st_coh(&bar->runs, any);
        st_coh(&bar->bad, bad);
        if (bad)
            __hip_atomic_fetch_or(status, ST_BAD_INPUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Served form: a device-resident request loop.
//
// Requests r = 0..nreq-1 own the contiguous op ranges [off[r], off[r+1]) and
// arrive at arrival[r] (ticks of the 100 MHz constant clock after the loop
// starts, nondecreasing).  Each tick admits, in arrival order, every request
// that has arrived, up to max_tick_ops ops and max_tick_reqs requests,
// provided the slab pool keeps headroom for it: a request reserves one slab
// per put, the most its puts can allocate (an insert allocates at most one
// slab).  A request whose reservation can never fit (the pool only shrinks)
// is rejected.  The other blocks run the admitted ops as one two-phase batch
// while block 0 admits the next tick; the last block stamps each request's
// completion time.  Requests arriving while a tick runs join the next tick.
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
    uint64_t op_lo, op_hi;  // the tick's op range: off[r_lo], off[r_hi]
    uint64_t ticks;
    uint64_t rejected;
    uint64_t served_ops;
    uint64_t end;
    uint32_t done;
    uint32_t pad;
};

constexpr uint32_t SERVE_REJECTED = 0xffffffffu;

// Inclusive prefix sum of v over the threads of a block (up to
// PERSIST_THREADS), and the block's total in *total.  All threads call it.
__device__ __forceinline__ uint64_t block_scan(uint64_t v, uint64_t* total) {
    __shared__ uint64_t wsum[PERSIST_THREADS / 64];
    const int lane = (int)__lane_id();
    const int w = (int)(threadIdx.x / 64);
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t lo = __shfl_up((uint32_t)v, d, 64);
        const uint32_t hi = __shfl_up((uint32_t)(v >> 32), d, 64);
        if (lane >= d) v += ((uint64_t)hi << 32) | lo;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint64_t before = 0, all = 0;
    const int nw = (int)(blockDim.x / 64);
    for (int i = 0; i < nw; ++i) {
        if (i < w) before += wsum[i];
        all += wsum[i];
    }
    __syncthreads();  // wsum is reused by the next call
    *total = all;
    return v + before;
}

// A tick's admission, by all threads of block 0, in arrival order: the
// requests that have arrived and fit the op cap, the request cap and the
// slab-pool headroom, the rules of a request-by-request admission with the
// clock read once per chunk.  The first request of a tick is taken by thread
// 0 alone: it may wait for the request's arrival (an idle loop), or reject it
// when the pool could never hold its reservation.  The following requests
// are taken by the whole block in chunks of one request per thread: each
// thread loads its request's arrival, size and reservation, prefix sums give
// the running op and slab totals, and the tick admits the longest prefix
// that passes every check.
// Block 0's scheduler state for the whole launch (its LDS): the next request
// to consider, an upper bound of the slabs in use once every admitted tick
// has finished (the last exact count plus the reservations admitted since),
// and the admission being built.
struct SchedState {
    uint64_t r;
    uint64_t bound;
    uint64_t used;  // slabs in use the admission counts on
    uint64_t lo;    // first request of the admission
    uint64_t n, ops, res;
    uint64_t off0;  // op offset of the admission's first request
    int open;       // the admission stopped only at a request not yet arrived
    int bad;        // request offsets decrease at request r: the loop ends there
};

__device__ __forceinline__ SchedState* sched_state() {
    __shared__ SchedState ss;
    return &ss;
}

// Starts an empty admission.  exact: no admitted tick is still writing back,
// so the pool's counters are final and are read; otherwise the headroom
// check uses the bound, and a request the bound refuses is not admitted yet:
// it waits for a later admission, at the latest the exact one that an empty
// tick leads to.
__device__ inline void sched_begin(const TableDev& t, bool exact) {
    SchedState* S = sched_state();
    if (threadIdx.x == 0) {
        S->used = exact ? ld_agent(t.alloc_ctr) - ld_agent(t.free_count) : S->bound;
        S->lo = S->r;
        S->n = S->ops = S->res = S->off0 = 0;
        S->open = 1;
        S->bad = 0;
    }
    __syncthreads();
}

// Adds to the admission the requests that qualify now.  wait: with nothing
// admitted yet, wait for the next request's arrival, however far away.  A
// request whose reservation exceeds the pool's exact headroom can never fit
// (the pool only shrinks) and is rejected.
__device__ inline void sched_fill(const TableDev& t, const ServeDev& sv, ServeState* ss,
                                  BarState* bar, bool exact, bool wait) {
    __shared__ uint64_t st_el, st_ps, st_pn;
    __shared__ uint64_t st_pos[PERSIST_THREADS / 64];
    __shared__ int st_go;
    SchedState* S = sched_state();
    if (S->n == 0) {  // block-uniform (LDS, written before a block barrier)
        if (threadIdx.x == 0) {
            uint64_t r = S->r;
            while (r < sv.nreq) {
                if (sv.arrival[r] > now_ticks() - ss->t0) {
                    if (!wait) break;
                    // Idle: wait for the next arrival with a heartbeat every
                    // millisecond, so that the blocks waiting in the grid
                    // barrier meanwhile do not take the loop for stuck.
                    uint64_t last = now_ticks();
                    for (;;) {
                        const uint64_t now = now_ticks();
                        if (now - ss->t0 >= sv.arrival[r]) break;
                        if (now - last >= TICKS_PER_MS) {
                            last = now;
                            __hip_atomic_fetch_add(&bar->beat, 1u, __ATOMIC_RELAXED,
                                                   __HIP_MEMORY_SCOPE_AGENT);
                        }
                        __builtin_amdgcn_s_sleep(8);
                    }
                }
                const uint64_t off0 = sv.off[r];
                const uint64_t off1 = sv.off[r + 1];
                if (off1 < off0) {  // refused: no op of it or after it runs
                    S->bad = 1;
                    S->open = 0;
                    __hip_atomic_fetch_or(t.status, ST_BAD_INPUT, __ATOMIC_RELAXED,
                                          __HIP_MEMORY_SCOPE_AGENT);
                    break;
                }
                const uint64_t sz = off1 - off0;
                const uint64_t need = sv.reserve ? sv.reserve[r] : sz;
                if (S->used + need > t.pool && !exact) {  // wait for the exact count
                    S->open = 0;
                    break;
                }
                if (S->used + need > t.pool) {  // can never fit: the pool only shrinks
                    sv.tick_of[r] = SERVE_REJECTED;
                    sv.complete[r] = now_ticks() - ss->t0;
                    ss->rejected++;
                    r++;
                    S->lo = r;
                    continue;
                }
                S->n = 1;
                S->ops = sz;
                S->res = need;
                S->off0 = off0;
                r++;
                break;
            }
            if (r >= sv.nreq) S->open = 0;
            S->r = r;
            st_go = S->n > 0 && S->open;
        }
    } else if (threadIdx.x == 0) {
        st_go = S->open;
    }
    __syncthreads();
    while (st_go) {
        const uint64_t r = S->r;
        if (threadIdx.x == 0) st_el = now_ticks() - ss->t0;
        __syncthreads();
        const uint64_t i = r + threadIdx.x;
        const bool valid = i < sv.nreq;
        uint64_t sz = 0, need = 0;
        bool arrived = false;
        bool dec = false;  // decreasing offsets: left to the first-request check
        if (valid) {
            const uint64_t o0 = sv.off[i], o1 = sv.off[i + 1];
            dec = o1 < o0;
            sz = dec ? 0 : o1 - o0;
            need = sv.reserve ? sv.reserve[i] : sz;
            arrived = sv.arrival[i] <= st_el;
        }
        uint64_t tot;
        const uint64_t ps = block_scan(sz, &tot);
        const uint64_t pn = block_scan(need, &tot);
        const bool fits = valid && !dec && S->ops + ps <= sv.max_tick_ops &&
                          S->n + threadIdx.x + 1 <= (uint64_t)sv.max_tick_reqs &&
                          S->used + S->res + pn <= t.pool;
        const bool ok = fits && arrived;
        // the first request that fails: the tick admits the ones before it
        uint64_t f = ok ? (uint64_t)blockDim.x : (uint64_t)threadIdx.x;
        for (int d = 32; d > 0; d >>= 1) {
            const uint32_t lo = __shfl_xor((uint32_t)f, d, 64);
            f = min(f, (uint64_t)lo);
        }
        if (__lane_id() == 0) st_pos[threadIdx.x / 64] = f;
        __syncthreads();
        uint64_t k = blockDim.x;
        for (int w = 0; w < (int)(blockDim.x / 64); ++w) k = min(k, st_pos[w]);
        if (k > 0 && threadIdx.x == k - 1) {
            st_ps = ps;
            st_pn = pn;
        }
        // a later fill can only add requests when the first failure was an
        // arrival still to come
        if (k < blockDim.x && threadIdx.x == k) S->open = fits && !arrived;
        __syncthreads();
        if (threadIdx.x == 0) {
            if (k > 0) {
                S->n += k;
                S->ops += st_ps;
                S->res += st_pn;
                S->r = r + k;
            }
            if (S->r >= sv.nreq) S->open = 0;
            st_go = k == blockDim.x && S->open;
        }
        __syncthreads();
    }
}

// Publishes the admission as the next tick (read by every block after the
// grid barrier: agent-scope stores).  The admitted requests are contiguous,
// so their ops are [off[r_lo], off[r_lo] + ops).  No admission with requests
// left is an empty tick; with none left the loop is done.
__device__ inline void sched_publish(const ServeDev& sv, ServeState* ss) {
    SchedState* S = sched_state();
    if (threadIdx.x == 0) {
        const uint64_t op_lo = S->n ? S->off0 : 0;
        st_coh(&ss->r_lo, S->lo);
        st_coh(&ss->r_hi, S->r);
        st_coh(&ss->op_lo, op_lo);
        st_coh(&ss->op_hi, op_lo + S->ops);
        ss->next_req = S->r;
        ss->served_ops += S->ops;
        if (S->n == 0 && (S->r >= sv.nreq || S->bad)) st_coh(&ss->done, 1u);
        S->bound = S->used + S->res;  // the admitted tick allocates at most its reservations
    }
}

// Served loop.  With more than one block, block 0 only schedules: it builds
// the next tick's admission while the other blocks run the current tick's
// phase R, adds the requests that arrived meanwhile while they run phase W,
// and publishes it at the grid barrier that ends the tick, so admission is
// off the tick's critical path.  Each tick: phase R, grid barrier,
// completion stamps (last block), phase W, grid barrier; a tick larger than
// a sub-batch runs in sub-batches, with a barrier between.  After an empty
// tick (nothing admitted) block 0 admits with the pool's exact counts,
// waiting for the next arrival if there is none.
template <class L, int G, bool BC, bool RC>
__global__ __launch_bounds__(PERSIST_THREADS) void k_serve(TableDev t, OpsDev o, OutDev out, ServeDev sv,
                                               ServeState* ss, BarState* bar, uint64_t submax) {
    const bool split = gridDim.x > 1;  // block 0 schedules only
    const uint64_t wb = split ? blockIdx.x - (blockIdx.x > 0) : blockIdx.x;  // worker index
    const uint64_t nwb = split ? gridDim.x - 1 : gridDim.x;
    const uint64_t ng = nwb * blockDim.x / G;
    const uint64_t gid = (wb * blockDim.x + threadIdx.x) / G;
    __shared__ uint64_t stamp_s, tick_req[2];
    uint32_t* rc = rep_counter();
    RepEnt* rb = t.rep + block_rep_base_fixed(submax, ng, blockDim.x / G, wb);
    SchedState* S = sched_state();
    if (threadIdx.x == 0) *rc = 0;
    if constexpr (BC)
        comb_init<PERSIST_THREADS>();
    else
        __syncthreads();
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st_coh(&ss->t0, now_ticks());  // read by the stamping block
        S->r = 0;
        S->bound = 0;
    }
    if (split && blockIdx.x == 0) {
        // The scheduling block: the same grid barriers as the workers, and
        // admission in between.  A tick's op cap counts the workers' share
        // of the grid, so that a tick takes as many phase-R rounds on the
        // workers as the cap would on the whole grid.
        if (sv.max_tick_ops < (1ull << 48))  // larger caps do not bind
            sv.max_tick_ops = max(sv.max_tick_ops * (gridDim.x - 1) / gridDim.x, (uint64_t)1);
        bool running = false;  // the last tick's admission is already built
        for (;;) {
            if (!running) {
                sched_begin(t, true);
                sched_fill(t, sv, ss, bar, true, true);
            }
            sched_publish(sv, ss);
            if (!grid_sync(bar, gridDim.x, t.spin_ticks, running, rc)) goto fail;
            running = false;
            if (ld_agent(&ss->done)) break;
            const uint64_t lo = ld_coh(&ss->op_lo), hi = ld_coh(&ss->op_hi);
            if (lo == hi) continue;  // empty tick: admit again, exactly
            sched_begin(t, false);  // the next tick's admission, during this tick
            sched_fill(t, sv, ss, bar, false, false);
            for (uint64_t slo = lo;;) {
                const uint64_t m = min(submax, hi - slo);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, false)) goto fail;
                if (slo + m >= hi)  // requests that arrived during the tick
                    sched_fill(t, sv, ss, bar, false, false);
                slo += m;
                if (slo >= hi) break;
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc)) goto fail;
            }
            running = true;
        }
    } else {
        bool running = false;  // a tick wrote back since the last exact admission
        uint64_t ntick = 0;    // ticks that admitted requests (counted by the stamping block)
        // The stamping block: the tick's requests complete now.
        auto stamp_tick = [&]() {
            if (threadIdx.x == 0) stamp_s = now_ticks() - ld_coh(&ss->t0);
            __syncthreads();
            for (uint64_t r = tick_req[0] + threadIdx.x; r < tick_req[1]; r += blockDim.x) {
                sv.complete[r] = stamp_s;
                sv.tick_of[r] = (uint32_t)ntick;
            }
            __syncthreads();
            if (threadIdx.x == 0) ss->ticks = ntick + 1;
        };
        for (;;) {
            if (!split) {  // one block: it admits each tick itself
                sched_begin(t, !running);
                sched_fill(t, sv, ss, bar, !running, true);
                sched_publish(sv, ss);
            }
            if (!grid_sync(bar, gridDim.x, t.spin_ticks, running, rc)) goto fail;
            running = false;
            if (ld_agent(&ss->done)) break;
            // The tick's fields are read now: block 0 publishes the next
            // tick's while this one writes back.  The stamping block keeps
            // the tick's requests in LDS until phase R is done.
            if (blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
                tick_req[0] = ld_coh(&ss->r_lo);
                tick_req[1] = ld_coh(&ss->r_hi);
            }
            const uint64_t lo = uniform64(ld_coh(&ss->op_lo));
            const uint64_t hi = uniform64(ld_coh(&ss->op_hi));
            if (lo == hi) {  // no ops: admit again, exactly
                if (blockIdx.x == gridDim.x - 1) {
                    __syncthreads();
                    if (tick_req[0] < tick_req[1]) {  // requests of no ops complete at once
                        stamp_tick();
                        ++ntick;
                    }
                }
                continue;
            }
            for (uint64_t slo = lo;;) {
                const uint64_t m = min(submax, hi - slo);
                phase_r_range<L, G, PERSIST_THREADS, BC>(t, o, out, slo, m, gid, ng, rb,
                                                         rep_lds(), REP_LDS);
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, false)) goto fail;
                // Every result of the tick is written in phase R: its
                // requests complete here.  Phase W only brings the table up
                // to date for the next tick.  The last block stamps them.
                if (slo + m >= hi && blockIdx.x == gridDim.x - 1) stamp_tick();
                phase_w_block<L, G, RC>(t, o, slo, rb, *rc, rep_lds(), REP_LDS);
                slo += m;
                if (slo >= hi) break;
                if (!grid_sync(bar, gridDim.x, t.spin_ticks, true, rc)) goto fail;
            }
            ++ntick;
            running = true;
        }
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) ss->end = now_ticks() - ss->t0;
    return;
fail:
    if (threadIdx.x == 0)
        __hip_atomic_fetch_or(t.status, ST_TIMEOUT, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <class L>
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
    unsigned long long hist[LSLAB_CHAIN_BINS];  // buckets by chain length (slabs), last bin open
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
    // Chain-length histogram: per block in LDS, then one atomic per bin.
    __shared__ unsigned int hist_s[LSLAB_CHAIN_BINS];
    if (threadIdx.x < LSLAB_CHAIN_BINS) hist_s[threadIdx.x] = 0;
    __syncthreads();
    if (b < t.nbuckets)
        atomicAdd(&hist_s[len < LSLAB_CHAIN_BINS ? len - 1 : LSLAB_CHAIN_BINS - 1], 1u);
    // wave reduction then one atomic per wave
    for (int off = 32; off > 0; off >>= 1) {
        present += __shfl_down(present, off, 64);
        len = max(len, (unsigned long long)__shfl_down(len, off, 64));
    }
    if (__lane_id() == 0) {
        atomicAdd(&st->present, present);
        atomicMax(&st->max_chain, len);
    }
    __syncthreads();
    if (threadIdx.x < LSLAB_CHAIN_BINS && hist_s[threadIdx.x])
        atomicAdd(&st->hist[threadIdx.x], (unsigned long long)hist_s[threadIdx.x]);
}

template <class L>
__global__ void k_dump(TableDev t, typename L::K* keys, typename L::V* vals, uint64_t cap,
                       unsigned long long* counter) {
    const uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= t.nbuckets) return;
    uint64_t s = b, hops = 0;
    for (;;) {
        const uint32_t v = *valid_word<L>(t, s) & L::CAPMASK;
        for (int i = 0; i < 32; ++i) {
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


// Device-side state of a table besides its slabs: bump pointer, status bits,
// barrier, introspection results and the served loop's state.
struct DevState {
    uint64_t alloc_ctr;
    uint32_t status;
    uint32_t pad;
    uint64_t free_head;   // reclaim policy: tagged head of the free pool slabs
    uint64_t free_count;  // reclaim policy: slabs on the free list
    BarState bar;
    StatsDev stats;
    unsigned long long dump_counter;
    ServeState serve;
};

using L_M3232 = Layout<uint32_t, uint32_t, true>;
using L_M3264 = Layout<uint32_t, uint64_t, true>;
using L_M6432 = Layout<uint64_t, uint32_t, true>;
using L_M6464 = Layout<uint64_t, uint64_t, true>;
using L_S32 = Layout<uint32_t, uint32_t, false>;
using L_S64 = Layout<uint64_t, uint32_t, false>;

// What a launcher needs to know of a table.
struct LaunchCtx {
    TableDev dev;
    lslab_tuning tun;  // normalised and checked
    uint64_t submax;   // two-phase ops per sub-batch (scratch capacity)
    uint64_t run_cap;  // ops a run of batches may hold (run scratch capacity)
    int num_cus;
    // Recorded by the call's last kernel launch itself when the launch
    // completes (a completion event of the dispatch: no separate marker in
    // the stream); nullptr: none (graph capture, or calls that record their
    // own).
    hipEvent_t done_ev;
};

// Ops a run of the batched kernel may hold: the sub-batch size, at most 2^20
// (the run tables take about 130 B per op of this capacity per set).
constexpr uint64_t RUN_CAP_MAX = 1ull << 20;

// Host launchers of one layout; each returns an LSLAB_* status.
template <class L>
struct Launch {
    // n ops: the lock-free plain kernel (read_only), the locked plain kernel,
    // or (two_phase) sub-batches of the two-phase kernels.
    static int ops(const LaunchCtx& c, const OpsDev& o, const OutDev& out, uint64_t n,
                   bool two_phase, bool read_only, hipStream_t s);
    // The persistent batched kernel over nbatches batches.
    static int batches(const LaunchCtx& c, const OpsDev& o, const OutDev& out,
                       const uint64_t* offsets, uint64_t nbatches, BarState* bar, hipStream_t s);
    // The persistent served loop.
    static int serve(const LaunchCtx& c, const OpsDev& o, const OutDev& out, const ServeDev& sv,
                     ServeState* ss, BarState* bar, hipStream_t s);
    // Present keys and longest chain into *st (zeroed by the caller).
    static int stats(const LaunchCtx& c, StatsDev* st, hipStream_t s);
    // Every present key (and value) into keys/vals, up to cap; *counter
    // (zeroed by the caller) ends as the number present.
    static int dump(const LaunchCtx& c, void* keys, void* vals, uint64_t cap,
                    unsigned long long* counter, hipStream_t s);
};

extern template struct Launch<L_M3232>;
extern template struct Launch<L_M3264>;
extern template struct Launch<L_M6432>;
extern template struct Launch<L_M6464>;
extern template struct Launch<L_S32>;
extern template struct Launch<L_S64>;
