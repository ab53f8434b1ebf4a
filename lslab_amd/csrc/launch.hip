// lslab: host launchers of one table layout.  The build compiles this unit
// once per layout with -DLSLAB_LAYOUT=<layout alias>, so that the six
// layouts' kernels are separate translation units built in parallel.
#include <mutex>
#include <unordered_map>

#include <hip/hip_ext.h>

#include "kernels.hpp"

#ifndef LSLAB_LAYOUT
#error "LSLAB_LAYOUT must name the table layout this unit instantiates (e.g. L_M6464)"
#endif

namespace {

template <typename F>
int dispatch_g(int g, F&& f) {
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
int dispatch_c(int combine, F&& f) {
    if (combine == LSLAB_COMBINE_BLOCK) return f(std::true_type{});
    return f(std::false_type{});
}

// Reclaim policy as a compile-time flag.
template <class F>
int dispatch_r(int reclaim, F&& f) {
    if (reclaim == LSLAB_RECLAIM_ON) return f(std::true_type{});
    return f(std::false_type{});
}

int launched() { return hipGetLastError() == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP; }

// Blocks of the per-batch kernels for n ops: one G-lane group per op, capped
// by max_blocks.
uint64_t grid_for(const LaunchCtx& c, uint64_t n) {
    const uint64_t groups_per_block = (uint64_t)c.tun.block_threads / c.tun.lanes_per_op;
    uint64_t blocks = (n + groups_per_block - 1) / groups_per_block;
    if (blocks == 0) blocks = 1;
    if (c.tun.max_blocks > 0) blocks = std::min<uint64_t>(blocks, (uint64_t)c.tun.max_blocks);
    return std::min<uint64_t>(blocks, 1ull << 30);
}

// Grid of a persistent kernel whose blocks wait on each other at grid
// barriers: every block must be resident at once.  One block of
// PERSIST_THREADS (4 waves per SIMD) per CU: the kernels are built for that
// (launch bounds 1024 cap them at 128 VGPRs; their ~106 SGPRs allow 7 waves
// per SIMD), and the occupancy query must confirm at least one block per CU.
// max_blocks (tuning) can shrink the grid for small batches.
// The occupancy answer depends on the kernel only, so it is asked once per
// kernel and process and kept.
int persistent_grid(const LaunchCtx& c, const void* kernel, uint64_t* grid) {
    static std::mutex mu;
    static std::unordered_map<const void*, int> checked;
    int per_cu = 0;
    {
        std::lock_guard<std::mutex> g(mu);
        auto it = checked.find(kernel);
        if (it != checked.end()) per_cu = it->second;
    }
    if (per_cu == 0) {
        if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, PERSIST_THREADS, 0) !=
            hipSuccess)
            return LSLAB_ERR_HIP;
        std::lock_guard<std::mutex> g(mu);
        checked[kernel] = per_cu;
    }
    if (per_cu < 1) return LSLAB_ERR_HIP;
    uint64_t g = (uint64_t)c.num_cus;
    if (c.tun.max_blocks > 0) g = std::min<uint64_t>(g, (uint64_t)c.tun.max_blocks);
    *grid = g;
    return LSLAB_OK;
}

}  // namespace

template <class L>
int Launch<L>::ops(const LaunchCtx& c, const OpsDev& o, const OutDev& out, uint64_t n,
                   bool two_phase, bool read_only, hipStream_t s) {
    const int bt = c.tun.block_threads;
    return dispatch_g(c.tun.lanes_per_op, [&](auto gc) {
        constexpr int G = decltype(gc)::value;
        if (read_only) {
            hipExtLaunchKernelGGL((k_plain<L, G, false, false>), dim3((unsigned)grid_for(c, n)),
                                  dim3(bt), 0, s, nullptr, c.done_ev, 0, c.dev, o, out, n);
            return launched();
        }
        return dispatch_r(c.tun.reclaim, [&](auto rr) {
            constexpr bool RC = decltype(rr)::value;
            if (!two_phase) {
                hipExtLaunchKernelGGL((k_plain<L, G, true, RC>), dim3((unsigned)grid_for(c, n)),
                                      dim3(bt), 0, s, nullptr, c.done_ev, 0, c.dev, o, out, n);
                return launched();
            }
            return dispatch_c(c.tun.combine, [&](auto cc) {
                constexpr bool BC = decltype(cc)::value;
                for (uint64_t lo = 0; lo < n; lo += c.submax) {
                    const uint64_t m = std::min<uint64_t>(c.submax, n - lo);
                    const uint64_t grid = grid_for(c, m);
                    hipLaunchKernelGGL((k_phase_r<L, G, BC>), dim3((unsigned)grid), dim3(bt), 0,
                                       s, c.dev, o, out, lo, m);
                    hipExtLaunchKernelGGL((k_phase_w<L, G, RC>), dim3((unsigned)grid), dim3(bt),
                                          0, s, nullptr, lo + m >= n ? c.done_ev : nullptr, 0,
                                          c.dev, o, lo, m);
                }
                return launched();
            });
        });
    });
}

template <class L>
int Launch<L>::batches(const LaunchCtx& c, const OpsDev& o, const OutDev& out,
                       const uint64_t* offsets, uint64_t nbatches, BarState* bar, hipStream_t s) {
    return dispatch_g(c.tun.lanes_per_op, [&](auto gc) {
        constexpr int G = decltype(gc)::value;
        return dispatch_c(c.tun.combine, [&](auto cc) {
            constexpr bool BC = decltype(cc)::value;
            return dispatch_r(c.tun.reclaim, [&](auto rr) {
                constexpr bool RC = decltype(rr)::value;
                const int mode = c.tun.runs;
                const bool alone = mode != LSLAB_RUNS_ON, runs = mode != LSLAB_RUNS_OFF;
                uint64_t grid = 0, grid_r = 0;
                int rc = LSLAB_OK;
                if (alone)
                    rc = persistent_grid(
                        c, reinterpret_cast<const void*>(k_batches<L, G, BC, RC, false>), &grid);
                if (rc == LSLAB_OK && runs)
                    rc = persistent_grid(
                        c, reinterpret_cast<const void*>(k_batches<L, G, BC, RC, true>), &grid_r);
                if (rc) return rc;
                // The reset kernel sets bar->runs (found out on the device in
                // auto mode); of two launches the one that does not match it
                // returns at once.
                hipLaunchKernelGGL(k_reset_batches<L>, dim3(1), dim3(1024), 0, s, bar, offsets,
                                   nbatches, c.run_cap, mode, c.dev.status);
                if (alone)
                    hipExtLaunchKernelGGL((k_batches<L, G, BC, RC, false>), dim3((unsigned)grid),
                                          dim3(PERSIST_THREADS), 0, s, nullptr,
                                          runs ? nullptr : c.done_ev, 0, c.dev, o, out, offsets,
                                          nbatches, c.submax, c.run_cap, bar);
                if (runs)
                    hipExtLaunchKernelGGL((k_batches<L, G, BC, RC, true>), dim3((unsigned)grid_r),
                                          dim3(PERSIST_THREADS), 0, s, nullptr, c.done_ev, 0,
                                          c.dev, o, out, offsets, nbatches, c.submax, c.run_cap,
                                          bar);
                return launched();
            });
        });
    });
}

template <class L>
int Launch<L>::serve(const LaunchCtx& c, const OpsDev& o, const OutDev& out, const ServeDev& sv,
                     ServeState* ss, BarState* bar, hipStream_t s) {
    return dispatch_g(c.tun.lanes_per_op, [&](auto gc) {
        constexpr int G = decltype(gc)::value;
        return dispatch_c(c.tun.combine, [&](auto cc) {
            constexpr bool BC = decltype(cc)::value;
            return dispatch_r(c.tun.reclaim, [&](auto rr) {
                constexpr bool RC = decltype(rr)::value;
                uint64_t grid = 0;
                const int rc = persistent_grid(
                    c, reinterpret_cast<const void*>(k_serve<L, G, BC, RC>), &grid);
                if (rc) return rc;
                hipLaunchKernelGGL(k_reset_serve<L>, dim3(1), dim3(64), 0, s, ss, bar);
                hipExtLaunchKernelGGL((k_serve<L, G, BC, RC>), dim3((unsigned)grid),
                                      dim3(PERSIST_THREADS), 0, s, nullptr, c.done_ev, 0, c.dev,
                                      o, out, sv, ss, bar, c.submax);
                return launched();
            });
        });
    });
}

template <class L>
int Launch<L>::stats(const LaunchCtx& c, StatsDev* st, hipStream_t s) {
    const uint64_t grid = (c.dev.nbuckets + 255) / 256;
    hipLaunchKernelGGL(k_stats<L>, dim3((unsigned)grid), dim3(256), 0, s, c.dev, st);
    return launched();
}

template <class L>
int Launch<L>::dump(const LaunchCtx& c, void* keys, void* vals, uint64_t cap,
                    unsigned long long* counter, hipStream_t s) {
    const uint64_t grid = (c.dev.nbuckets + 255) / 256;
    hipLaunchKernelGGL(k_dump<L>, dim3((unsigned)grid), dim3(256), 0, s, c.dev,
                       static_cast<typename L::K*>(keys), static_cast<typename L::V*>(vals), cap,
                       counter);
    return launched();
}

template struct Launch<LSLAB_LAYOUT>;
