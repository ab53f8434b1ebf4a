// lslab: the C ABI declared in include/lslab_c.h (table lifetime, tuning,
// batch calls, introspection).  Kernels and their launchers live in
// kernels.hpp / launch.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>

#include "kernels.hpp"

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
    uint64_t gt_cap, submax, run_cap;
    hipEvent_t order_ev;
    int device;
    int num_cus;
    bool own_slabs;  // false: the slabs live in caller memory (create_at)
};

#define HIPCHK(x)                                   \
    do {                                            \
        hipError_t e_ = (x);                        \
        if (e_ != hipSuccess) return LSLAB_ERR_HIP; \
    } while (0)

// ---------------------------------------------------------------------------
// Static scratch of lslab_exec_batches_noalloc: the persistent batched
// kernel's grid-barrier state, combining table and final-candidate lists for
// sub-batches of at most LSLAB_NOALLOC_MAX_SUBBATCH ops, in one file-scope
// device array.  Every GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel: each
// device's address comes from hipGetSymbolAddress, looked up once per device
// at table creation with that device current.
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t NA_SUB = LSLAB_NOALLOC_MAX_SUBBATCH;
constexpr uint64_t NA_GT_CAP = 2 * NA_SUB;  // combining-table entries (power of two)
constexpr uint64_t NA_BLOCKS = LSLAB_NOALLOC_MAX_BLOCKS;
constexpr uint64_t na_align(uint64_t x) { return (x + 255) & ~255ull; }
constexpr uint64_t NA_BAR_OFF = 0;
constexpr uint64_t NA_GTK_OFF = na_align(NA_BAR_OFF + sizeof(BarState));
constexpr uint64_t NA_GTH_OFF = na_align(NA_GTK_OFF + (NA_GT_CAP + 1) * sizeof(uint64_t));
constexpr uint64_t NA_REP_OFF = na_align(NA_GTH_OFF + (NA_GT_CAP + 1) * sizeof(uint32_t));
constexpr uint64_t NA_REP_LEN = NA_SUB + NA_BLOCKS * PERSIST_THREADS;
// two sets of run tables of the batched kernel (runs of at most NA_SUB ops)
constexpr uint64_t NA_RK_LEN = NA_GT_CAP + 2;
constexpr uint64_t NA_RC_LEN = 4 * NA_SUB + RUN_MAX_BATCHES;
constexpr uint64_t na_clr_len() {
    uint64_t len = 0;
    for (uint64_t nb = 1; nb <= NA_BLOCKS; ++nb)
        len = len > nb * run_clr_stride(NA_SUB, nb) ? len : nb * run_clr_stride(NA_SUB, nb);
    return len;
}
constexpr uint64_t NA_CLR_LEN = na_clr_len();
constexpr uint64_t NA_SET_BYTES = na_align(NA_RK_LEN * sizeof(RunKey)) +
                                  na_align(NA_RC_LEN * sizeof(RunChain)) +
                                  na_align(NA_CLR_LEN * sizeof(uint32_t));
constexpr uint64_t NA_SET_OFF = na_align(NA_REP_OFF + NA_REP_LEN * sizeof(RepEnt));
constexpr uint64_t NA_PRED_OFF = NA_SET_OFF + 2 * NA_SET_BYTES;
constexpr uint64_t NA_BYTES = na_align(NA_PRED_OFF + NA_SUB * sizeof(uint32_t));
static_assert(NA_BYTES <= 32ull << 20, "static scratch stays under 32 MiB");
}  // namespace

static __device__ __align__(256) unsigned char scratch[NA_BYTES];

// Device address of `scratch` per device (nullptr until looked up).
static constexpr int NA_MAX_DEVICES = 64;
static std::mutex na_mutex;
static char* na_ptr[NA_MAX_DEVICES];

// Looks up this device's copy of `scratch` (the device must be current).
static int noalloc_setup(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return LSLAB_ERR_INVALID;
    std::lock_guard<std::mutex> g(na_mutex);
    if (na_ptr[device]) return LSLAB_OK;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return LSLAB_ERR_HIP;
    na_ptr[device] = static_cast<char*>(p);
    return LSLAB_OK;
}

static char* noalloc_ptr(int device) {
    if (device < 0 || device >= NA_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> g(na_mutex);
    return na_ptr[device];
}

// Empties the static combining table and one set of run tables (keys all
// ones, heads and bitmaps 0) before a call: a kernel, not a memset, so the
// call can be captured into a graph.
__global__ void k_noalloc_clear(uint64_t* gt_key, uint32_t* gt_head, RunKey* rk, RunChain* rc) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < NA_RC_LEN;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (gt_key && i < NA_GT_CAP + 1) {
            gt_key[i] = GT_EMPTY;
            gt_head[i] = 0;
        }
        if (i < NA_RK_LEN) rk[i] = RunKey{GT_EMPTY, {0, 0}, 0};
        rc[i] = RunChain{GT_EMPTY, 0, 0};
    }
}

// Set p of the static run tables at base.
static RunTabs noalloc_set(char* base, int p) {
    char* s = base + NA_SET_OFF + (uint64_t)p * NA_SET_BYTES;
    RunTabs r;
    r.rk = reinterpret_cast<RunKey*>(s);
    r.rc = reinterpret_cast<RunChain*>(s + na_align(NA_RK_LEN * sizeof(RunKey)));
    r.clr = reinterpret_cast<uint32_t*>(s + na_align(NA_RK_LEN * sizeof(RunKey)) +
                                        na_align(NA_RC_LEN * sizeof(RunChain)));
    return r;
}

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

static void normalise(lslab_tuning& tu) {
    if (tu.lanes_per_op == 0) tu.lanes_per_op = 4;
    if (tu.block_threads == 0) tu.block_threads = 256;
    if (tu.max_subbatch == 0) tu.max_subbatch = 1ull << 20;
    if (tu.combine == 0) tu.combine = LSLAB_COMBINE_BLOCK;
    if (tu.hash == 0) tu.hash = LSLAB_HASH_IDENTITY;
    if (tu.reclaim == 0) tu.reclaim = LSLAB_RECLAIM_OFF;
    if (tu.spin_ms == 0) tu.spin_ms = (int)(SPIN_LIMIT_TICKS / TICKS_PER_MS);
}

static bool tuning_ok(const lslab_tuning& tu) {
    const int g = tu.lanes_per_op;
    if (g != 1 && g != 2 && g != 4 && g != 8 && g != 16 && g != 32) return false;
    if (tu.engine != LSLAB_ENGINE_PLAIN && tu.engine != LSLAB_ENGINE_TWO_PHASE) return false;
    if (tu.block_threads < 64 || tu.block_threads > 256 || (tu.block_threads % 64)) return false;
    if (tu.max_blocks < 0) return false;
    if (tu.max_subbatch < 64 || tu.max_subbatch > (1ull << 30)) return false;
    if (tu.combine != LSLAB_COMBINE_WAVE && tu.combine != LSLAB_COMBINE_BLOCK) return false;
    if (tu.hash != LSLAB_HASH_IDENTITY && tu.hash != LSLAB_HASH_MIX) return false;
    if (tu.reclaim != LSLAB_RECLAIM_OFF && tu.reclaim != LSLAB_RECLAIM_ON) return false;
    if (tu.spin_ms < 10 || tu.spin_ms > 3600000) return false;
    if (tu.slot_heads != 0 && tu.slot_heads != LSLAB_SLOT_HEADS_ON &&
        tu.slot_heads != LSLAB_SLOT_HEADS_OFF)
        return false;
    if (tu.runs != 0 && tu.runs != LSLAB_RUNS_ON && tu.runs != LSLAB_RUNS_OFF) return false;
    return true;
}

static uint64_t slot_heads_len(const lslab_table* t) { return (t->nbuckets + t->pool) * 32; }

// Whether the tuning asks for per-slot chain heads (and the slot indices fit
// below LOC_G).
static bool slot_heads_wanted(const lslab_table* t, const lslab_tuning& tu) {
    const bool on = tu.slot_heads == LSLAB_SLOT_HEADS_ON ||
                    (tu.slot_heads == 0 && tu.engine == LSLAB_ENGINE_TWO_PHASE);
    return on && slot_heads_len(t) <= LOC_G;
}

// Allocates (zeroed) or frees the per-slot chain heads as the tuning asks.
// Freeing waits for the device: queued calls may still use them.
static int sync_slot_heads(lslab_table* t) {
    const bool want = slot_heads_wanted(t, t->tun);
    if (want && !t->dev.loc_head) {
        uint32_t* h = nullptr;
        if (hipMalloc(&h, slot_heads_len(t) * sizeof(uint32_t)) != hipSuccess)
            return LSLAB_ERR_OOM;
        if (hipMemset(h, 0, slot_heads_len(t) * sizeof(uint32_t)) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            (void)hipFree(h);
            return LSLAB_ERR_HIP;
        }
        t->dev.loc_head = h;
    } else if (!want && t->dev.loc_head) {
        if (hipDeviceSynchronize() != hipSuccess) return LSLAB_ERR_HIP;
        (void)hipFree(t->dev.loc_head);
        t->dev.loc_head = nullptr;
    }
    return LSLAB_OK;
}

static void free_table(lslab_table* t) {
    if (!t) return;
    if (t->own_slabs) (void)hipFree(t->dev.slabs);
    (void)hipFree(t->dev.locks);
    (void)hipFree(t->state);
    (void)hipFree(t->dev.gt_key);
    (void)hipFree(t->dev.gt_head);
    (void)hipFree(t->dev.rep);
    (void)hipFree(t->dev.rep_cnt);
    for (RunTabs& r : t->dev.run) {
        (void)hipFree(r.rk);
        (void)hipFree(r.rc);
        (void)hipFree(r.clr);
    }
    (void)hipFree(t->dev.run_pred);
    (void)hipFree(t->dev.loc_head);
    if (t->order_ev) (void)hipEventDestroy(t->order_ev);
    delete t;
}

// Blocks of a k_phase_r launch over up to submax ops: at least 2 groups per
// block (64 threads, 32 lanes per op).
static uint64_t rep_cnt_len(uint64_t submax) { return submax / 2 + 1; }

// Final-candidate list entries: submax for a k_phase_r launch; up to submax
// + (groups of the grid) for a persistent launch, whose blocks keep fixed
// regions (one block of PERSIST_THREADS per CU, 1 lane per op at most).
static uint64_t rep_len(const lslab_table* t) {
    return t->tun.max_subbatch + (uint64_t)t->num_cus * PERSIST_THREADS;
}

// Run tables of the batched kernel, two sets: key table (rk_cap + 2 entries),
// chain tables (4 * run_cap + RUN_MAX_BATCHES entries) and per-block clear
// lists (as many entries as pass A can hand the blocks of any grid up to the
// CU count: run_clr_stride); and the per-op predecessors (run_cap).
static uint64_t rk_cap_of(uint64_t run_cap) {
    uint64_t cap = 1;
    while (cap < 2 * run_cap) cap <<= 1;
    return cap;
}
static uint64_t run_clr_len(const lslab_table* t, uint64_t run_cap) {
    uint64_t len = 0;
    for (uint64_t nb = 1; nb <= (uint64_t)t->num_cus; ++nb)
        len = std::max(len, nb * run_clr_stride(run_cap, nb));
    return len;
}

// Device memory the library holds for t (caller-provided slab memory excluded).
static uint64_t table_device_bytes(const lslab_table* t) {
    return (t->own_slabs ? (t->nbuckets + t->pool) * t->slab_bytes : 0) + t->nbuckets * 4 +
           sizeof(DevState) + (t->gt_cap + 1) * 12 + rep_len(t) * sizeof(RepEnt) +
           rep_cnt_len(t->submax) * 4 + t->run_cap * 4 +
           2 * ((rk_cap_of(t->run_cap) + 2) * sizeof(RunKey) +
                (4 * t->run_cap + RUN_MAX_BATCHES) * sizeof(RunChain) +
                run_clr_len(t, t->run_cap) * 4) +
           (t->dev.loc_head ? slot_heads_len(t) * 4 : 0);
}

// Empties run tables (setup only; the kernels leave them empty after every run).
__global__ void k_init_run(RunKey* rk, uint64_t nk, RunChain* rc, uint64_t nc) {
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nk + nc;
         i += (uint64_t)gridDim.x * blockDim.x) {
        if (i < nk) rk[i] = RunKey{GT_EMPTY, {0, 0}, 0};
        else rc[i - nk] = RunChain{GT_EMPTY, 0, 0};
    }
}

// Sizes the two-phase scratch for t->tun.max_subbatch.  The new buffers are
// allocated before the old ones are freed, so on failure the table keeps its
// previous, consistent scratch.
static int alloc_scratch(lslab_table* t) {
    uint64_t cap = 1;
    while (cap < 2 * t->tun.max_subbatch) cap <<= 1;
    const uint64_t run_cap = std::min<uint64_t>(t->tun.max_subbatch, RUN_CAP_MAX);
    const uint64_t rk_cap = rk_cap_of(run_cap);
    const uint64_t rc_len = 4 * run_cap + RUN_MAX_BATCHES;
    uint64_t* gt_key = nullptr;
    uint32_t* gt_head = nullptr;
    RepEnt* rep = nullptr;
    uint32_t* rep_cnt = nullptr;
    RunTabs run[2] = {};
    uint32_t* run_pred = nullptr;
    bool ok = hipMalloc(&gt_key, (cap + 1) * sizeof(uint64_t)) == hipSuccess &&
              hipMalloc(&gt_head, (cap + 1) * sizeof(uint32_t)) == hipSuccess &&
              hipMalloc(&rep, rep_len(t) * sizeof(RepEnt)) == hipSuccess &&
              hipMalloc(&rep_cnt, rep_cnt_len(t->tun.max_subbatch) * sizeof(uint32_t)) ==
                  hipSuccess &&
              hipMalloc(&run_pred, run_cap * sizeof(uint32_t)) == hipSuccess;
    for (RunTabs& r : run) {
        ok = ok && hipMalloc(&r.rk, (rk_cap + 2) * sizeof(RunKey)) == hipSuccess &&
             hipMalloc(&r.rc, rc_len * sizeof(RunChain)) == hipSuccess &&
             hipMalloc(&r.clr, run_clr_len(t, run_cap) * sizeof(uint32_t)) == hipSuccess;
        if (ok) {
            hipLaunchKernelGGL(k_init_run, dim3(1024), dim3(256), 0, 0, r.rk, rk_cap + 2, r.rc,
                               rc_len);
            ok = hipGetLastError() == hipSuccess;
        }
    }
    ok = ok && hipMemset(gt_key, 0xff, (cap + 1) * sizeof(uint64_t)) == hipSuccess &&
         hipMemset(gt_head, 0, (cap + 1) * sizeof(uint32_t)) == hipSuccess &&
         hipDeviceSynchronize() == hipSuccess;
    auto release = [](uint64_t* gk, uint32_t* gh, RepEnt* rp, uint32_t* rcn, RunTabs* rn,
                      uint32_t* rpd) {
        (void)hipFree(gk);
        (void)hipFree(gh);
        (void)hipFree(rp);
        (void)hipFree(rcn);
        for (int i = 0; i < 2; ++i) {
            (void)hipFree(rn[i].rk);
            (void)hipFree(rn[i].rc);
            (void)hipFree(rn[i].clr);
        }
        (void)hipFree(rpd);
    };
    if (!ok) {
        release(gt_key, gt_head, rep, rep_cnt, run, run_pred);
        return LSLAB_ERR_OOM;
    }
    release(t->dev.gt_key, t->dev.gt_head, t->dev.rep, t->dev.rep_cnt, t->dev.run,
            t->dev.run_pred);
    t->dev.gt_key = gt_key;
    t->dev.gt_head = gt_head;
    t->dev.rep = rep;
    t->dev.rep_cnt = rep_cnt;
    t->dev.run[0] = run[0];
    t->dev.run[1] = run[1];
    t->dev.run_pred = run_pred;
    t->dev.rc_cap = 4 * run_cap;
    t->dev.rk_mask = (uint32_t)(rk_cap - 1);
    t->gt_cap = cap;
    t->submax = t->tun.max_subbatch;
    t->run_cap = run_cap;
    t->dev.gt_mask = (uint32_t)(cap - 1);
    return LSLAB_OK;
}

static uint64_t layout_slab_bytes(TableType type) {
    switch (type) {
        case T_MAP_32_32: return L_M3232::FOOTPRINT;
        case T_MAP_32_64: return L_M3264::FOOTPRINT;
        case T_MAP_64_32: return L_M6432::FOOTPRINT;
        case T_MAP_64_64: return L_M6464::FOOTPRINT;
        case T_SET_32: return L_S32::FOOTPRINT;
        case T_SET_64: return L_S64::FOOTPRINT;
    }
    return 0;
}

static bool table_type(bool is_map, int kb, int vb, TableType* ty) {
    if (is_map) {
        if (kb == 4 && vb == 4) *ty = T_MAP_32_32;
        else if (kb == 4 && vb == 8) *ty = T_MAP_32_64;
        else if (kb == 8 && vb == 4) *ty = T_MAP_64_32;
        else if (kb == 8 && vb == 8) *ty = T_MAP_64_64;
        else return false;
    } else {
        if (kb == 4) *ty = T_SET_32;
        else if (kb == 8) *ty = T_SET_64;
        else return false;
    }
    return true;
}

// Bytes of the slab array (bucket heads, then the pool); 0 if invalid.
static uint64_t slab_array_bytes(TableType ty, unsigned n_log2, uint64_t pool) {
    if (n_log2 > 30) return 0;
    const uint64_t nslabs = (1ull << n_log2) + pool;
    if (nslabs >= (1ull << 32)) return 0;
    return nslabs * layout_slab_bytes(ty);
}

static int create_common(lslab_table** out, TableType type, bool is_map, int kb, int vb,
                         unsigned n_log2, uint64_t pool, const lslab_tuning* tuning,
                         void* slab_mem = nullptr, uint64_t slab_mem_bytes = 0) {
    if (!out) return LSLAB_ERR_INVALID;
    *out = nullptr;
    if (n_log2 > 30) return LSLAB_ERR_INVALID;
    if (slab_mem && ((uintptr_t)slab_mem % 256 != 0 ||
                     slab_mem_bytes < slab_array_bytes(type, n_log2, pool)))
        return LSLAB_ERR_INVALID;
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
        t->slab_bytes = decltype(l)::FOOTPRINT;
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
    // Set-up of lslab_exec_batches_noalloc: this device's static scratch
    // address, looked up once per device with the device current.
    if (int rc0 = noalloc_setup(t->device)) {
        delete t;
        return rc0;
    }
    int rc = LSLAB_OK;
    const uint64_t slab_total = nslabs * t->slab_bytes;
    t->own_slabs = slab_mem == nullptr;
    if (slab_mem) t->dev.slabs = static_cast<char*>(slab_mem);
    if ((t->own_slabs && hipMalloc(&t->dev.slabs, slab_total) != hipSuccess) ||
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
    if (rc == LSLAB_OK) rc = sync_slot_heads(t);
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
    t->dev.free_head = &t->state->free_head;
    t->dev.free_count = &t->state->free_count;
    t->dev.nbuckets = t->nbuckets;
    t->dev.pool = pool;
    t->dev.mask = (uint32_t)(t->nbuckets - 1);
    t->dev.hash_mix = t->tun.hash == LSLAB_HASH_MIX;
    t->dev.spin_ticks = (uint64_t)t->tun.spin_ms * TICKS_PER_MS;
    t->device_bytes = table_device_bytes(t);
    *out = t;
    return LSLAB_OK;
}

extern "C" int lslab_map_create(lslab_table** out, int key_bytes, int value_bytes, unsigned n_log2,
                                uint64_t pool_slabs, const lslab_tuning* tuning) {
    TableType ty;
    if (!table_type(true, key_bytes, value_bytes, &ty)) return LSLAB_ERR_INVALID;
    return create_common(out, ty, true, key_bytes, value_bytes, n_log2, pool_slabs, tuning);
}

extern "C" int lslab_set_create(lslab_table** out, int key_bytes, unsigned n_log2,
                                uint64_t pool_slabs, const lslab_tuning* tuning) {
    TableType ty;
    if (!table_type(false, key_bytes, 0, &ty)) return LSLAB_ERR_INVALID;
    return create_common(out, ty, false, key_bytes, 0, n_log2, pool_slabs, tuning);
}

extern "C" int lslab_table_info(const lslab_table* t, int* is_map, int* key_bytes,
                                int* value_bytes, int* device) {
    if (!t) return LSLAB_ERR_INVALID;
    if (is_map) *is_map = t->is_map ? 1 : 0;
    if (key_bytes) *key_bytes = t->key_bytes;
    if (value_bytes) *value_bytes = t->is_map ? t->val_bytes : 0;
    if (device) *device = t->device;
    return LSLAB_OK;
}

extern "C" uint64_t lslab_slab_bytes(int is_map, int key_bytes, int value_bytes, unsigned n_log2,
                                     uint64_t pool_slabs) {
    TableType ty;
    if (!table_type(is_map != 0, key_bytes, value_bytes, &ty)) return 0;
    return slab_array_bytes(ty, n_log2, pool_slabs);
}

extern "C" int lslab_map_create_at(lslab_table** out, int key_bytes, int value_bytes,
                                   unsigned n_log2, uint64_t pool_slabs, void* slab_mem,
                                   uint64_t slab_mem_bytes, const lslab_tuning* tuning) {
    TableType ty;
    if (!slab_mem || !table_type(true, key_bytes, value_bytes, &ty)) return LSLAB_ERR_INVALID;
    return create_common(out, ty, true, key_bytes, value_bytes, n_log2, pool_slabs, tuning,
                         slab_mem, slab_mem_bytes);
}

extern "C" int lslab_set_create_at(lslab_table** out, int key_bytes, unsigned n_log2,
                                   uint64_t pool_slabs, void* slab_mem, uint64_t slab_mem_bytes,
                                   const lslab_tuning* tuning) {
    TableType ty;
    if (!slab_mem || !table_type(false, key_bytes, 0, &ty)) return LSLAB_ERR_INVALID;
    return create_common(out, ty, false, key_bytes, 0, n_log2, pool_slabs, tuning, slab_mem,
                         slab_mem_bytes);
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
    if (tu.hash == 0) tu.hash = t->tun.hash;  // 0 keeps the table's hash
    if (tu.reclaim == 0) tu.reclaim = t->tun.reclaim;
    normalise(tu);
    if (!tuning_ok(tu)) return LSLAB_ERR_INVALID;
    if (tu.hash != t->tun.hash) return LSLAB_ERR_INVALID;  // the table's buckets depend on it
    if (tu.reclaim != t->tun.reclaim) return LSLAB_ERR_INVALID;  // fixed with the table
    const bool resize = tu.max_subbatch != t->tun.max_subbatch;
    const lslab_tuning prev = t->tun;
    auto restore = [&]() {
        t->tun = prev;
        t->dev.spin_ticks = (uint64_t)prev.spin_ms * TICKS_PER_MS;
    };
    t->tun = tu;
    t->dev.spin_ticks = (uint64_t)tu.spin_ms * TICKS_PER_MS;
    int rc = sync_slot_heads(t);
    if (rc == LSLAB_OK && resize) {
        // The scratch may still be in use by queued calls of this table.
        rc = hipDeviceSynchronize() == hipSuccess ? alloc_scratch(t) : LSLAB_ERR_HIP;
    }
    if (rc) {  // the previous scratch stays, sized for the previous tuning
        restore();
        (void)sync_slot_heads(t);  // heads as the previous tuning had them
    }
    t->device_bytes = table_device_bytes(t);
    return rc;
}

extern "C" int lslab_get_tuning(const lslab_table* t, lslab_tuning* out) {
    if (!t || !out) return LSLAB_ERR_INVALID;
    *out = t->tun;
    return LSLAB_OK;
}

// Orders this call after the table's previous call, whatever its stream: the
// call waits on the table's order event, and its last kernel launch records
// that event on completion (call_ctx), which costs the stream no marker of
// its own (a separate hipEventRecord added ~4 us of device time to a call).
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
static int end_call(lslab_table*, hipStream_t) {
    HIPCHK(hipGetLastError());
    return LSLAB_OK;
}

static LaunchCtx ctx(const lslab_table* t) {
    LaunchCtx c;
    c.dev = t->dev;
    c.tun = t->tun;
    c.submax = t->submax;
    c.run_cap = t->run_cap;
    c.num_cus = t->num_cus;
    c.done_ev = nullptr;
    return c;
}

// Launch context of a call on stream s whose last kernel records the table's
// order event (none while s is being captured).
static LaunchCtx call_ctx(const lslab_table* t, hipStream_t s) {
    LaunchCtx c = ctx(t);
    if (!capturing(s)) c.done_ev = t->order_ev;
    return c;
}

// Runs n ops through the plain kernel (locked or lock-free) or the two-phase
// kernels, depending on `engine`.
static int run_ops(lslab_table* t, const OpsDev& o, const OutDev& out, uint64_t n, int engine,
                   bool read_only, hipStream_t s) {
    if (n == 0) return LSLAB_OK;
    int rc = begin_call(t, s);
    if (rc) return rc;
    rc = dispatch(t, [&](auto l) {
        return Launch<decltype(l)>::ops(call_ctx(t, s), o, out, n,
                                        engine == LSLAB_ENGINE_TWO_PHASE, read_only, s);
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
        return Launch<decltype(l)>::batches(call_ctx(t, stream), o, d, offsets, (uint64_t)nbatches,
                                            bar, stream);
    });
    if (rc) return rc;
    return end_call(t, stream);
}

extern "C" int lslab_exec_batches_noalloc(lslab_table* t, size_t nbatches, const uint64_t* offsets,
                                          const uint8_t* ops, const void* keys,
                                          const void* values, void* out, hipStream_t stream) {
    if (!t || (nbatches && (!offsets || !ops || !keys || !out))) return LSLAB_ERR_INVALID;
    if (t->is_map && nbatches && !values) return LSLAB_ERR_INVALID;
    // Sizes the static scratch cannot hold are refused before any device work.
    if (t->submax > NA_SUB) return LSLAB_ERR_INVALID;
    const uint64_t blocks = t->tun.max_blocks > 0
                                ? std::min<uint64_t>((uint64_t)t->tun.max_blocks, t->num_cus)
                                : (uint64_t)t->num_cus;
    if (blocks > NA_BLOCKS) return LSLAB_ERR_INVALID;
    if (nbatches == 0) return LSLAB_OK;
    char* base = noalloc_ptr(t->device);
    if (!base) return LSLAB_ERR_HIP;  // set-up did not run on this device
    int rc = begin_call(t, stream);
    if (rc) return rc;
    LaunchCtx c = call_ctx(t, stream);
    c.dev.gt_key = reinterpret_cast<uint64_t*>(base + NA_GTK_OFF);
    c.dev.gt_head = reinterpret_cast<uint32_t*>(base + NA_GTH_OFF);
    c.dev.gt_mask = (uint32_t)(NA_GT_CAP - 1);
    c.dev.rep = reinterpret_cast<RepEnt*>(base + NA_REP_OFF);
    c.dev.rep_cnt = nullptr;  // only the per-call kernel pair uses it
    c.dev.run[0] = noalloc_set(base, 0);
    c.dev.run[1] = noalloc_set(base, 1);
    c.dev.rk_mask = (uint32_t)(NA_GT_CAP - 1);
    c.dev.rc_cap = 4 * NA_SUB;
    c.dev.run_pred = reinterpret_cast<uint32_t*>(base + NA_PRED_OFF);
    BarState* bar = reinterpret_cast<BarState*>(base + NA_BAR_OFF);
    hipLaunchKernelGGL(k_noalloc_clear, dim3(256), dim3(256), 0, stream, c.dev.gt_key,
                       c.dev.gt_head, c.dev.run[0].rk, c.dev.run[0].rc);
    hipLaunchKernelGGL(k_noalloc_clear, dim3(256), dim3(256), 0, stream, nullptr, nullptr,
                       c.dev.run[1].rk, c.dev.run[1].rc);
    OpsDev o = make_ops(ops, 0, keys, t->key_bytes, t->is_map ? values : nullptr, t->val_bytes);
    OutDev d = t->is_map ? make_out(out, nullptr, nullptr)
                         : make_out(nullptr, nullptr, static_cast<uint8_t*>(out));
    rc = dispatch(t, [&](auto l) {
        return Launch<decltype(l)>::batches(c, o, d, offsets, (uint64_t)nbatches, bar, stream);
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
        return Launch<decltype(l)>::serve(call_ctx(t, stream), o, d, sv, ss, bar, stream);
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
    if (st & ST_BAD_INPUT) return LSLAB_ERR_INVALID;
    if (st & ST_CORRUPT) return LSLAB_ERR_CORRUPT;
    if (st & ST_TIMEOUT) return LSLAB_ERR_TIMEOUT;
    if (st & ST_POOL_EXHAUSTED) return LSLAB_ERR_POOL_EXHAUSTED;
    return LSLAB_OK;
}

extern "C" int lslab_stats_get(lslab_table* t, lslab_stats* out, hipStream_t stream) {
    if (!t || !out) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipMemsetAsync(&t->state->stats, 0, sizeof(StatsDev), stream));
    int rc = dispatch(t, [&](auto l) {
        return Launch<decltype(l)>::stats(ctx(t), &t->state->stats, stream);
    });
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    DevState hs;
    HIPCHK(hipMemcpy(&hs, t->state, sizeof(DevState), hipMemcpyDeviceToHost));
    out->buckets = t->nbuckets;
    out->pool_slabs = t->pool;
    out->slabs_used = std::min<uint64_t>(hs.alloc_ctr, t->pool) - hs.free_count;
    out->present = hs.stats.present;
    out->max_chain = hs.stats.max_chain;
    out->slab_bytes = t->slab_bytes;
    out->device_bytes = t->device_bytes;
    HIPCHK(hipEventRecord(t->order_ev, stream));
    return LSLAB_OK;
}

extern "C" int lslab_chain_histogram(lslab_table* t, uint64_t* hist, unsigned nbins,
                                     hipStream_t stream) {
    if (!t || !hist || nbins < 1 || nbins > LSLAB_CHAIN_BINS) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipMemsetAsync(&t->state->stats, 0, sizeof(StatsDev), stream));
    int rc = dispatch(t, [&](auto l) {
        return Launch<decltype(l)>::stats(ctx(t), &t->state->stats, stream);
    });
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(stream));
    StatsDev hs;
    HIPCHK(hipMemcpy(&hs, &t->state->stats, sizeof(hs), hipMemcpyDeviceToHost));
    for (unsigned i = 0; i < nbins; ++i) hist[i] = hs.hist[i];
    for (unsigned i = nbins; i < LSLAB_CHAIN_BINS; ++i) hist[nbins - 1] += hs.hist[i];
    HIPCHK(hipEventRecord(t->order_ev, stream));
    return LSLAB_OK;
}

extern "C" int lslab_dump(lslab_table* t, void* keys_out, void* vals_out, uint64_t cap,
                          uint64_t* count, hipStream_t stream) {
    if (!t || !count || (cap && !keys_out)) return LSLAB_ERR_INVALID;
    HIPCHK(hipStreamWaitEvent(stream, t->order_ev, 0));
    HIPCHK(hipMemsetAsync(&t->state->dump_counter, 0, sizeof(unsigned long long), stream));
    int rc = dispatch(t, [&](auto l) {
        return Launch<decltype(l)>::dump(ctx(t), keys_out, vals_out, cap, &t->state->dump_counter,
                                         stream);
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
