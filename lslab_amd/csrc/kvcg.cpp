// lslab KVCG hot/cold split: a host-memory tier for the hottest keys in
// front of a device table, for batches whose requests arrive in host memory.
//
// KVCG, the key-value system the reference implements the GPU side of
// (README.md:3, CITATION.bib:1-17), serves skewed workloads by keeping hot
// keys in a CPU cache and sending only the rest over PCIe to the GPU slab
// table.  The reference repository holds only the GPU table; this file is
// the split, written against the C ABI of the table.
//
// Ownership, not caching: while a key is hot, the host tier is its only
// authority and the table's copy of it is stale; the table is brought up to
// date when the key is demoted (its host state is written back as a put or a
// remove) and by lslab_kvcg_flush.  Every op of a batch on a hot key runs on
// the host in batch order; every other op goes to the table in one mixed call.
// A key's ops therefore run entirely on one side, so per key the results
// follow a serial order: batch order for hot keys, the table engine's order
// for cold ones.  A batch of distinct keys gives exactly the sequential
// replay's results.
//
// One call: the workers classify the ops against the hot index and gather the
// cold ones into pinned staging; the cold ops are copied to the device, run
// and copied back on the caller's stream while the workers run the hot ops;
// after the stream completes, the cold results are scattered to their ops.
// Hot ops are sharded by hot slot over the workers, each shard in batch order.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "lslab_c.h"

namespace {

// murmur3 64-bit finalizer: the hot index's hash.
inline uint64_t fmix(uint64_t k) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return k;
}

inline uint64_t ldw(const void* base, uint64_t i, int w) {
    if (w == 4) {
        uint32_t v;
        std::memcpy(&v, static_cast<const char*>(base) + i * 4, 4);
        return v;
    }
    uint64_t v;
    std::memcpy(&v, static_cast<const char*>(base) + i * 8, 8);
    return v;
}

inline void stw(void* base, uint64_t i, int w, uint64_t v) {
    if (w == 4) {
        const uint32_t x = (uint32_t)v;
        std::memcpy(static_cast<char*>(base) + i * 4, &x, 4);
    } else {
        std::memcpy(static_cast<char*>(base) + i * 8, &v, 8);
    }
}

// A fixed set of worker threads; run(f) calls f(w) for every w in [0, n),
// f(0) on the calling thread, and returns when all have returned.
class Workers {
public:
    explicit Workers(int n) : n_(n) {
        for (int w = 1; w < n; ++w) th_.emplace_back([this, w] { loop(w); });
    }
    ~Workers() {
        {
            std::lock_guard<std::mutex> g(m_);
            stop_ = true;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }
    int size() const { return n_; }

    void run(const std::function<void(int)>& f) {
        if (n_ > 1) {
            std::lock_guard<std::mutex> g(m_);
            job_ = &f;
            pending_ = n_ - 1;
            ++gen_;
        }
        cv_.notify_all();
        f(0);
        if (n_ > 1) {
            std::unique_lock<std::mutex> g(m_);
            done_.wait(g, [&] { return pending_ == 0; });
            job_ = nullptr;
        }
    }

private:
    void loop(int w) {
        uint64_t seen = 0;
        for (;;) {
            const std::function<void(int)>* job;
            {
                std::unique_lock<std::mutex> g(m_);
                cv_.wait(g, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = gen_;
                job = job_;
            }
            (*job)(w);
            std::lock_guard<std::mutex> g(m_);
            if (--pending_ == 0) done_.notify_one();
        }
    }

    int n_;
    std::vector<std::thread> th_;
    std::mutex m_;
    std::condition_variable cv_, done_;
    const std::function<void(int)>* job_ = nullptr;
    uint64_t gen_ = 0;
    int pending_ = 0;
    bool stop_ = false;
};

// State of one hot key (or one op's result): value and presence.
struct Cell {
    uint64_t val;
    uint64_t present;
};

// The hot keys: dense slots and an open-addressing index key -> slot at load
// factor <= 1/2, fronted by a bitmap of 16 bits per index entry that turns
// away most keys that are not hot before they touch the index.  Slot s
// belongs to shard s % shards (the worker that runs its ops), and its state
// lives in that shard's own array, so workers never write to a shared cache
// line.
struct HotTier {
    struct Ent {
        uint64_t key;
        uint64_t slot;  // slot + 1; 0 = empty
    };
    std::vector<uint64_t> key;
    std::vector<std::vector<Cell>> cell;  // [shard][slot / shards]
    std::vector<Ent> index;
    std::vector<uint64_t> filter;
    uint64_t imask = 0, fmask = 0;
    int shards = 1;

    void reset(uint64_t cap, int nshards) {
        key.clear();
        shards = nshards;
        cell.assign(nshards, {});
        uint64_t ic = 2;
        while (ic < 2 * cap) ic <<= 1;
        index.assign(ic, Ent{0, 0});
        imask = ic - 1;
        filter.assign(std::max<uint64_t>(ic / 4, 1), 0);  // 16 bits per entry
        fmask = filter.size() * 64 - 1;
    }
    uint64_t size() const { return key.size(); }
    Cell& at(uint64_t s) { return cell[s % shards][s / shards]; }
    const Cell& at(uint64_t s) const { return cell[s % shards][s / shards]; }
    int64_t find(uint64_t k) const {
        const uint64_t h = fmix(k);
        const uint64_t b = (h >> 20) & fmask;
        if (!((filter[b >> 6] >> (b & 63)) & 1)) return -1;
        for (uint64_t i = h & imask;; i = (i + 1) & imask) {
            const Ent& e = index[i];
            if (e.slot == 0) return -1;
            if (e.key == k) return (int64_t)e.slot - 1;
        }
    }
    // Adds k (absent) with the given state; returns its slot.
    uint64_t add(uint64_t k, bool p, uint64_t v) {
        const uint64_t s = key.size();
        key.push_back(k);
        cell[s % shards].push_back(Cell{v, p ? 1u : 0u});
        const uint64_t h = fmix(k);
        const uint64_t b = (h >> 20) & fmask;
        filter[b >> 6] |= 1ull << (b & 63);
        uint64_t i = h & imask;
        while (index[i].slot) i = (i + 1) & imask;
        index[i] = Ent{k, s + 1};
        return s;
    }
};

}  // namespace

struct lslab_kvcg {
    lslab_table* t = nullptr;
    int is_map = 0, kb = 0, vb = 0, device = 0;
    uint64_t hot_cap = 0, max_batch = 0;
    uint32_t ostride = 0;  // bytes per result: pair<bool,V> for maps, bool for sets
    HotTier hot;
    Workers* workers = nullptr;
    // pinned staging and device buffers of one slice of cold ops
    uint8_t *h_ops = nullptr, *d_ops = nullptr;
    char *h_keys = nullptr, *h_vals = nullptr, *h_out = nullptr;
    char *d_keys = nullptr, *d_vals = nullptr, *d_out = nullptr;
    // per chunk: cold op indices; per (chunk, shard): hot ops as slot << 32 |
    // index, and their results in the same order
    std::vector<std::vector<uint32_t>> cold;
    std::vector<std::vector<std::vector<uint64_t>>> hotl;
    std::vector<std::vector<std::vector<Cell>>> hres;
    std::vector<uint64_t> off;
    lslab_kvcg_stats st{};
};

namespace {

void free_kvcg(lslab_kvcg* kv) {
    (void)hipHostFree(kv->h_ops);
    (void)hipHostFree(kv->h_keys);
    (void)hipHostFree(kv->h_vals);
    (void)hipHostFree(kv->h_out);
    (void)hipFree(kv->d_ops);
    (void)hipFree(kv->d_keys);
    (void)hipFree(kv->d_vals);
    (void)hipFree(kv->d_out);
    delete kv->workers;
    delete kv;
}

// Runs staged ops [0, n) on the table (ops/keys/values already in the pinned
// staging) and copies the results into h_out; asynchronous on s.
int run_staged(lslab_kvcg* kv, uint64_t n, hipStream_t s) {
    if (n == 0) return LSLAB_OK;
    if (hipMemcpyAsync(kv->d_ops, kv->h_ops, n, hipMemcpyHostToDevice, s) != hipSuccess ||
        hipMemcpyAsync(kv->d_keys, kv->h_keys, n * kv->kb, hipMemcpyHostToDevice, s) !=
            hipSuccess)
        return LSLAB_ERR_HIP;
    int rc;
    if (kv->is_map) {
        if (hipMemcpyAsync(kv->d_vals, kv->h_vals, n * kv->vb, hipMemcpyHostToDevice, s) !=
            hipSuccess)
            return LSLAB_ERR_HIP;
        rc = lslab_map_mixed(kv->t, kv->d_ops, kv->d_keys, kv->d_vals, kv->d_out, n, s);
    } else {
        rc = lslab_set_mixed(kv->t, kv->d_ops, kv->d_keys, reinterpret_cast<uint8_t*>(kv->d_out),
                             n, s);
    }
    if (rc) return rc;
    if (hipMemcpyAsync(kv->h_out, kv->d_out, n * kv->ostride, hipMemcpyDeviceToHost, s) !=
        hipSuccess)
        return LSLAB_ERR_HIP;
    return LSLAB_OK;
}

int finish(hipStream_t s) {
    return hipStreamSynchronize(s) == hipSuccess ? LSLAB_OK : LSLAB_ERR_HIP;
}

// Result of staged op p: (found, value).
inline std::pair<bool, uint64_t> staged_result(const lslab_kvcg* kv, uint64_t p) {
    if (!kv->is_map) return {kv->h_out[p] != 0, 0};
    const char* q = kv->h_out + p * kv->ostride;
    return {q[0] != 0, ldw(q + kv->vb, 0, kv->vb)};
}

inline void write_out(const lslab_kvcg* kv, void* out, uint64_t i, bool f, uint64_t v) {
    if (!kv->is_map) {
        static_cast<uint8_t*>(out)[i] = f ? 1 : 0;
        return;
    }
    char* q = static_cast<char*>(out) + i * kv->ostride;
    q[0] = f ? 1 : 0;
    stw(q + kv->vb, 0, kv->vb, v);
}

inline uint8_t opcode(const uint8_t* ops, uint64_t i) {
    return ops[i] <= LSLAB_OP_REMOVE ? ops[i] : (uint8_t)LSLAB_OP_GET;
}

// Writes back the state of hot slots `slots` to the table (put when present,
// remove when absent), max_batch at a time; synchronous.
int write_back(lslab_kvcg* kv, const std::vector<uint64_t>& slots, hipStream_t s) {
    for (uint64_t lo = 0; lo < slots.size(); lo += kv->max_batch) {
        const uint64_t m = std::min<uint64_t>(kv->max_batch, slots.size() - lo);
        for (uint64_t j = 0; j < m; ++j) {
            const uint64_t sl = slots[lo + j];
            const Cell& c = kv->hot.at(sl);
            kv->h_ops[j] = c.present ? LSLAB_OP_PUT : LSLAB_OP_REMOVE;
            stw(kv->h_keys, j, kv->kb, kv->hot.key[sl]);
            if (kv->is_map) stw(kv->h_vals, j, kv->vb, c.val);
        }
        int rc = run_staged(kv, m, s);
        if (!rc) rc = finish(s);
        if (rc) return rc;
    }
    return LSLAB_OK;
}

// One slice of a batch (m <= max_batch ops).
int run_slice(lslab_kvcg* kv, const uint8_t* ops, const void* keys, const void* vals, void* out,
              uint64_t m, hipStream_t s) {
    Workers& W = *kv->workers;
    const int T = W.size();
    const int kb = kv->kb, vb = kv->vb;
    auto lo_of = [&](int c) { return m * (uint64_t)c / (uint64_t)T; };
    // 1. classify
    W.run([&](int c) {
        auto& cl = kv->cold[c];
        cl.clear();
        for (auto& h : kv->hotl[c]) h.clear();
        const bool any_hot = kv->hot.size() != 0;
        for (uint64_t i = lo_of(c), e = lo_of(c + 1); i < e; ++i) {
            const int64_t sl = any_hot ? kv->hot.find(ldw(keys, i, kb)) : -1;
            if (sl < 0) cl.push_back((uint32_t)i);
            else kv->hotl[c][(uint64_t)sl % (uint64_t)T].push_back((uint64_t)sl << 32 | i);
        }
    });
    uint64_t ncold = 0, nhot = 0;
    for (int c = 0; c < T; ++c) {
        kv->off[c] = ncold;
        ncold += kv->cold[c].size();
    }
    nhot = m - ncold;
    // 2. gather the cold ops into the staging
    if (ncold) {
        W.run([&](int c) {
            uint64_t p = kv->off[c];
            for (uint32_t i : kv->cold[c]) {
                const uint8_t op = opcode(ops, i);
                kv->h_ops[p] = op;
                stw(kv->h_keys, p, kb, ldw(keys, i, kb));
                if (kv->is_map) stw(kv->h_vals, p, vb, op == LSLAB_OP_PUT ? ldw(vals, i, vb) : 0);
                ++p;
            }
        });
    }
    // 3. cold ops on the device, hot ops on the workers meanwhile (results
    // into per-shard lists: ops of one output cache line may be in any shard)
    int rc = run_staged(kv, ncold, s);
    if (rc) {
        (void)finish(s);  // the hot tier stays as it was
        return rc;
    }
    if (nhot) {
        HotTier& H = kv->hot;
        const bool is_map = kv->is_map != 0;
        W.run([&](int sh) {
            std::vector<Cell>& mine = H.cell[sh];
            for (int c = 0; c < T; ++c) {
                std::vector<Cell>& res = kv->hres[c][sh];
                res.clear();
                for (uint64_t e : kv->hotl[c][sh]) {
                    const uint64_t sl = e >> 32, i = e & 0xffffffffull;
                    Cell& st = mine[sl / (uint64_t)T];
                    const uint8_t op = opcode(ops, i);
                    const uint64_t old = st.present ? st.val : 0;
                    // a set insert reports true for present and new keys alike
                    res.push_back(Cell{old, (op == LSLAB_OP_PUT && !is_map) ? 1u : st.present});
                    if (op == LSLAB_OP_PUT) {
                        st.present = 1;
                        st.val = is_map ? ldw(vals, i, vb) : 0;
                    } else if (op == LSLAB_OP_REMOVE) {
                        st.present = 0;
                        st.val = 0;
                    }
                }
            }
        });
    }
    const int frc = finish(s);
    if (frc) return frc;
    // 4. every chunk writes the results of its own ops
    W.run([&](int c) {
        uint64_t p = kv->off[c];
        for (uint32_t i : kv->cold[c]) {
            const auto r = staged_result(kv, p++);
            write_out(kv, out, i, r.first, r.second);
        }
        if (nhot) {
            for (int sh = 0; sh < T; ++sh) {
                const auto& hl = kv->hotl[c][sh];
                const auto& hr = kv->hres[c][sh];
                for (size_t k = 0; k < hl.size(); ++k)
                    write_out(kv, out, hl[k] & 0xffffffffull, hr[k].present != 0, hr[k].val);
            }
        }
    });
    kv->st.hot_ops += nhot;
    kv->st.cold_ops += ncold;
    return LSLAB_OK;
}

}  // namespace

extern "C" int lslab_kvcg_create(lslab_kvcg** out, lslab_table* t, uint64_t hot_capacity,
                                 uint64_t max_batch, int threads) {
    if (!out) return LSLAB_ERR_INVALID;
    *out = nullptr;
    int is_map, kb, vb, dev;
    if (!t || lslab_table_info(t, &is_map, &kb, &vb, &dev) != LSLAB_OK) return LSLAB_ERR_INVALID;
    if (max_batch == 0) max_batch = 1ull << 20;
    if (max_batch > (1ull << 31) || hot_capacity > (1ull << 30) || threads < 0 || threads > 256)
        return LSLAB_ERR_INVALID;
    if (threads == 0)
        threads = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    lslab_kvcg* kv = new (std::nothrow) lslab_kvcg;
    if (!kv) return LSLAB_ERR_OOM;
    kv->t = t;
    kv->is_map = is_map;
    kv->kb = kb;
    kv->vb = is_map ? vb : 0;
    kv->device = dev;
    kv->hot_cap = hot_capacity;
    kv->max_batch = max_batch;
    kv->ostride = is_map ? 2u * (uint32_t)vb : 1u;
    kv->hot.reset(hot_capacity, threads);
    int prev = 0;
    if (hipGetDevice(&prev) != hipSuccess || hipSetDevice(dev) != hipSuccess) {
        delete kv;
        return LSLAB_ERR_HIP;
    }
    const uint64_t vbytes = is_map ? max_batch * vb : 8;
    const bool ok =
        hipHostMalloc((void**)&kv->h_ops, max_batch, hipHostMallocDefault) == hipSuccess &&
        hipHostMalloc((void**)&kv->h_keys, max_batch * kb, hipHostMallocDefault) == hipSuccess &&
        hipHostMalloc((void**)&kv->h_vals, vbytes, hipHostMallocDefault) == hipSuccess &&
        hipHostMalloc((void**)&kv->h_out, max_batch * kv->ostride, hipHostMallocDefault) ==
            hipSuccess &&
        hipMalloc((void**)&kv->d_ops, max_batch) == hipSuccess &&
        hipMalloc((void**)&kv->d_keys, max_batch * kb) == hipSuccess &&
        hipMalloc((void**)&kv->d_vals, vbytes) == hipSuccess &&
        hipMalloc((void**)&kv->d_out, max_batch * kv->ostride) == hipSuccess;
    (void)hipSetDevice(prev);
    if (!ok) {
        free_kvcg(kv);
        return LSLAB_ERR_OOM;
    }
    kv->workers = new (std::nothrow) Workers(threads);
    if (!kv->workers) {
        free_kvcg(kv);
        return LSLAB_ERR_OOM;
    }
    kv->cold.resize(threads);
    kv->hotl.assign(threads, std::vector<std::vector<uint64_t>>(threads));
    kv->hres.assign(threads, std::vector<std::vector<Cell>>(threads));
    kv->off.assign(threads, 0);
    kv->st.hot_capacity = hot_capacity;
    *out = kv;
    return LSLAB_OK;
}

extern "C" int lslab_kvcg_flush(lslab_kvcg* kv, hipStream_t stream) {
    if (!kv) return LSLAB_ERR_INVALID;
    std::vector<uint64_t> all(kv->hot.size());
    for (uint64_t i = 0; i < all.size(); ++i) all[i] = i;
    return write_back(kv, all, stream);
}

extern "C" int lslab_kvcg_destroy(lslab_kvcg* kv) {
    if (!kv) return LSLAB_ERR_INVALID;
    const int rc = lslab_kvcg_flush(kv, nullptr);
    free_kvcg(kv);
    return rc;
}

extern "C" int lslab_kvcg_set_hot(lslab_kvcg* kv, const void* keys, uint64_t n,
                                  hipStream_t stream) {
    if (!kv || (n && !keys)) return LSLAB_ERR_INVALID;
    HotTier next;
    next.reset(kv->hot_cap, kv->workers->size());
    std::vector<uint64_t> promote;  // slots of `next` whose state the table holds
    for (uint64_t i = 0; i < n; ++i) {
        const uint64_t k = ldw(keys, i, kv->kb);
        if (next.find(k) >= 0) continue;  // duplicate
        if (next.size() == kv->hot_cap) return LSLAB_ERR_INVALID;
        const int64_t o = kv->hot.find(k);
        if (o >= 0) {
            const Cell& c = kv->hot.at((uint64_t)o);
            next.add(k, c.present != 0, c.val);
        } else {
            promote.push_back(next.add(k, false, 0));
        }
    }
    // demoted keys' state goes back to the table
    std::vector<uint64_t> demote;
    for (uint64_t s = 0; s < kv->hot.size(); ++s)
        if (next.find(kv->hot.key[s]) < 0) demote.push_back(s);
    int rc = write_back(kv, demote, stream);
    if (rc) return rc;
    // promoted keys' state comes from the table
    for (uint64_t lo = 0; lo < promote.size(); lo += kv->max_batch) {
        const uint64_t m = std::min<uint64_t>(kv->max_batch, promote.size() - lo);
        for (uint64_t j = 0; j < m; ++j) {
            kv->h_ops[j] = LSLAB_OP_GET;
            stw(kv->h_keys, j, kv->kb, next.key[promote[lo + j]]);
            if (kv->is_map) stw(kv->h_vals, j, kv->vb, 0);
        }
        rc = run_staged(kv, m, stream);
        if (!rc) rc = finish(stream);
        if (rc) return rc;
        for (uint64_t j = 0; j < m; ++j) {
            const auto r = staged_result(kv, j);
            next.at(promote[lo + j]) = Cell{r.second, r.first ? 1u : 0u};
        }
    }
    kv->hot = std::move(next);
    kv->st.hot_keys = kv->hot.size();
    return LSLAB_OK;
}

extern "C" int lslab_kvcg_sample_hot(lslab_kvcg* kv, const void* sample_keys, uint64_t n,
                                     uint64_t k, hipStream_t stream) {
    if (!kv || (n && !sample_keys) || k > kv->hot_cap) return LSLAB_ERR_INVALID;
    std::unordered_map<uint64_t, uint64_t> cnt;
    cnt.reserve(n < (1ull << 22) ? n : (1ull << 22));
    for (uint64_t i = 0; i < n; ++i) ++cnt[ldw(sample_keys, i, kv->kb)];
    std::vector<std::pair<uint64_t, uint64_t>> v(cnt.begin(), cnt.end());  // (key, count)
    auto hotter = [](const std::pair<uint64_t, uint64_t>& a,
                     const std::pair<uint64_t, uint64_t>& b) {
        return a.second != b.second ? a.second > b.second : a.first < b.first;
    };
    const uint64_t take = std::min<uint64_t>(k, v.size());
    std::partial_sort(v.begin(), v.begin() + take, v.end(), hotter);
    std::vector<uint64_t> keys(take);
    for (uint64_t i = 0; i < take; ++i) keys[i] = v[i].first;
    if (kv->kb == 4) {
        std::vector<uint32_t> k32(take);
        for (uint64_t i = 0; i < take; ++i) k32[i] = (uint32_t)keys[i];
        return lslab_kvcg_set_hot(kv, k32.data(), take, stream);
    }
    return lslab_kvcg_set_hot(kv, keys.data(), take, stream);
}

extern "C" int lslab_kvcg_mixed(lslab_kvcg* kv, const uint8_t* ops, const void* keys,
                                const void* values, void* out, uint64_t n, hipStream_t stream) {
    if (!kv || (n && (!ops || !keys || !out || (kv->is_map && !values)))) return LSLAB_ERR_INVALID;
    for (uint64_t lo = 0; lo < n; lo += kv->max_batch) {
        const uint64_t m = std::min<uint64_t>(kv->max_batch, n - lo);
        const int rc = run_slice(
            kv, ops + lo, static_cast<const char*>(keys) + lo * kv->kb,
            kv->is_map ? static_cast<const char*>(values) + lo * kv->vb : nullptr,
            static_cast<char*>(out) + lo * kv->ostride, m, stream);
        if (rc) return rc;
    }
    ++kv->st.batches;
    return LSLAB_OK;
}

extern "C" int lslab_kvcg_stats_get(const lslab_kvcg* kv, lslab_kvcg_stats* out) {
    if (!kv || !out) return LSLAB_ERR_INVALID;
    *out = kv->st;
    return LSLAB_OK;
}
