// Multithreaded host baseline (test / benchmark infrastructure only): the
// same op stream applied to std::unordered_map shards on the host cores.
//
// 16 * T shards chosen by the bucket bits of the key (the reference's bucket
// = key & (2^n - 1), include/lslab/map.h:95-96), one std::shared_mutex per
// shard (shared for gets, exclusive for put/remove, as the reference's
// warp_mutex modes, include/lslab/warp_mutex.h:13-59), T worker threads that
// split each batch into contiguous slices, and a barrier between batches so
// that batch b sees every effect of batches < b.  Results have the same
// meaning as the oracle's (found, value) per op.
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <memory>
#include <mutex>
#include <shared_mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

struct Shard {
    std::shared_mutex mu;
    std::unordered_map<uint64_t, uint64_t> m;
};

struct Barrier {
    std::mutex mu;
    std::condition_variable cv;
    int n, waiting = 0;
    uint64_t gen = 0;
    explicit Barrier(int n_) : n(n_) {}
    void wait() {
        std::unique_lock<std::mutex> lk(mu);
        uint64_t g = gen;
        if (++waiting == n) {
            waiting = 0;
            gen++;
            cv.notify_all();
        } else {
            cv.wait(lk, [&] { return gen != g; });
        }
    }
};

struct Mtb {
    int threads;
    uint64_t mask;
    bool is_map;
    std::vector<std::unique_ptr<Shard>> shards;
};

inline Shard& shard_of(Mtb* h, uint64_t key) {
    return *h->shards[(key & h->mask) % h->shards.size()];
}

void apply(Mtb* h, uint8_t op, uint64_t key, uint64_t val, uint8_t* f, uint64_t* o) {
    Shard& s = shard_of(h, key);
    if (op == 0) {
        std::shared_lock<std::shared_mutex> lk(s.mu);
        auto it = s.m.find(key);
        *f = it != s.m.end();
        *o = (it != s.m.end() && h->is_map) ? it->second : 0;
    } else if (op == 1) {
        std::unique_lock<std::shared_mutex> lk(s.mu);
        auto r = s.m.try_emplace(key, val);
        *f = h->is_map ? !r.second : 1;
        *o = (!r.second && h->is_map) ? r.first->second : 0;
        if (!r.second) r.first->second = val;
    } else {
        std::unique_lock<std::shared_mutex> lk(s.mu);
        auto it = s.m.find(key);
        *f = it != s.m.end();
        *o = (it != s.m.end() && h->is_map) ? it->second : 0;
        if (it != s.m.end()) s.m.erase(it);
    }
}

}  // namespace

extern "C" {

void* mtb_create(int threads, unsigned n_log2, int is_map) {
    if (threads < 1) threads = 1;
    Mtb* h = new Mtb;
    h->threads = threads;
    h->mask = (n_log2 >= 64) ? ~0ull : ((1ull << n_log2) - 1ull);
    h->is_map = is_map != 0;
    for (int i = 0; i < 16 * threads; ++i) h->shards.emplace_back(new Shard);
    return h;
}

void mtb_destroy(void* p) { delete static_cast<Mtb*>(p); }

// Sizes every shard for `total` keys spread evenly (no rehash while filling).
void mtb_reserve(void* p, uint64_t total) {
    Mtb* h = static_cast<Mtb*>(p);
    for (auto& s : h->shards) s->m.reserve(total / h->shards.size() + 1);
}

// Runs nbatches batches ([offsets[b], offsets[b+1])) and returns the wall time
// in seconds.  found_out / val_out may be NULL.
double mtb_run(void* p, uint64_t nbatches, const uint64_t* offsets, const uint8_t* ops,
               const uint64_t* keys, const uint64_t* vals, uint8_t* found_out, uint64_t* val_out) {
    Mtb* h = static_cast<Mtb*>(p);
    const int T = h->threads;
    Barrier bar(T);
    auto t0 = std::chrono::steady_clock::now();
    std::vector<std::thread> th;
    for (int w = 0; w < T; ++w) {
        th.emplace_back([&, w] {
            for (uint64_t b = 0; b < nbatches; ++b) {
                const uint64_t lo = offsets[b], hi = offsets[b + 1], n = hi - lo;
                const uint64_t a = lo + n * w / T, e = lo + n * (w + 1) / T;
                for (uint64_t i = a; i < e; ++i) {
                    uint8_t f;
                    uint64_t o;
                    apply(h, ops[i], keys[i], vals ? vals[i] : 0, &f, &o);
                    if (found_out) found_out[i] = f;
                    if (val_out) val_out[i] = o;
                }
                bar.wait();
            }
        });
    }
    for (auto& t : th) t.join();
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

uint64_t mtb_size(void* p) {
    Mtb* h = static_cast<Mtb*>(p);
    uint64_t n = 0;
    for (auto& s : h->shards) n += s->m.size();
    return n;
}

}  // extern "C"
