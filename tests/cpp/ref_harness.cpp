// The reference's benchmark harness (benchmark/benchmark.cu:18-177) restated
// over the header-only C++ API, so that its workload is measured the way it
// measures it, from C++: a map<unsigned long long, unsigned> of 2^20 buckets,
// keys 1..2^19 put with value 0 in one call, then batches of 1024 keys drawn
// uniformly from [0, 2^21] (a seeded generator; the reference draws rand()),
// first gets, then puts of value 0.  Each call is timed by wall clock from
// before the launch to after hipStreamSynchronize, as the reference times it.
// Prints one JSON line: per series the p50 / p99 latency (us) and the p50
// rate (Mops/s) over `reps` calls, after 3 warm-up calls each.
//   usage: ref_harness [reps]
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include <lslab/map.h>

#define HIPOK(x)                                                             \
    do {                                                                     \
        hipError_t e_ = (x);                                                 \
        if (e_ != hipSuccess) throw std::runtime_error(hipGetErrorString(e_)); \
    } while (0)

using K = unsigned long long;
using V = unsigned;

struct Series {
    double p50_us, p99_us, mops;
};

template <typename F>
static Series time_calls(int reps, hipStream_t s, F&& call) {
    std::vector<double> us;
    for (int i = 0; i < reps + 3; ++i) {
        HIPOK(hipStreamSynchronize(s));
        const auto t0 = std::chrono::steady_clock::now();
        call(i);
        HIPOK(hipStreamSynchronize(s));
        const auto t1 = std::chrono::steady_clock::now();
        if (i >= 3) us.push_back(std::chrono::duration<double, std::micro>(t1 - t0).count());
    }
    std::sort(us.begin(), us.end());
    const double p50 = us[us.size() / 2];
    const double p99 = us[std::min(us.size() - 1, (size_t)(0.99 * (double)us.size()))];
    return {p50, p99, 1024.0 / p50};
}

int main(int argc, char** argv) {
    try {
        const int reps = argc > 1 ? std::max(1, atoi(argv[1])) : 100;
        const size_t pop = 1u << 19, batch = 1024;
        const K range = 1ull << 21;
        hipStream_t s;
        HIPOK(hipStreamCreate(&s));

        // the batches' keys, drawn on the host and copied once
        std::mt19937_64 rng(0xBE7Cu);
        std::uniform_int_distribution<K> pick(0, range);
        std::vector<K> hk((size_t)(reps + 3) * batch);
        for (K& k : hk) k = pick(rng);
        K* keys;
        V* zeros;
        V* old;
        lslab::found_value<V>* got;
        HIPOK(hipMalloc(&keys, hk.size() * sizeof(K)));
        HIPOK(hipMemcpy(keys, hk.data(), hk.size() * sizeof(K), hipMemcpyHostToDevice));
        HIPOK(hipMalloc(&zeros, pop * sizeof(V)));
        HIPOK(hipMemset(zeros, 0, pop * sizeof(V)));
        HIPOK(hipMalloc(&old, pop * sizeof(V)));
        HIPOK(hipMalloc(&got, batch * sizeof(lslab::found_value<V>)));
        std::vector<K> hp(pop);
        for (size_t i = 0; i < pop; ++i) hp[i] = i + 1;
        K* pre;
        HIPOK(hipMalloc(&pre, pop * sizeof(K)));
        HIPOK(hipMemcpy(pre, hp.data(), pop * sizeof(K), hipMemcpyHostToDevice));

        std::string out = "{\"workload\": \"map<u64,u32> 2^20 buckets, keys 1..2^19 prefilled; "
                          "1024-key calls uniform over [0, 2^21]; C++ header API\", \"reps\": " +
                          std::to_string(reps);
        const struct {
            const char* name;
            int engine;
        } engines[] = {{"plain", LSLAB_ENGINE_PLAIN}, {"two_phase", LSLAB_ENGINE_TWO_PHASE}};
        for (const auto& e : engines) {
            lslab_tuning tu{};
            tu.engine = e.engine;
            tu.lanes_per_op = 8;
            tu.max_subbatch = pop;
            lslab::map<K, V> m(20, 10000, &tu);
            m.put(pre, zeros, old, pop, s);
            HIPOK(hipStreamSynchronize(s));
            char buf[256];
            if (e.engine == LSLAB_ENGINE_PLAIN) {
                tu.lanes_per_op = 2;  // gets: the plain engine's tuned width for one-line probes
                m.set_tuning(tu);
                const Series g = time_calls(reps, s, [&](int i) {
                    m.get(keys + (size_t)i * batch, got, batch, s);
                });
                snprintf(buf, sizeof(buf),
                         ", \"get\": {\"latency_us_p50\": %.2f, \"latency_us_p99\": %.2f, "
                         "\"mops_p50\": %.2f}",
                         g.p50_us, g.p99_us, g.mops);
                out += buf;
                tu.lanes_per_op = 8;
                m.set_tuning(tu);
            }
            const Series p = time_calls(reps, s, [&](int i) {
                m.put(keys + (size_t)i * batch, zeros, old, batch, s);
            });
            snprintf(buf, sizeof(buf),
                     ", \"put_%s\": {\"latency_us_p50\": %.2f, \"latency_us_p99\": %.2f, "
                     "\"mops_p50\": %.2f}",
                     e.name, p.p50_us, p.p99_us, p.mops);
            out += buf;
            if (m.status() != LSLAB_OK) throw std::runtime_error("device status");
        }
        printf("%s}\n", out.c_str());
        (void)hipFree(keys);
        (void)hipFree(zeros);
        (void)hipFree(old);
        (void)hipFree(got);
        (void)hipFree(pre);
        (void)hipStreamDestroy(s);
        return 0;
    } catch (const std::exception& ex) {
        fprintf(stderr, "ref_harness: %s\n", ex.what());
        return 1;
    }
}
