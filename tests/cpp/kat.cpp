// The reference's known-answer tests, restated over the header-only C++ API:
//   GetPutTest                 test/slabunified_test.cu:42-89
//   GetPutTestGroupAllocator   test/slabunified_test.cu:156-205 (a pool of a
//                              different size stands in for the foreign allocator)
//   ContainsInsertRemoveTest   test/set_test.cu:65-106
// GetPutTest also runs on a table created with lslab::mix_hash (the
// reference's Hash template parameter with a non-identity functor).
// Exit code 0 = all passed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include <lslab/map.h>
#include <lslab/set.h>

#define ASSERT(x, y) \
    if (!(x)) throw std::runtime_error((y));

#define HIPOK(x)                                                                \
    do {                                                                        \
        hipError_t e = (x);                                                     \
        if (e != hipSuccess) throw std::runtime_error(hipGetErrorString(e));    \
    } while (0)

template <typename T>
struct dvec {
    T* p = nullptr;
    size_t n;
    explicit dvec(size_t n_) : n(n_) { HIPOK(hipMalloc(&p, n * sizeof(T))); }
    ~dvec() { (void)hipFree(p); }
    void from(const std::vector<T>& h) { HIPOK(hipMemcpy(p, h.data(), n * sizeof(T), hipMemcpyHostToDevice)); }
    std::vector<T> to() const {
        std::vector<T> h(n);
        HIPOK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
    std::vector<uint8_t> bytes() const {
        std::vector<uint8_t> h(n * sizeof(T));
        HIPOK(hipMemcpy(h.data(), p, n * sizeof(T), hipMemcpyDeviceToHost));
        return h;
    }
};

template <typename Hash = lslab::hash<int>>
static void get_put(uint64_t pool, int engine) {
    constexpr int size = 1024;
    lslab_tuning tu{};
    tu.engine = engine;
    lslab::map<int, int, Hash> m(10, pool, &tu);
    std::vector<int> hk(size);
    for (int i = 0; i < size; ++i) hk[i] = i + 1;  // thrust::sequence(.., 1)
    dvec<int> keys(size), values(size), res1(size);
    dvec<lslab::found_value<int>> res2(size);
    keys.from(hk);
    for (int rep = 0; rep < 100; rep++) {
        values.from(std::vector<int>(size, rep));
        m.put(keys.p, values.p, res1.p, size);
        HIPOK(hipDeviceSynchronize());
        auto r1 = res1.to();
        if (rep != 0)
            for (int r : r1) ASSERT(r == rep - 1, "put result error");
        m.get(keys.p, res2.p, size);
        HIPOK(hipDeviceSynchronize());
        auto r2 = res2.to();
        for (int i = 0; i < size; ++i)
            ASSERT(r2[i].first && r2[i].second == rep,
                   "get result error key " + std::to_string(i + 1) + " rep " + std::to_string(rep));
    }
    ASSERT(m.status() == LSLAB_OK, "device status");
}

static void contains_insert_remove(int engine) {
    constexpr int size = 1024;
    lslab_tuning tu{};
    tu.engine = engine;
    lslab::set<int> s(10, 10000, &tu);
    std::vector<int> hk(size);
    for (int i = 0; i < size; ++i) hk[i] = i + 1;
    dvec<int> keys(size);
    dvec<bool> res1(size), res2(size);
    keys.from(hk);
    for (int rep = 0; rep < 100; rep++) {
        s.insert(keys.p, res1.p, size);
        s.contains(keys.p, res2.p, size);
        HIPOK(hipDeviceSynchronize());
        for (uint8_t b : res2.bytes()) ASSERT(b == 1, "contains after insert");
        s.remove(keys.p, res1.p, size);
        s.contains(keys.p, res2.p, size);
        HIPOK(hipDeviceSynchronize());
        for (uint8_t b : res2.bytes()) ASSERT(b == 0, "contains after remove");
    }
    ASSERT(s.status() == LSLAB_OK, "device status");
}

int main() {
    try {
        for (int engine = 0; engine < 2; ++engine) {
            get_put(10000, engine);
            std::printf("GetPutTest engine %d ok\n", engine);
            get_put(2097152 / 256, engine);
            std::printf("GetPutTestGroupAllocator engine %d ok\n", engine);
            contains_insert_remove(engine);
            std::printf("ContainsInsertRemoveTest engine %d ok\n", engine);
            get_put<lslab::mix_hash<int>>(10000, engine);
            std::printf("GetPutTest mix_hash engine %d ok\n", engine);
        }
    } catch (const std::exception& e) {
        std::printf("FAILED: %s\n", e.what());
        return 1;
    }
    return 0;
}
