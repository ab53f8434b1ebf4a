// The reference's known-answer tests, restated over the header-only C++ API:
//   GetPutTest                 test/slabunified_test.cu:42-89
//   GetPutTestGroupAllocator   test/slabunified_test.cu:156-205 (the foreign
//                              allocator is a caller-owned arena: two maps'
//                              slab arrays carved from one hipMalloc)
//   ContainsInsertRemoveTest   test/set_test.cu:65-106, over device pointers and
//                              over thrust::device_vector (set.h:98-123)
//   PutRemoveTest_128B         test/slabunified_test.cu.old:386-474 (128-B keys
//   PutRemoveTest_128Bto128B   test/slabunified_test.cu.old:476-541  on lslab::wide_map)
// GetPutTest also runs on a table created with lslab::mix_hash (the
// reference's Hash template parameter with a non-identity functor).
// Exit code 0 = all passed.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include <thrust/device_vector.h>
#include <thrust/host_vector.h>

#include <lslab/map.h>
#include <lslab/set.h>
#include <lslab/wide.h>

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
static void get_put_on(lslab::map<int, int, Hash>& m) {
    constexpr int size = 1024;
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

template <typename Hash = lslab::hash<int>>
static void get_put(uint64_t pool, int engine) {
    lslab_tuning tu{};
    tu.engine = engine;
    lslab::map<int, int, Hash> m(10, pool, &tu);
    get_put_on(m);
}

// Two maps whose slab arrays share one caller arena; the first keeps its
// contents while the second runs GetPutTest, and a short buffer is refused.
static void get_put_group(uint64_t pool, int engine) {
    using map_t = lslab::map<int, int>;
    lslab_tuning tu{};
    tu.engine = engine;
    const uint64_t need = map_t::slab_bytes(10, pool);
    ASSERT(need == ((1ull << 10) + pool) * 256, "slab_bytes");
    const uint64_t stride = (need + 255) / 256 * 256;
    dvec<uint8_t> arena(2 * stride);
    bool refused = false;
    try {
        map_t bad(10, pool, lslab::slab_memory{arena.p, need - 1}, &tu);
    } catch (const lslab::error&) {
        refused = true;
    }
    ASSERT(refused, "short slab memory accepted");
    map_t other(10, pool, lslab::slab_memory{arena.p, stride}, &tu);
    constexpr int size = 1024;
    std::vector<int> hk(size), hv(size);
    for (int i = 0; i < size; ++i) hk[i] = -(i + 1), hv[i] = 7 * i;
    dvec<int> keys(size), values(size), old(size);
    dvec<lslab::found_value<int>> res(size);
    keys.from(hk);
    values.from(hv);
    other.put(keys.p, values.p, old.p, size);
    {
        map_t m(10, pool, lslab::slab_memory{arena.p + stride, stride}, &tu);
        get_put_on(m);
    }
    other.get(keys.p, res.p, size);
    HIPOK(hipDeviceSynchronize());
    auto r = res.to();
    for (int i = 0; i < size; ++i) ASSERT(r[i].first && r[i].second == 7 * i, "arena neighbour");
    ASSERT(other.status() == LSLAB_OK, "device status");
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

// The same test through the set's thrust::device_vector overloads.
static void contains_insert_remove_device_vector(int engine) {
    constexpr int size = 1024;
    lslab_tuning tu{};
    tu.engine = engine;
    lslab::set<int> s(10, 10000, &tu);
    std::vector<int> hk(size);
    for (int i = 0; i < size; ++i) hk[i] = i + 1;
    thrust::device_vector<int> keys(hk.begin(), hk.end());
    thrust::device_vector<bool> res1(size), res2(size);
    for (int rep = 0; rep < 100; rep++) {
        s.insert(keys, res1);
        s.contains(keys, res2);
        HIPOK(hipDeviceSynchronize());
        thrust::host_vector<bool> h1 = res1, h2 = res2;
        for (int i = 0; i < size; ++i) ASSERT(h1[i] && h2[i], "insert / contains");
        s.remove(keys, res1);
        s.contains(keys, res2);
        HIPOK(hipDeviceSynchronize());
        h1 = res1;
        h2 = res2;
        for (int i = 0; i < size; ++i) ASSERT(h1[i] && !h2[i], "remove / contains after remove");
    }
    bool threw = false;
    try {
        thrust::device_vector<bool> small(size - 1);
        s.contains(keys, small);
    } catch (const lslab::error&) {
        threw = true;
    }
    ASSERT(threw, "output shorter than keys is refused");
    ASSERT(s.status() == LSLAB_OK, "device status");
}

// The legacy test's 128-byte Key: i in the first 8 bytes, zeros after
// (test/slabunified_test.cu.old:344-364).
struct Key128 {
    uint64_t w[16];
};
static Key128 key128(uint64_t i) {
    Key128 k{};
    k.w[0] = i;
    return k;
}

// 100 rounds: insert keys 1..1000 with fresh values (each reads back "absent"),
// then remove them (each returns the value that round inserted).
template <typename V, typename MakeV>
static void put_remove_128b(MakeV make_value) {
    constexpr int size = 1000;
    lslab::wide_map<Key128, V> m(10, 10000);
    std::vector<Key128> hk(size);
    for (int j = 0; j < size; ++j) hk[j] = key128(j + 1);
    dvec<Key128> keys(size);
    dvec<V> vals(size), old(size);
    dvec<bool> found(size);
    keys.from(hk);
    for (int rep = 0; rep < 100; ++rep) {
        std::vector<V> hv(size);
        for (int j = 0; j < size; ++j) hv[j] = make_value(rep, j);
        vals.from(hv);
        m.put(keys.p, vals.p, found.p, old.p, size);
        HIPOK(hipDeviceSynchronize());
        for (uint8_t b : found.bytes()) ASSERT(b == 0, "insert of an absent key found it");
        for (uint8_t b : old.bytes()) ASSERT(b == 0, "insert of an absent key read a value");
        m.remove(keys.p, found.p, old.p, size);
        HIPOK(hipDeviceSynchronize());
        for (uint8_t b : found.bytes()) ASSERT(b == 1, "remove missed a key");
        ASSERT(old.bytes() == vals.bytes(), "remove returned another value");
    }
    ASSERT(m.status() == LSLAB_OK, "device status");
    ASSERT(m.stats().present == 0, "keys left after the last round");
}

int main() {
    try {
        for (int engine = 0; engine < 2; ++engine) {
            get_put(10000, engine);
            std::printf("GetPutTest engine %d ok\n", engine);
            get_put_group(10000, engine);
            std::printf("GetPutTestGroupAllocator engine %d ok\n", engine);
            contains_insert_remove(engine);
            std::printf("ContainsInsertRemoveTest engine %d ok\n", engine);
            contains_insert_remove_device_vector(engine);
            std::printf("ContainsInsertRemoveTest device_vector engine %d ok\n", engine);
            get_put<lslab::mix_hash<int>>(10000, engine);
            std::printf("GetPutTest mix_hash engine %d ok\n", engine);
        }
        put_remove_128b<uint64_t>([](int rep, int j) { return ((uint64_t)rep << 32) | (j + 1); });
        std::printf("PutRemoveTest_128B ok\n");
        put_remove_128b<Key128>([](int rep, int) { return key128(rep); });
        std::printf("PutRemoveTest_128Bto128B ok\n");
    } catch (const std::exception& e) {
        std::printf("FAILED: %s\n", e.what());
        return 1;
    }
    return 0;
}
