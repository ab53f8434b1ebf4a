// Host-only build of the KVCG split for CPU tests: the HIP runtime calls it
// makes are replaced by host memory operations, and the table behind it by a
// sequential host map with the same op semantics (map: get/put/remove with
// {found, old value}; set: contains/insert/remove).  Linked with
// lslab_amd/csrc/kvcg.cpp into libkvcg_host.so; tests/test_kvcg_host.py
// drives it through the lslab_kvcg_* C ABI and checks it against the oracle.
#include <cstdlib>
#include <cstring>
#include <unordered_map>

#include "lslab_c.h"

struct lslab_table {
    int is_map, kb, vb;
    std::unordered_map<uint64_t, uint64_t> m;
    uint64_t calls = 0, ops = 0;
};

static uint64_t ld(const void* p, uint64_t i, int w) {
    uint64_t v = 0;
    std::memcpy(&v, static_cast<const char*>(p) + i * w, w);
    return v;
}

extern "C" {
hipError_t hipHostMalloc(void** p, size_t n, unsigned) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipHostFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMalloc(void** p, size_t n) {
    *p = std::malloc(n ? n : 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
hipError_t hipFree(void* p) {
    std::free(p);
    return hipSuccess;
}
hipError_t hipMemcpyAsync(void* d, const void* s, size_t n, hipMemcpyKind, hipStream_t) {
    std::memcpy(d, s, n);
    return hipSuccess;
}
hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
hipError_t hipGetDevice(int* d) {
    *d = 0;
    return hipSuccess;
}
hipError_t hipSetDevice(int) { return hipSuccess; }

lslab_table* stub_table_create(int is_map, int kb, int vb) {
    lslab_table* t = new lslab_table;
    t->is_map = is_map;
    t->kb = kb;
    t->vb = is_map ? vb : 0;
    return t;
}
void stub_table_destroy(lslab_table* t) { delete t; }
uint64_t stub_table_size(lslab_table* t) { return t->m.size(); }
uint64_t stub_table_ops(lslab_table* t) { return t->ops; }
// Present keys and values, in any order; returns the count written.
uint64_t stub_table_dump(lslab_table* t, uint64_t* keys, uint64_t* vals, uint64_t cap) {
    uint64_t n = 0;
    for (const auto& kv : t->m) {
        if (n == cap) break;
        keys[n] = kv.first;
        vals[n] = kv.second;
        ++n;
    }
    return n;
}

int lslab_table_info(const lslab_table* t, int* is_map, int* kb, int* vb, int* dev) {
    if (!t) return LSLAB_ERR_INVALID;
    if (is_map) *is_map = t->is_map;
    if (kb) *kb = t->kb;
    if (vb) *vb = t->vb;
    if (dev) *dev = 0;
    return LSLAB_OK;
}

static void apply(lslab_table* t, uint8_t op, uint64_t k, uint64_t v, bool& f, uint64_t& old) {
    auto it = t->m.find(k);
    f = it != t->m.end();
    old = f ? it->second : 0;
    if (op == LSLAB_OP_PUT) {
        t->m[k] = t->is_map ? v : 0;
        if (!t->is_map) f = true;
    } else if (op == LSLAB_OP_REMOVE) {
        if (f) t->m.erase(it);
    }
}

int lslab_map_mixed(lslab_table* t, const uint8_t* ops, const void* keys, const void* values,
                    void* out, size_t n, hipStream_t) {
    ++t->calls;
    t->ops += n;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t op = ops[i] <= LSLAB_OP_REMOVE ? ops[i] : LSLAB_OP_GET;
        bool f;
        uint64_t old;
        apply(t, op, ld(keys, i, t->kb), ld(values, i, t->vb), f, old);
        char* q = static_cast<char*>(out) + i * 2 * t->vb;
        q[0] = f ? 1 : 0;
        std::memcpy(q + t->vb, &old, t->vb);
    }
    return LSLAB_OK;
}

int lslab_set_mixed(lslab_table* t, const uint8_t* ops, const void* keys, uint8_t* out, size_t n,
                    hipStream_t) {
    ++t->calls;
    t->ops += n;
    for (size_t i = 0; i < n; ++i) {
        const uint8_t op = ops[i] <= LSLAB_OP_REMOVE ? ops[i] : LSLAB_OP_GET;
        bool f;
        uint64_t old;
        apply(t, op, ld(keys, i, t->kb), 0, f, old);
        out[i] = f ? 1 : 0;
    }
    return LSLAB_OK;
}
}
