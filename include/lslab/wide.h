/**
 * @file
 * lslab::wide_map<K,V> and lslab::wide_set<K> -- header-only host API over
 * the wide-key tables of the lslab C ABI (lslab_wide_*).
 *
 * For key types of 4..128 bytes that are not a single integer word, such as
 * the reference's legacy 128-byte Key (test/slabunified_test.cu.old:344-364),
 * stored as SlabUnified<Key, int*> and SlabUnified<Key, Key> there.  K and V
 * are trivially copyable records whose sizes are multiples of 4 bytes; keys
 * compare by their bytes.  Results come back as a found flag per op and a
 * value per op (the previous value for put and remove, 0 bytes when absent).
 * The bucket is the key's first 8 bytes (the reference's identity hash on the
 * key's leading word) or, with hash_kind = LSLAB_HASH_MIX, a hash of all of it.
 */
#pragma once

#include "lslab.h"

namespace lslab {

namespace detail {
template <typename T>
constexpr bool wide_record() {
    return std::is_trivially_copyable<T>::value && sizeof(T) % 4 == 0 && sizeof(T) >= 4 &&
           sizeof(T) <= 128;
}
}  // namespace detail

template <typename K, typename V>
class wide_map {
    static_assert(detail::wide_record<K>(), "K: trivially copyable, 4..128 bytes, 4-byte multiple");
    static_assert(detail::wide_record<V>(), "V: trivially copyable, 4..128 bytes, 4-byte multiple");

public:
    explicit wide_map(unsigned n_log_2 = 10, uint64_t pool_slabs = 10000,
                      int hash_kind = LSLAB_HASH_IDENTITY)
        : n_log_2_(n_log_2) {
        check(lslab_wide_create(&w_, sizeof(K), sizeof(V), n_log_2, pool_slabs, hash_kind),
              "wide_create");
    }
    wide_map(const wide_map&) = delete;
    wide_map& operator=(const wide_map&) = delete;
    wide_map(wide_map&& o) noexcept : w_(o.w_), n_log_2_(o.n_log_2_) { o.w_ = nullptr; }
    ~wide_map() {
        if (w_) lslab_wide_destroy(w_);
    }

    /// found[i] = key i was present; old[i] = its previous value (zeros if absent).
    void put(const K* keys, const V* values, bool* found, V* old, size_t size,
             hipStream_t stream = 0) {
        check(lslab_wide_put(w_, keys, values, reinterpret_cast<uint8_t*>(found), old, size,
                             stream),
              "wide_map::put");
    }
    void get(const K* keys, bool* found, V* values, size_t size, hipStream_t stream = 0) {
        check(lslab_wide_get(w_, keys, reinterpret_cast<uint8_t*>(found), values, size, stream),
              "wide_map::get");
    }
    void remove(const K* keys, bool* found, V* old, size_t size, hipStream_t stream = 0) {
        check(lslab_wide_remove(w_, keys, reinterpret_cast<uint8_t*>(found), old, size, stream),
              "wide_map::remove");
    }
    void mixed(const uint8_t* ops, const K* keys, const V* values, bool* found, V* out,
               size_t size, hipStream_t stream = 0) {
        check(lslab_wide_mixed(w_, ops, keys, values, reinterpret_cast<uint8_t*>(found), out, size,
                               stream),
              "wide_map::mixed");
    }

    unsigned buckets() const { return 1u << n_log_2_; }
    int status(hipStream_t stream = 0) { return lslab_wide_status(w_, stream); }
    lslab_stats stats(hipStream_t stream = 0) {
        lslab_stats s;
        check(lslab_wide_stats(w_, &s, stream), "wide_map::stats");
        return s;
    }
    /// Present keys and values into device arrays of cap entries; returns the count.
    uint64_t dump(K* keys, V* values, uint64_t cap, hipStream_t stream = 0) {
        uint64_t n = 0;
        check(lslab_wide_dump(w_, keys, values, cap, &n, stream), "wide_map::dump");
        return n;
    }
    lslab_wide* handle() { return w_; }

private:
    lslab_wide* w_ = nullptr;
    unsigned n_log_2_;
};

template <typename K>
class wide_set {
    static_assert(detail::wide_record<K>(), "K: trivially copyable, 4..128 bytes, 4-byte multiple");

public:
    explicit wide_set(unsigned n_log_2 = 10, uint64_t pool_slabs = 10000,
                      int hash_kind = LSLAB_HASH_IDENTITY)
        : n_log_2_(n_log_2) {
        check(lslab_wide_create(&w_, sizeof(K), 0, n_log_2, pool_slabs, hash_kind),
              "wide_create");
    }
    wide_set(const wide_set&) = delete;
    wide_set& operator=(const wide_set&) = delete;
    wide_set(wide_set&& o) noexcept : w_(o.w_), n_log_2_(o.n_log_2_) { o.w_ = nullptr; }
    ~wide_set() {
        if (w_) lslab_wide_destroy(w_);
    }

    /// Insert reports true for present and newly inserted keys alike.
    void insert(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        check(lslab_wide_put(w_, keys, nullptr, reinterpret_cast<uint8_t*>(output), nullptr, size,
                             stream),
              "wide_set::insert");
    }
    void contains(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        check(lslab_wide_get(w_, keys, reinterpret_cast<uint8_t*>(output), nullptr, size, stream),
              "wide_set::contains");
    }
    void remove(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        check(lslab_wide_remove(w_, keys, reinterpret_cast<uint8_t*>(output), nullptr, size,
                                stream),
              "wide_set::remove");
    }

    unsigned buckets() const { return 1u << n_log_2_; }
    int status(hipStream_t stream = 0) { return lslab_wide_status(w_, stream); }
    lslab_stats stats(hipStream_t stream = 0) {
        lslab_stats s;
        check(lslab_wide_stats(w_, &s, stream), "wide_set::stats");
        return s;
    }
    lslab_wide* handle() { return w_; }

private:
    lslab_wide* w_ = nullptr;
    unsigned n_log_2_;
};

}  // namespace lslab
