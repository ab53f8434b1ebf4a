/**
 * @file
 * lslab::map<K,V> -- header-only host API over the lslab C ABI.
 *
 * Mirrors the reference's host batch API (include/lslab/map.h:24-234 of the
 * reference): same class name, constructor arguments (n_log_2, then the
 * allocator's pool size, default 10000 slabs as device_allocator.h:16),
 * method names, argument order and result layouts.  Pointers are caller-owned
 * device memory; every call is asynchronous on `stream`.  The device-side
 * per-op API (map::get/put called inside user kernels) is not offered.
 */
#pragma once

#include "lslab.h"

namespace lslab {

template <typename K, typename V, typename Hash = hash<K>>
class map {
    static_assert(detail::slab_word<K>(), "K must be a 4- or 8-byte integer");
    static_assert(detail::slab_word<V>(), "V must be a 4- or 8-byte integer");

public:
    using this_t = map<K, V, Hash>;

    map() : map(10) {}

    explicit map(unsigned n_log_2, uint64_t pool_slabs = 10000,
                 const lslab_tuning* tuning = nullptr)
        : n_log_2_(n_log_2) {
        const lslab_tuning tu = detail::with_hash<Hash>(tuning);
        check(lslab_map_create(&t_, sizeof(K), sizeof(V), n_log_2, pool_slabs, &tu),
              "map_create");
    }

    /// Table whose slab array lives in caller memory (the reference's
    /// map(n_log_2, Allocator&&), map.h:57).
    map(unsigned n_log_2, uint64_t pool_slabs, slab_memory mem,
        const lslab_tuning* tuning = nullptr)
        : n_log_2_(n_log_2) {
        const lslab_tuning tu = detail::with_hash<Hash>(tuning);
        check(lslab_map_create_at(&t_, sizeof(K), sizeof(V), n_log_2, pool_slabs, mem.ptr,
                                  mem.bytes, &tu),
              "map_create_at");
    }

    /// Bytes of slab_memory a map(n_log_2, pool_slabs, mem) needs.
    static uint64_t slab_bytes(unsigned n_log_2, uint64_t pool_slabs) {
        return lslab_slab_bytes(1, sizeof(K), sizeof(V), n_log_2, pool_slabs);
    }

    map(const map&) = delete;
    map& operator=(const map&) = delete;
    map(map&& o) noexcept : t_(o.t_), n_log_2_(o.n_log_2_) { o.t_ = nullptr; }
    map& operator=(map&& o) noexcept {
        if (this != &o) {
            if (t_) lslab_destroy(t_);
            t_ = o.t_;
            n_log_2_ = o.n_log_2_;
            o.t_ = nullptr;
        }
        return *this;
    }
    ~map() {
        if (t_) lslab_destroy(t_);
    }

    /// Batch of puts from pair<K,V> (reference map.h:189-192); output[i] is
    /// the previous value, 0 if the key was absent.
    template <int block_size = 0>
    void put(key_value<K, V>* operations, V* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_map_put_pairs(t_, operations, output, size, stream), "map::put");
    }

    /// Batch of puts from separate arrays (reference map.h:198-201).
    template <int block_size = 0>
    void put(K* operations_keys, V* operations_values, V* output, size_t size,
             hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_map_put(t_, operations_keys, operations_values, output, size, stream),
              "map::put");
    }

    /// Batch of gets (reference map.h:208-211): output[i] = {found, value}.
    template <int block_size = 0>
    void get(K* operations, found_value<V>* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_map_get(t_, operations, output, size, stream), "map::get");
    }

    /// Batch of removes: output[i] = {was present, removed value}.  The
    /// reference traversal supports REMOVE (detail/traverse.h:88-91) but its
    /// map does not expose it.
    template <int block_size = 0>
    void remove(K* operations, found_value<V>* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_map_remove(t_, operations, output, size, stream), "map::remove");
    }

    /// Mixed batch: ops[i] in LSLAB_OP_GET / PUT / REMOVE.
    void mixed(const uint8_t* ops, const K* keys, const V* values, found_value<V>* output,
               size_t size, hipStream_t stream = 0) {
        check(lslab_map_mixed(t_, ops, keys, values, output, size, stream), "map::mixed");
    }

    /// Many mixed batches in one persistent launch (capturable as one graph node).
    void exec_batches(size_t nbatches, const uint64_t* offsets, const uint8_t* ops, const K* keys,
                      const V* values, found_value<V>* output, hipStream_t stream = 0) {
        check(lslab_exec_batches(t_, nbatches, offsets, ops, keys, values, output, stream),
              "map::exec_batches");
    }

    unsigned buckets() const { return 1u << n_log_2_; }

    /// Sticky device status since the last call (pool exhaustion, timeouts).
    int status(hipStream_t stream = 0) { return lslab_status(t_, stream); }

    lslab_stats stats(hipStream_t stream = 0) {
        lslab_stats s;
        check(lslab_stats_get(t_, &s, stream), "map::stats");
        return s;
    }

    /// hist[i] = buckets whose chain has i+1 slabs (the last bin also counts
    /// longer chains); nbins <= LSLAB_CHAIN_BINS.
    void chain_histogram(uint64_t* hist, unsigned nbins = LSLAB_CHAIN_BINS,
                         hipStream_t stream = 0) {
        check(lslab_chain_histogram(t_, hist, nbins, stream), "map::chain_histogram");
    }

    void set_tuning(const lslab_tuning& tu) { check(lslab_set_tuning(t_, &tu), "set_tuning"); }

    lslab_table* handle() { return t_; }

private:
    lslab_table* t_ = nullptr;
    unsigned n_log_2_;
};

}  // namespace lslab
