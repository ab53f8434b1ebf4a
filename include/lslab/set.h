/**
 * @file
 * lslab::set<K> -- header-only host API over the lslab C ABI.
 *
 * Mirrors the reference's set (include/lslab/set.h:21-123 of the reference):
 * contains / insert / remove batches over device pointers, bool results.
 * insert reports true for present and newly inserted keys alike.  The
 * reference's thrust::device_vector overloads (set.h:98-123 of the reference)
 * are kept without a Thrust dependency: contains / insert / remove also take
 * any pair of device vectors whose data() is a device-pointer wrapper with
 * get(), thrust::device_vector<K> and thrust::device_vector<bool> among them.
 * Host containers, whose data() is a raw pointer, do not match.
 */
#pragma once

#include "lslab.h"

namespace lslab {

template <typename K, typename Hash = hash<K>>
class set {
    static_assert(detail::slab_word<K>(), "K must be a 4- or 8-byte integer");

public:
    set() : set(10) {}

    explicit set(unsigned n_log_2, uint64_t pool_slabs = 10000,
                 const lslab_tuning* tuning = nullptr)
        : n_log_2_(n_log_2) {
        const lslab_tuning tu = detail::with_hash<Hash>(tuning);
        check(lslab_set_create(&t_, sizeof(K), n_log_2, pool_slabs, &tu), "set_create");
    }

    /// Table whose slab array lives in caller memory (the reference's
    /// set(n_log_2, Allocator&&), set.h:42).
    set(unsigned n_log_2, uint64_t pool_slabs, slab_memory mem,
        const lslab_tuning* tuning = nullptr)
        : n_log_2_(n_log_2) {
        const lslab_tuning tu = detail::with_hash<Hash>(tuning);
        check(lslab_set_create_at(&t_, sizeof(K), n_log_2, pool_slabs, mem.ptr, mem.bytes, &tu),
              "set_create_at");
    }

    /// Bytes of slab_memory a set(n_log_2, pool_slabs, mem) needs.
    static uint64_t slab_bytes(unsigned n_log_2, uint64_t pool_slabs) {
        return lslab_slab_bytes(0, sizeof(K), 0, n_log_2, pool_slabs);
    }

    set(const set&) = delete;
    set& operator=(const set&) = delete;
    set(set&& o) noexcept : t_(o.t_), n_log_2_(o.n_log_2_) { o.t_ = nullptr; }
    set& operator=(set&& o) noexcept {
        if (this != &o) {
            if (t_) lslab_destroy(t_);
            t_ = o.t_;
            n_log_2_ = o.n_log_2_;
            o.t_ = nullptr;
        }
        return *this;
    }
    ~set() {
        if (t_) lslab_destroy(t_);
    }

    template <int block_size = 0>
    void contains(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_set_contains(t_, keys, reinterpret_cast<uint8_t*>(output), size, stream),
              "set::contains");
    }

    template <int block_size = 0>
    void insert(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_set_insert(t_, keys, reinterpret_cast<uint8_t*>(output), size, stream),
              "set::insert");
    }

    template <int block_size = 0>
    void remove(const K* keys, bool* output, size_t size, hipStream_t stream = 0) {
        detail::apply_block_size<block_size>(t_);
        check(lslab_set_remove(t_, keys, reinterpret_cast<uint8_t*>(output), size, stream),
              "set::remove");
    }

    // Device-vector forms: keys.size() ops, results into output (which must
    // hold at least as many).
    template <int block_size = 0, typename KV, typename BV>
    auto contains(const KV& keys, BV& output, hipStream_t stream = 0)
        -> decltype((void)keys.data().get(), (void)output.data().get()) {
        check_room(keys.size(), output.size());
        contains<block_size>(keys.data().get(), output.data().get(), keys.size(), stream);
    }
    template <int block_size = 0, typename KV, typename BV>
    auto insert(const KV& keys, BV& output, hipStream_t stream = 0)
        -> decltype((void)keys.data().get(), (void)output.data().get()) {
        check_room(keys.size(), output.size());
        insert<block_size>(keys.data().get(), output.data().get(), keys.size(), stream);
    }
    template <int block_size = 0, typename KV, typename BV>
    auto remove(const KV& keys, BV& output, hipStream_t stream = 0)
        -> decltype((void)keys.data().get(), (void)output.data().get()) {
        check_room(keys.size(), output.size());
        remove<block_size>(keys.data().get(), output.data().get(), keys.size(), stream);
    }

    void mixed(const uint8_t* ops, const K* keys, bool* output, size_t size,
               hipStream_t stream = 0) {
        check(lslab_set_mixed(t_, ops, keys, reinterpret_cast<uint8_t*>(output), size, stream),
              "set::mixed");
    }

    void exec_batches(size_t nbatches, const uint64_t* offsets, const uint8_t* ops, const K* keys,
                      bool* output, hipStream_t stream = 0) {
        check(lslab_exec_batches(t_, nbatches, offsets, ops, keys, nullptr, output, stream),
              "set::exec_batches");
    }

    unsigned buckets() const { return 1u << n_log_2_; }
    int status(hipStream_t stream = 0) { return lslab_status(t_, stream); }
    lslab_stats stats(hipStream_t stream = 0) {
        lslab_stats s;
        check(lslab_stats_get(t_, &s, stream), "set::stats");
        return s;
    }
    void chain_histogram(uint64_t* hist, unsigned nbins = LSLAB_CHAIN_BINS,
                         hipStream_t stream = 0) {
        check(lslab_chain_histogram(t_, hist, nbins, stream), "set::chain_histogram");
    }
    lslab_table* handle() { return t_; }

private:
    static void check_room(size_t n, size_t room) {
        if (room < n) check(LSLAB_ERR_INVALID, "set: output shorter than keys");
    }

    lslab_table* t_ = nullptr;
    unsigned n_log_2_;
};

}  // namespace lslab
