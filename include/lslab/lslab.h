/**
 * @file
 * Common pieces of the header-only C++ API over the lslab C ABI.
 *
 * The reference reports nothing: host errors are ignored or exit() the
 * process (include/lslab/lslab.h:28-41 of the reference).  Here every C ABI
 * status other than LSLAB_OK becomes an lslab::error exception.
 */
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <type_traits>

#include "../lslab_c.h"

namespace lslab {

struct error : std::runtime_error {
    int code;
    explicit error(int c, const char* what = "")
        : std::runtime_error(std::string("lslab: ") + lslab_error_string(c) +
                             (what && *what ? std::string(" in ") + what : std::string())),
          code(c) {}
};

inline void check(int rc, const char* what = "") {
    if (rc != LSLAB_OK) throw error(rc, what);
}

/// Identity hash: the reference's hash<T> (include/lslab/hash.h:10-15).  A
/// table created with it buckets by (size_t)key & (2^n - 1) on the device.
template <typename T>
struct hash {
    static constexpr int kind = LSLAB_HASH_IDENTITY;
    size_t operator()(T x) const { return static_cast<size_t>(x); }
};

/// Mixing hash (new): buckets by the murmur3 64-bit finalizer of the key, for
/// keys whose low bits are not uniform.  The device computes it; this host
/// functor gives the same value.
template <typename T>
struct mix_hash {
    static constexpr int kind = LSLAB_HASH_MIX;
    size_t operator()(T x) const {
        uint64_t z = static_cast<uint64_t>(x);
        z ^= z >> 33;
        z *= 0xff51afd7ed558ccdull;
        z ^= z >> 33;
        z *= 0xc4ceb9fe1a85ec53ull;
        z ^= z >> 33;
        return static_cast<size_t>(z);
    }
};

namespace detail {
/// The caller's tuning (or the defaults) with the hash of the Hash functor.
template <typename Hash>
inline lslab_tuning with_hash(const lslab_tuning* tuning) {
    lslab_tuning tu{};
    if (tuning) tu = *tuning;
    tu.hash = Hash::kind;
    return tu;
}
}  // namespace detail

/// Caller-owned device memory for a table's slab array (bucket heads, then
/// the pool).  It stands where the reference's Allocator argument supplies
/// the pool (reference map.h:57, set.h:42): allocate it as you like (at least
/// map::slab_bytes / set::slab_bytes bytes, 256-B aligned) and keep it alive
/// until the table is destroyed; the table never frees it.
struct slab_memory {
    void* ptr;
    size_t bytes;
};

/// Layout of cuda::std::pair<bool, V>: the reference's get() result
/// (include/lslab/map.h:208-211).
template <typename V>
struct found_value {
    bool first;
    V second;
};

/// Layout of cuda::std::pair<K, V>: the reference's AoS put() input
/// (include/lslab/map.h:189-192).
template <typename K, typename V>
struct key_value {
    K first;
    V second;
};

namespace detail {
template <typename T>
constexpr bool slab_word() {
    return std::is_integral<T>::value && (sizeof(T) == 4 || sizeof(T) == 8);
}

/// The reference chooses the CTA size per call as a template argument
/// (map.h:189-220, default 256).  Here the default, 0, keeps the table's own
/// block size (set at construction or by set_tuning); an explicit size (wave64
/// kernels take 64..256 threads per block, so it is clamped to that) becomes
/// the table's block size, as set_tuning would make it -- a host-side change
/// of shared table state, so callers that share a table across host threads
/// pass sizes through set_tuning instead.
template <int block_size>
inline void apply_block_size(lslab_table* t) {
    static_assert(block_size == 0 || (block_size >= 32 && block_size <= 1024), "block_size");
    if constexpr (block_size != 0) {
        constexpr int bt = block_size < 64 ? 64 : (block_size > 256 ? 256 : block_size);
        lslab_tuning tu;
        check(lslab_get_tuning(t, &tu), "get_tuning");
        if (tu.block_threads != bt) {
            tu.block_threads = bt;
            check(lslab_set_tuning(t, &tu), "set_tuning");
        }
    }
}
}  // namespace detail

}  // namespace lslab
