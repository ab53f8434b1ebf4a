/*
 * lslab C ABI -- MI355X (gfx950) slab-list hash map and set.
 *
 * The drop-in boundary of the LSlab batch API.  The reference exposes a
 * header-only CUDA C++ template API (include/lslab/map.h:24-25,
 * include/lslab/set.h:21-22) whose host batch methods launch one kernel per
 * batch on a caller stream.  This ABI carries the same calls as plain C
 * functions over an opaque table handle, with key/value widths chosen at
 * creation (4 or 8 bytes each), caller-owned DEVICE pointers for every input
 * and output, a hipStream_t per call, and an int status instead of the
 * reference's silent failures and device traps.  The header-only C++ classes
 * in include/lslab/map.h and include/lslab/set.h sit on top of it.
 *
 * Threading: every call is asynchronous on its stream.  Calls on one table
 * are ordered by the library (each call waits for the table's previous call,
 * whatever stream it used), which is the linearisation the reference obtains
 * from its per-bucket locks (include/lslab/warp_mutex.h:13-59).
 *
 * Device-side failures (slab pool exhausted, a lock or barrier spin that hit
 * its time bound) are sticky per table and reported by lslab_status().
 */
#ifndef LSLAB_C_H
#define LSLAB_C_H

#include <stddef.h>
#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  The reference has none: allocation errors are ignored
 * (map.h:45-51), pool exhaustion traps (detail/device_allocator.h:65-68). */
#define LSLAB_OK 0
#define LSLAB_ERR_INVALID 1        /* bad argument or unsupported width */
#define LSLAB_ERR_HIP 2            /* a HIP runtime call failed */
#define LSLAB_ERR_OOM 3            /* device allocation failed */
#define LSLAB_ERR_POOL_EXHAUSTED 4 /* an insert needed a slab beyond the pool */
#define LSLAB_ERR_TIMEOUT 5        /* a device spin exceeded its time bound */
#define LSLAB_ERR_CORRUPT 6        /* a chain link or batch index failed its bounds check */

/* Opcodes of the mixed-op entry points.  Maps: get/put/remove; sets:
 * contains/insert/remove.  The reference's legacy tests tag requests the same
 * way (REQUEST_INSERT / REQUEST_REMOVE / REQUEST_EMPTY,
 * test/slabunified_test.cu.old:60-67) without the repository defining their
 * values, so the numbering is this ABI's own.  Any other opcode value is
 * treated as GET. */
#define LSLAB_OP_GET 0
#define LSLAB_OP_PUT 1
#define LSLAB_OP_REMOVE 2

/* Engines.  PLAIN: the reference protocol, one wave64 group per op holding
 * its bucket's reader/writer lock for the traversal.  TWO_PHASE: lock-free
 * reads of the pre-batch table plus one atomic exchange per mutation on a
 * batch-local per-key chain, then one write-back per mutated key. */
#define LSLAB_ENGINE_PLAIN 0
#define LSLAB_ENGINE_TWO_PHASE 1

/* Two-phase combining scope: mutations of one key are chained within a
 * wavefront (WAVE), or also across the wavefronts of a block through LDS
 * (BLOCK), before the one global exchange per chain.  BLOCK adds one block
 * barrier per round and pays off under heavy key skew. */
#define LSLAB_COMBINE_WAVE 1
#define LSLAB_COMBINE_BLOCK 2

/* Bucket hash, fixed when the table is created.  IDENTITY is the reference's
 * hash<T> (include/lslab/hash.h:10-15): bucket = key & (2^n - 1).  MIX buckets
 * by the murmur3 64-bit finalizer of the key (zero-extended for 32-bit keys),
 * for key sets whose low bits are not uniform. */
#define LSLAB_HASH_IDENTITY 1
#define LSLAB_HASH_MIX 2

/* Slot and slab reuse, fixed when the table is created.  OFF is the
 * reference's policy (detail/traverse.h:54-55,98-124;
 * detail/device_allocator.h:62-73): an insert uses only the last slab's
 * empty slots, and pool slabs are never freed.  ON (new): an insert takes the
 * lowest empty slot of the first slab of its chain that has one; a remove that
 * leaves a pool slab empty unlinks and frees it; new slabs come from the
 * freed ones before the bump pointer.  Long runs with removes then keep their
 * pool headroom.  stats.slabs_used counts slabs in chains. */
#define LSLAB_RECLAIM_OFF 1
#define LSLAB_RECLAIM_ON 2

/* Chain heads per key slot (new).  The two-phase engine orders each key's
 * mutations of a sub-batch by one atomic exchange on that key's chain head.
 * With ON the table keeps one 4-B head per key slot of its slab array
 * (128 B per slab: +25% on a u64 map's 512-B slabs), and a key already in
 * the table heads its chain at its own slot; otherwise (OFF, and keys not in
 * the table) the head is an entry of the batch combining table, claimed by
 * one compare-and-swap first.  One device round trip fewer per mutation:
 * +5.8% on the 95/5 get/put Zipf-0.99 stream.  0 = AUTO: ON while the
 * tuning's engine is TWO_PHASE.  create and set_tuning allocate or free the
 * heads (LSLAB_ERR_OOM when they do not fit); a table of more than 2^26
 * slabs has none. */
#define LSLAB_SLOT_HEADS_ON 1
#define LSLAB_SLOT_HEADS_OFF 2
/* Runs of batches in lslab_exec_batches (lslab_tuning.runs). */
#define LSLAB_RUNS_ON 1
#define LSLAB_RUNS_OFF 2

typedef struct lslab_table lslab_table;

typedef struct {
    int lanes_per_op;      /* lanes probing one slab per op: 1, 2, 4, 8, 16, 32 (0 = 4) */
    int engine;            /* LSLAB_ENGINE_* for put/remove/mixed calls (0 = PLAIN) */
    int block_threads;     /* threads per block (0 = 256) */
    int max_blocks;        /* grid cap for the per-batch kernels (0 = auto) */
    uint64_t max_subbatch; /* two-phase ops per sub-batch; sizes the combining
                              table and per-op scratch (0 = 2^20) */
    int combine;           /* LSLAB_COMBINE_* for the two-phase engine (0 = BLOCK) */
    int hash;              /* LSLAB_HASH_* at create (0 = IDENTITY); set_tuning keeps it
                              (0 there means "unchanged", another value is an error) */
    int reclaim;           /* LSLAB_RECLAIM_* at create (0 = OFF); set_tuning keeps it */
    int spin_ms;           /* bound on every device spin -- a bucket lock, a grid barrier --
                              in ms, 10..3600000 (0 = 10000): a spin that makes no progress
                              for that long ends its kernel with LSLAB_ERR_TIMEOUT.  The
                              served loop's wait for an arrival is not bounded by it. */
    int slot_heads;        /* LSLAB_SLOT_HEADS_* (0 = AUTO: ON while engine is TWO_PHASE) */
    int runs;              /* LSLAB_RUNS_* for lslab_exec_batches.  0 = AUTO: the call checks on
                              the device whether any two consecutive batches fit one run
                              (max_subbatch ops, at most 2^20) and launches the run kernel and
                              the all-alone kernel, of which the one that does not apply
                              returns at once.  ON: the run kernel only.  OFF: the all-alone
                              kernel only (every batch runs by itself; for streams whose
                              batches never fit a run, one launch less).  Every setting gives
                              the same results. */
} lslab_tuning;

typedef struct {
    uint64_t buckets;      /* 2^n_log2 */
    uint64_t pool_slabs;   /* pool capacity */
    uint64_t slabs_used;   /* pool slabs handed out (min(bump pointer, pool)) */
    uint64_t present;      /* keys whose valid bit is set */
    uint64_t max_chain;    /* longest bucket chain, in slabs */
    uint64_t slab_bytes;   /* device bytes per slab (a u32 set: 128-B key line + 4-B link) */
    uint64_t device_bytes; /* bytes of device memory the table holds */
} lslab_stats;

/* Creation / destruction.  Reference: map(n_log_2) / set(n_log_2)
 * (map.h:43-52, set.h:31-40) with device_allocator(stack_size = 10000)
 * (device_allocator.h:16, detail/device_allocator.h:9-14).  Unlike the
 * reference the destructor frees everything.  n_log2 <= 30 as in the
 * reference's int shift (map.h:44).  tuning may be NULL. */
int lslab_map_create(lslab_table** out, int key_bytes, int value_bytes, unsigned n_log2,
                     uint64_t pool_slabs, const lslab_tuning* tuning);
int lslab_set_create(lslab_table** out, int key_bytes, unsigned n_log2, uint64_t pool_slabs,
                     const lslab_tuning* tuning);
/* Pluggable slab memory.  The reference takes an Allocator type that owns the
 * pool slabs (map.h:24 and map.h:57, set.h:21 and set.h:42; its default is a
 * device bump pool, device_allocator.h:13-33, and its tests plug in one over
 * a group allocator, test/slabunified_test.cu:92-162).  Here the caller hands
 * over the memory of the whole slab array instead (the 2^n bucket heads, then
 * the pool), allocated however it likes: device memory, 256-B aligned, at
 * least lslab_slab_bytes() bytes, owned by the caller and kept alive until
 * lslab_destroy (which does not free it).  The library zeroes it.  Locks and
 * scratch stay library-allocated.  slab_bytes returns 0 for an invalid
 * shape; create_at returns LSLAB_ERR_INVALID for a misaligned or short
 * buffer. */
uint64_t lslab_slab_bytes(int is_map, int key_bytes, int value_bytes, unsigned n_log2,
                          uint64_t pool_slabs);
int lslab_map_create_at(lslab_table** out, int key_bytes, int value_bytes, unsigned n_log2,
                        uint64_t pool_slabs, void* slab_mem, uint64_t slab_mem_bytes,
                        const lslab_tuning* tuning);
int lslab_set_create_at(lslab_table** out, int key_bytes, unsigned n_log2, uint64_t pool_slabs,
                        void* slab_mem, uint64_t slab_mem_bytes, const lslab_tuning* tuning);
int lslab_destroy(lslab_table* t);
int lslab_set_tuning(lslab_table* t, const lslab_tuning* tuning);
int lslab_get_tuning(const lslab_table* t, lslab_tuning* out);

/* Map batch calls.
 * put_pairs: map::put(pair<K,V>* ops, V* out, n, stream)   map.h:189-192
 *            (pairs laid out as cuda::std::pair<K,V>: K, then V at alignof)
 * put:       map::put(K* keys, V* vals, V* out, n, stream)  map.h:198-201
 *            out[i] = previous value, 0 if the key was absent (map.h:136-152)
 * get:       map::get(K* keys, pair<bool,V>* out, n, stream) map.h:208-211
 *            out[i] = {found, value}; value 0 when absent
 * remove:    new; the reference traversal supports REMOVE on map nodes
 *            (detail/traverse.h:88-91) but map.h does not expose it.
 *            out[i] = {was present, removed value}
 * mixed:     new; per-op opcodes (LSLAB_OP_*), out[i] = {found, value} with the
 *            meaning of the op's single-type call (put: {was present, old}).
 *            values holds n entries, one per op (ignored for non-puts but
 *            readable), or is null when no op is a put. */
int lslab_map_put_pairs(lslab_table* t, const void* pairs, void* out_old, size_t n,
                        hipStream_t stream);
int lslab_map_put(lslab_table* t, const void* keys, const void* values, void* out_old, size_t n,
                  hipStream_t stream);
int lslab_map_get(lslab_table* t, const void* keys, void* out_pairs, size_t n,
                  hipStream_t stream);
int lslab_map_remove(lslab_table* t, const void* keys, void* out_pairs, size_t n,
                     hipStream_t stream);
int lslab_map_mixed(lslab_table* t, const uint8_t* ops, const void* keys, const void* values,
                    void* out_pairs, size_t n, hipStream_t stream);

/* Set batch calls: set::insert / contains / remove(const K*, bool*, n, stream)
 * (set.h:92-123).  insert reports true for present and newly inserted keys
 * alike (detail/traverse.h:204-212,226-232); remove reports whether the key
 * was present.  mixed: per-op opcodes (GET = contains, PUT = insert). */
int lslab_set_insert(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                     hipStream_t stream);
int lslab_set_contains(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                       hipStream_t stream);
int lslab_set_remove(lslab_table* t, const void* keys, uint8_t* out, size_t n,
                     hipStream_t stream);
int lslab_set_mixed(lslab_table* t, const uint8_t* ops, const void* keys, uint8_t* out, size_t n,
                    hipStream_t stream);

/* Batched form (new): nbatches mixed-op batches laid end to end in ops/keys/
 * values/out, batch b = [offsets[b], offsets[b+1]) with offsets a DEVICE
 * array of nbatches+1 uint64.  One persistent launch runs them in order with
 * the two-phase engine (batch b sees every effect of batches < b; per key a
 * batch's gets come before its mutations), so the
 * call is one graph node when captured.  out as lslab_map_mixed (maps) or
 * bool per op (sets, values ignored).  Offsets must not decrease: the device
 * checks them first and refuses the whole call otherwise (no op runs; the
 * next lslab_status returns LSLAB_ERR_INVALID). */
int lslab_exec_batches(lslab_table* t, size_t nbatches, const uint64_t* offsets,
                       const uint8_t* ops, const void* keys, const void* values, void* out,
                       hipStream_t stream);

/* Batched form that allocates nothing (new; no counterpart in the reference).
 * For callers that cannot allocate device memory at the call -- code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * Same arguments, kernels and results as lslab_exec_batches, but the
 * batched kernel's scratch (grid-barrier state, combining table and
 * final-candidate lists) lives in one static device array of the library
 * (about 25 MiB per GPU, held until the process ends) instead of the table's.
 * Limits: the table's max_subbatch must be at most
 * LSLAB_NOALLOC_MAX_SUBBATCH ops and its persistent grid (CUs, or max_blocks
 * when set) at most LSLAB_NOALLOC_MAX_BLOCKS blocks; otherwise the call
 * returns LSLAB_ERR_INVALID before any device work.  Every call on a GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other tables) race: the caller orders them.  The array keeps the last
 * call's contents; each call empties the combining table and resets the
 * barrier with kernels of its own before the batched kernel runs. */
#define LSLAB_NOALLOC_MAX_SUBBATCH 65536
#define LSLAB_NOALLOC_MAX_BLOCKS 256
int lslab_exec_batches_noalloc(lslab_table* t, size_t nbatches, const uint64_t* offsets,
                               const uint8_t* ops, const void* keys, const void* values,
                               void* out, hipStream_t stream);

/* Served form (new): a device-resident request loop.  Request r owns ops
 * [req_offsets[r], req_offsets[r+1]) of ops/keys/values/out and arrives
 * arrival_ticks[r] ticks of the device's 100 MHz constant clock after the
 * loop starts (nondecreasing).  One persistent launch admits arrived requests
 * in order, tick by tick (up to max_tick_ops ops and max_tick_reqs requests
 * per tick, and only while the slab pool has headroom for req_reserve[r]
 * slabs -- one per put bounds what a request can allocate; NULL reserves one
 * per op), runs each tick as one two-phase batch, and writes
 * complete_ticks[r] (same clock) and tick_of[r] (LSLAB_SERVE_REJECTED when the
 * pool can never hold the request).  All arrays are DEVICE memory.  Request
 * offsets must not decrease: the loop ends at the first request whose end
 * lies before its start (it and later requests do not run; the next
 * lslab_status returns LSLAB_ERR_INVALID). */
#define LSLAB_SERVE_REJECTED 0xffffffffu
typedef struct {
    uint64_t nreq;
    const uint64_t* req_offsets;   /* nreq + 1 */
    const uint64_t* arrival_ticks; /* nreq */
    const uint32_t* req_reserve;   /* nreq, or NULL */
    uint64_t* complete_ticks;      /* nreq, written */
    uint32_t* tick_of;             /* nreq, written */
    uint64_t max_tick_ops;         /* 0 = the table's max_subbatch; with a grid of
                                      B > 1 blocks, one of which only admits, a
                                      tick holds at most max_tick_ops * (B-1) / B
                                      ops (the workers' share) */
    uint32_t max_tick_reqs;        /* 0 = no limit */
} lslab_serve_params;

typedef struct {
    uint64_t ticks;         /* batches run */
    uint64_t rejected;      /* requests refused for lack of pool headroom */
    uint64_t served_ops;    /* ops admitted */
    uint64_t elapsed_ticks; /* loop duration on the 100 MHz clock */
} lslab_serve_stats;

int lslab_serve(lslab_table* t, const lslab_serve_params* p, const uint8_t* ops, const void* keys,
                const void* values, void* out, hipStream_t stream);
/* Synchronises and reads the last serve loop's counters. */
int lslab_serve_result(lslab_table* t, lslab_serve_stats* out, hipStream_t stream);

/* Synchronises the stream, returns the table's sticky device status
 * (LSLAB_OK, LSLAB_ERR_POOL_EXHAUSTED, LSLAB_ERR_TIMEOUT) and clears it. */
int lslab_status(lslab_table* t, hipStream_t stream);

/* Introspection (synchronous).  dump writes up to cap present pairs into the
 * DEVICE arrays keys_out / vals_out (vals_out ignored for sets; order
 * unspecified) and the number of present keys into *count.
 * chain_histogram (new; the reference has no introspection) writes into the
 * HOST array hist[0..nbins) the number of buckets whose chain has i+1 slabs,
 * with the last bin counting every longer chain too; 1 <= nbins <=
 * LSLAB_CHAIN_BINS. */
#define LSLAB_CHAIN_BINS 16
int lslab_stats_get(lslab_table* t, lslab_stats* out, hipStream_t stream);
int lslab_chain_histogram(lslab_table* t, uint64_t* hist, unsigned nbins, hipStream_t stream);
int lslab_dump(lslab_table* t, void* keys_out, void* vals_out, uint64_t cap, uint64_t* count,
               hipStream_t stream);

/* Wide-key tables (new): map / set with keys of 4..128 bytes and values of
 * 0 (set) .. 128 bytes, both multiples of 4 -- the reference's legacy tests
 * of 128-B keys (test/slabunified_test.cu.old:344-553: Key -> int*, Key ->
 * Key).  A slab is one 128-B line of 30 key tags (32-bit hash
 * fingerprints), the next link and the valid bitmap; keys and values live in
 * arenas parallel to the slots, so a probe reads the tag line and confirms a
 * tag match against the key arena.  Same op semantics, bucket locks and slab
 * policy as the plain engine of the narrow tables (insert into the last
 * slab's lowest empty slot; pool slabs never freed).  hash: 0 or
 * LSLAB_HASH_IDENTITY (bucket = the key's first 8 bytes) or LSLAB_HASH_MIX.
 * Inputs and outputs are DEVICE arrays: keys n * key_bytes, values n *
 * value_bytes, out_found n bytes, out values n * value_bytes (either output
 * may be NULL).  get reads without locks; put / remove lock their buckets;
 * mixed runs its gets without locks against the table as it was before the
 * call, then its mutations under the locks (per key the serial order [gets]
 * [mutations], as the two-phase engine of the narrow tables).  A set reports
 * insert as true, like lslab_set_insert.
 * stats: slab_bytes counts the tag line plus the 30 slots' arena bytes. */
typedef struct lslab_wide lslab_wide;
int lslab_wide_create(lslab_wide** out, int key_bytes, int value_bytes, unsigned n_log2,
                      uint64_t pool_slabs, int hash);
int lslab_wide_destroy(lslab_wide* w);
int lslab_wide_put(lslab_wide* w, const void* keys, const void* values, uint8_t* out_found,
                   void* out_old, size_t n, hipStream_t stream);
int lslab_wide_get(lslab_wide* w, const void* keys, uint8_t* out_found, void* out_values,
                   size_t n, hipStream_t stream);
int lslab_wide_remove(lslab_wide* w, const void* keys, uint8_t* out_found, void* out_old,
                      size_t n, hipStream_t stream);
int lslab_wide_mixed(lslab_wide* w, const uint8_t* ops, const void* keys, const void* values,
                     uint8_t* out_found, void* out_values, size_t n, hipStream_t stream);
int lslab_wide_status(lslab_wide* w, hipStream_t stream);
int lslab_wide_stats(lslab_wide* w, lslab_stats* out, hipStream_t stream);
int lslab_wide_dump(lslab_wide* w, void* keys_out, void* vals_out, uint64_t cap, uint64_t* count,
                    hipStream_t stream);
/* Slabs (heads + pool) of a wide table; 0 if the shape is invalid. */
uint64_t lslab_wide_slab_count(unsigned n_log2, uint64_t pool_slabs);

/* Shape of a table (new): 1 for a map, key and value widths in bytes (0 for
 * a set's value), and the HIP device it lives on. */
int lslab_table_info(const lslab_table* t, int* is_map, int* key_bytes, int* value_bytes,
                     int* device);

/* KVCG hot/cold split (new).  KVCG, the system the reference's table comes
 * from (README.md:3, CITATION.bib:1-17), keeps the hottest keys of a skewed
 * workload in a CPU-side cache in front of the GPU slab table, for requests
 * that arrive in host memory; the reference repository holds only the GPU
 * table.  Here a lslab_kvcg owns a host tier of up to hot_capacity keys in
 * front of table t:
 *   - the hot set is chosen by the caller (set_hot) or as the k most frequent
 *     keys of a sample (sample_hot; ties go to the smaller key).  While a key
 *     is hot the host tier is its only authority; demoting it writes its
 *     state back to the table (put if present, remove if absent), and flush
 *     writes back every hot key (the table is then complete, e.g. for
 *     lslab_dump).
 *   - mixed runs a batch whose ops, keys, values and out are HOST memory
 *     (out laid out as for lslab_map_mixed / lslab_set_mixed): ops on hot
 *     keys run on `threads` host workers in batch order while the other ops
 *     are staged through pinned memory and run on the table in one mixed call
 *     on `stream`.  Per key the results follow a serial order (batch order
 *     for hot keys, the table engine's order for cold ones); a batch of
 *     distinct keys gives exactly the sequential replay's results.  Batches
 *     larger than max_batch run as consecutive slices.  The call returns
 *     when the batch is complete.
 *   - destroy flushes on the null stream, then frees; the table stays.
 * threads 0 = min(16, hardware threads); max_batch 0 = 2^20.  The table must
 * not be used directly while a kvcg holds hot keys of it, except after a
 * flush and before the next kvcg call. */
typedef struct lslab_kvcg lslab_kvcg;
typedef struct {
    uint64_t hot_keys;     /* keys in the hot set */
    uint64_t hot_capacity;
    uint64_t hot_ops;      /* ops run by the host tier */
    uint64_t cold_ops;     /* ops sent to the table */
    uint64_t batches;      /* mixed calls */
} lslab_kvcg_stats;
int lslab_kvcg_create(lslab_kvcg** out, lslab_table* t, uint64_t hot_capacity,
                      uint64_t max_batch, int threads);
int lslab_kvcg_destroy(lslab_kvcg* kv);
int lslab_kvcg_set_hot(lslab_kvcg* kv, const void* keys, uint64_t n, hipStream_t stream);
int lslab_kvcg_sample_hot(lslab_kvcg* kv, const void* sample_keys, uint64_t n, uint64_t k,
                          hipStream_t stream);
int lslab_kvcg_mixed(lslab_kvcg* kv, const uint8_t* ops, const void* keys, const void* values,
                     void* out, uint64_t n, hipStream_t stream);
int lslab_kvcg_flush(lslab_kvcg* kv, hipStream_t stream);
int lslab_kvcg_stats_get(const lslab_kvcg* kv, lslab_kvcg_stats* out);

const char* lslab_error_string(int status);

#ifdef __cplusplus
}
#endif
#endif /* LSLAB_C_H */
