/*
 * lslab oracle: a sequential host replay of the LSlab slab-list hash map and
 * set, restated in plain C from the reference's algorithm.
 *
 * TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, and only as the
 * checker or as the timed host baseline.  The product path (lslab_amd/) never
 * links, loads or calls it.
 *
 * Every op of a batch is applied in op-index order, one at a time, exactly as
 * one warp of the reference applies its lanes in lane order
 * (include/lslab/detail/traverse.h:39-48).  Keys and values are passed as
 * uint64_t arrays whatever the table's key/value width; 4-byte tables use the
 * low 32 bits.
 *
 * Opcodes (same numbering as the product's mixed-op entry point):
 *   0 = get (map) / contains (set)
 *   1 = put (map) / insert (set)
 *   2 = remove
 *
 * Outputs per op: found[i] (0/1) and val[i]:
 *   map get     -> (present, value or 0)             map.h:105-119, S3
 *   map put     -> (was present, old value or 0)     map.h:136-152, S4
 *   map remove  -> (was present, removed value or 0) traverse.h:88-91, S5
 *   set contains-> (present, 0)                       set.h:64-70
 *   set insert  -> (1, 0) (always true; 0 only when the pool is exhausted)
 *                                                     traverse.h:204-212,226-232
 *   set remove  -> (was present, 0)                   traverse.h:246-250
 */
#ifndef LSLAB_ORACLE_H
#define LSLAB_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_table orc_table;

/* slab_keys: keys per slab (31 in the reference's slab_node/set_node). */
orc_table* orc_create(int is_map, int key_bytes, int val_bytes, unsigned n_log2,
                      uint64_t pool_slabs, int slab_keys);
void orc_destroy(orc_table* t);

/* Apply n ops in index order.  Returns the number of ops that could not be
 * applied because the slab pool was exhausted (those report found=0, val=0
 * and leave the table unchanged). */
int64_t orc_apply(orc_table* t, uint64_t n, const uint8_t* ops, const uint64_t* keys,
                  const uint64_t* vals, uint8_t* found_out, uint64_t* val_out);

int orc_lookup(const orc_table* t, uint64_t key, uint64_t* val_out);
uint64_t orc_present(const orc_table* t);
uint64_t orc_slabs_used(const orc_table* t);
uint64_t orc_max_chain(const orc_table* t);
/* Buckets by fmix64(key) (mix != 0) instead of the reference's identity
 * hash; call before the first apply. */
void orc_set_hash(orc_table* t, int mix);
/* Reclaim policy (on != 0; call before the first apply): inserts take the
 * lowest empty slot of the first slab of the chain that has one; a remove
 * that empties a pool slab unlinks it and frees it; new slabs come from the
 * free list (most recently freed first) before the bump pointer.  Slabs used
 * = bump count - free slabs.  Returns 0 if the free list cannot be
 * allocated. */
int orc_set_reclaim(orc_table* t, int on);
/* Buckets by chain length in slabs: hist[i] = chains of i+1 slabs, the last
 * bin also counting longer ones. */
void orc_chain_hist(const orc_table* t, uint64_t* hist, unsigned nbins);
/* Writes up to cap (key, value) pairs in bucket/chain/slot order; returns the
 * number of present keys. */
uint64_t orc_dump(const orc_table* t, uint64_t* keys, uint64_t* vals, uint64_t cap);

/* Per-key serial-order checker for a batch whose ops may conflict on a key.
 * The table holds the state before the batch.  For every key of the batch it
 * checks that the observed results (found, val) match SOME serial order of
 * that key's ops (gets may sit anywhere; the mutations must form one chain of
 * states starting at the pre-batch state).  It then moves the table to the
 * end state that chain implies, so the next batch can be checked.
 * post_found/post_val (may be NULL) give, per op, the state of that op's key
 * observed after the batch; when given, the chain must end there.  For a set
 * the end state is not implied by the results alone, so pass them.
 * Returns the number of keys whose results match no serial order; the first
 * offending op index is written to *first_bad (or UINT64_MAX). */
int64_t orc_check_batch(orc_table* t, uint64_t n, const uint8_t* ops, const uint64_t* keys,
                        const uint64_t* vals, const uint8_t* found, const uint64_t* got,
                        const uint8_t* post_found, const uint64_t* post_val,
                        uint64_t* first_bad);

#ifdef __cplusplus
}
#endif
#endif
