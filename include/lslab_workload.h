/*
 * lslab synthetic workload generators (device side).
 *
 * Not part of the reference's interface: the reference's benchmark draws
 * keys with unseeded rand() on the host (benchmark/benchmark.cu:84-87).
 * These kernels produce the seeded op streams of the benchmark configurations
 * directly in device memory.  Every function is deterministic in its
 * arguments, asynchronous on `stream`, and returns LSLAB_OK or an LSLAB_ERR_*
 * code from lslab_c.h.
 */
#ifndef LSLAB_WORKLOAD_H
#define LSLAB_WORKLOAD_H

#include <stdint.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Distinct keys: out[i] = bijective mix of (first + i) -- murmur3's 32-bit or
 * 64-bit finaliser, seeded by xor -- so keys of distinct indices never
 * collide.  key_bytes 4 or 8. */
int lslab_gen_keys(void* out, int key_bytes, uint64_t first, uint64_t n, uint64_t seed,
                   hipStream_t stream);

/* Values: out[i] = mix64((first + i) ^ seed), truncated to val_bytes. */
int lslab_gen_values(void* out, int val_bytes, uint64_t first, uint64_t n, uint64_t seed,
                     hipStream_t stream);

typedef struct {
    uint64_t seed;
    uint64_t first_op;     /* global index of element 0 (random streams and values) */
    double theta;          /* Zipf exponent; 0 = uniform ranks */
    unsigned space_log2;   /* ranks 1..2^space_log2 map bijectively onto keys [0, 2^space_log2) */
    double p_get, p_put;   /* op mix; remove with probability 1 - p_get - p_put */
    int key_bytes, val_bytes;
} lslab_gen_params;

/* Mixed op stream: ops[i] in {LSLAB_OP_GET, PUT, REMOVE}; keys[i] =
 * perm(rank_i - 1) with rank_i ~ Zipf(theta, 2^space_log2) drawn by
 * rejection-inversion (Hoermann & Derflinger, 1996) and perm a fixed bijection
 * of [0, 2^space_log2); vals[i] = mix64(op index ^ seed) (vals may be NULL). */
int lslab_gen_ops(const lslab_gen_params* p, uint8_t* ops, void* keys, void* vals, uint64_t n,
                  hipStream_t stream);

/* Zipf ranks only (1-based), for distribution tests. */
int lslab_gen_zipf_ranks(uint64_t* out, uint64_t n, double theta, uint64_t space, uint64_t seed,
                         uint64_t first, hipStream_t stream);

#ifdef __cplusplus
}
#endif
#endif /* LSLAB_WORKLOAD_H */
