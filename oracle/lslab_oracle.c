/*
 * lslab oracle -- sequential host replay of the LSlab slab-list map and set.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the timed host baseline); the
 * product under lslab_amd/ never links or loads it.
 *
 * The table is restated from the reference's data structure and traversal:
 *   - bucket = (size_t)key & (2^n - 1)            include/lslab/hash.h:10-15,
 *                                                 include/lslab/map.h:95-96
 *   - each bucket heads a chain of slabs holding slab_keys (31) keys, a valid
 *     bitmap (bit i <-> key[i]) and a next link  detail/slab_node.h:13-35
 *   - a lookup scans a slab's valid keys, then follows next
 *                                                 detail/traverse.h:69-96
 *   - an insert of an absent key uses the LOWEST empty slot of the LAST slab
 *     of the chain (found_empty is reset per slab); when that slab is full it
 *     allocates one slab from a bump pool, links it, and inserts at slot 0
 *                                                 detail/traverse.h:54-55,98-124
 *   - remove clears the valid bit, leaving key and value in place
 *                                                 detail/traverse.h:88-90
 *   - the pool is a bump allocator that never reuses slabs
 *                                                 detail/device_allocator.h:62-73
 *
 * Parity pins for this file live in tests/test_oracle.py: the reference's own
 * known-answer tests (test/slabunified_test.cu:42-89, test/set_test.cu:65-106)
 * restated, a brute-force dict model, closed forms for slab counts and a
 * planted-mistake check of the serial-order checker.
 */
#include "lslab_oracle.h"

#include <stdlib.h>
#include <string.h>

struct orc_table {
    int is_map, key_bytes, val_bytes, cap;
    unsigned n_log2;
    uint64_t nbuckets, pool, used, present;
    uint64_t key_mask, val_mask;
    int hash_mix;     /* bucket by fmix64(key) instead of the key (LSLAB_HASH_MIX) */
    int reclaim;      /* LSLAB_RECLAIM_ON policy (see orc_set_reclaim) */
    uint32_t* freel;  /* free pool slabs, LIFO (reclaim) */
    uint64_t nfree;
    uint64_t* keys;   /* (nbuckets + pool) * 32 */
    uint64_t* vals;   /* maps only, same shape */
    uint32_t* valid;  /* per slab */
    uint32_t* next;   /* per slab, 0 = end of chain (slab 0 is a head, never a link target) */
};

orc_table* orc_create(int is_map, int key_bytes, int val_bytes, unsigned n_log2,
                      uint64_t pool_slabs, int slab_keys) {
    if (n_log2 > 30 || slab_keys < 1 || slab_keys > 31) return NULL;
    if (key_bytes != 4 && key_bytes != 8) return NULL;
    if (is_map && val_bytes != 4 && val_bytes != 8) return NULL;
    orc_table* t = (orc_table*)calloc(1, sizeof(orc_table));
    if (!t) return NULL;
    t->is_map = is_map;
    t->key_bytes = key_bytes;
    t->val_bytes = is_map ? val_bytes : 0;
    t->cap = slab_keys;
    t->n_log2 = n_log2;
    t->nbuckets = 1ull << n_log2;
    t->pool = pool_slabs;
    t->key_mask = key_bytes == 8 ? ~0ull : 0xffffffffull;
    t->val_mask = (is_map && val_bytes == 4) ? 0xffffffffull : ~0ull;
    uint64_t nslabs = t->nbuckets + pool_slabs;
    t->keys = (uint64_t*)calloc(nslabs * 32, sizeof(uint64_t));
    t->vals = is_map ? (uint64_t*)calloc(nslabs * 32, sizeof(uint64_t)) : NULL;
    t->valid = (uint32_t*)calloc(nslabs, sizeof(uint32_t));
    t->next = (uint32_t*)calloc(nslabs, sizeof(uint32_t));
    if (!t->keys || (is_map && !t->vals) || !t->valid || !t->next) {
        orc_destroy(t);
        return NULL;
    }
    return t;
}

void orc_destroy(orc_table* t) {
    if (!t) return;
    free(t->freel);
    free(t->keys);
    free(t->vals);
    free(t->valid);
    free(t->next);
    free(t);
}

/* murmur3's 64-bit finalizer (Appleby, MurmurHash3 fmix64). */
static uint64_t fmix64(uint64_t z) {
    z ^= z >> 33;
    z *= 0xff51afd7ed558ccdull;
    z ^= z >> 33;
    z *= 0xc4ceb9fe1a85ec53ull;
    z ^= z >> 33;
    return z;
}

void orc_set_hash(orc_table* t, int mix) { t->hash_mix = mix != 0; }

int orc_set_reclaim(orc_table* t, int on) {
    t->reclaim = on != 0;
    if (t->reclaim && !t->freel) {
        t->freel = (uint32_t*)calloc(t->pool ? t->pool : 1, sizeof(uint32_t));
        if (!t->freel) return 0;
    }
    return 1;
}

/* Bucket of a key: the reference's identity hash (hash.h:10-15) masked to
 * the table, or fmix64 of the key for a MIX table. */
static uint64_t bucket(const orc_table* t, uint64_t key) {
    return (t->hash_mix ? fmix64(key) : key) & (t->nbuckets - 1);
}

/* Find key: returns 1 and (slab, slot) when present.  Also reports the last
 * slab of the chain for a subsequent insert, the slab before the key's slab
 * (prev, NOSLAB for a head) and, for the reclaim policy, the first slab of
 * the chain with an empty slot (room, NOSLAB if none). */
#define NOSLAB (~0ull)
static int find(const orc_table* t, uint64_t key, uint64_t* slab_out, int* slot_out,
                uint64_t* last_out, uint64_t* prev_out, uint64_t* room_out) {
    const uint32_t capmask = (t->cap == 32) ? 0xffffffffu : ((1u << t->cap) - 1u);
    uint64_t s = bucket(t, key), prev = NOSLAB, room = NOSLAB;
    for (;;) {
        uint32_t v = t->valid[s];
        const uint64_t* k = t->keys + s * 32;
        if (room == NOSLAB && (~v & capmask)) room = s;
        for (int i = 0; i < t->cap; ++i) {
            if (((v >> i) & 1u) && k[i] == key) {
                *slab_out = s;
                *slot_out = i;
                *prev_out = prev;
                *room_out = room;
                return 1;
            }
        }
        uint32_t nx = t->next[s];
        if (nx == 0) {
            *last_out = s;
            *prev_out = prev;
            *room_out = room;
            return 0;
        }
        prev = s;
        s = nx;
    }
}

/* Insert an absent key.  Reference policy (traverse.h:54-55,98-124): the
 * lowest empty slot of the last slab, else slot 0 of a freshly linked pool
 * slab.  Reclaim policy: the lowest empty slot of the first slab with one,
 * else a pool slab taken from the free list (most recently freed first) or
 * the bump pointer, linked after the last slab.  Returns 0 when the pool is
 * full. */
static int insert_absent(orc_table* t, uint64_t last, uint64_t room, uint64_t key, uint64_t val) {
    uint32_t capmask = (t->cap == 32) ? 0xffffffffu : ((1u << t->cap) - 1u);
    uint64_t s = (t->reclaim && room != NOSLAB) ? room : last;
    uint32_t empty = ~t->valid[s] & capmask;
    int slot;
    if (empty) {
        slot = __builtin_ctz(empty);
    } else {
        if (t->reclaim && t->nfree) {
            s = t->freel[--t->nfree];
        } else {
            if (t->used >= t->pool) return 0;
            s = t->nbuckets + t->used++;
        }
        t->next[s] = 0;
        t->next[last] = (uint32_t)s;
        slot = 0;
    }
    t->keys[s * 32 + slot] = key;
    if (t->is_map) t->vals[s * 32 + slot] = val;
    t->valid[s] |= 1u << slot;
    t->present++;
    return 1;
}

/* Clear a present key's valid bit (traverse.h:88-90).  Reclaim policy: a
 * pool slab left empty is unlinked from its chain and freed. */
static void remove_at(orc_table* t, uint64_t slab, int slot, uint64_t prev) {
    t->valid[slab] ^= 1u << slot;
    t->present--;
    if (t->reclaim && slab >= t->nbuckets && t->valid[slab] == 0) {
        t->next[prev] = t->next[slab];
        t->next[slab] = 0;
        t->freel[t->nfree++] = (uint32_t)slab;
    }
}

static int apply_one(orc_table* t, uint8_t op, uint64_t key, uint64_t val, uint8_t* f,
                     uint64_t* out) {
    uint64_t slab = 0, last = 0, prev = NOSLAB, room = NOSLAB;
    int slot = 0;
    key &= t->key_mask;
    val &= t->val_mask;
    int hit = find(t, key, &slab, &slot, &last, &prev, &room);
    *f = 0;
    *out = 0;
    switch (op) {
        case 0: /* get / contains */
            if (hit) {
                *f = 1;
                if (t->is_map) *out = t->vals[slab * 32 + slot];
            }
            return 1;
        case 1: /* put / insert */
            if (hit) {
                *f = 1;
                if (t->is_map) {
                    *out = t->vals[slab * 32 + slot];
                    t->vals[slab * 32 + slot] = val;
                }
                return 1;
            }
            if (!insert_absent(t, last, room, key, val)) return 0;
            *f = t->is_map ? 0 : 1; /* set insert reports true for a new key too */
            return 1;
        case 2: /* remove */
            if (hit) {
                *f = 1;
                if (t->is_map) *out = t->vals[slab * 32 + slot];
                remove_at(t, slab, slot, prev);
            }
            return 1;
        default:
            return 1;
    }
}

int64_t orc_apply(orc_table* t, uint64_t n, const uint8_t* ops, const uint64_t* keys,
                  const uint64_t* vals, uint8_t* found_out, uint64_t* val_out) {
    int64_t failed = 0;
    for (uint64_t i = 0; i < n; ++i) {
        uint8_t f;
        uint64_t o;
        if (!apply_one(t, ops[i], keys[i], vals ? vals[i] : 0, &f, &o)) failed++;
        if (found_out) found_out[i] = f;
        if (val_out) val_out[i] = o;
    }
    return failed;
}

int orc_lookup(const orc_table* t, uint64_t key, uint64_t* val_out) {
    uint64_t slab, last, prev, room;
    int slot;
    key &= t->key_mask;
    if (find(t, key, &slab, &slot, &last, &prev, &room)) {
        if (val_out) *val_out = t->is_map ? t->vals[slab * 32 + slot] : 0;
        return 1;
    }
    if (val_out) *val_out = 0;
    return 0;
}

uint64_t orc_present(const orc_table* t) { return t->present; }
uint64_t orc_slabs_used(const orc_table* t) { return t->used - t->nfree; }

uint64_t orc_max_chain(const orc_table* t) {
    uint64_t best = 0;
    for (uint64_t b = 0; b < t->nbuckets; ++b) {
        uint64_t len = 1, s = b;
        while (t->next[s]) {
            s = t->next[s];
            len++;
        }
        if (len > best) best = len;
    }
    return best;
}

void orc_chain_hist(const orc_table* t, uint64_t* hist, unsigned nbins) {
    for (unsigned i = 0; i < nbins; ++i) hist[i] = 0;
    for (uint64_t b = 0; b < t->nbuckets; ++b) {
        uint64_t len = 1, s = b;
        while (t->next[s]) {
            s = t->next[s];
            len++;
        }
        hist[len < nbins ? len - 1 : nbins - 1]++;
    }
}

uint64_t orc_dump(const orc_table* t, uint64_t* keys, uint64_t* vals, uint64_t cap) {
    uint64_t c = 0;
    for (uint64_t b = 0; b < t->nbuckets; ++b) {
        uint64_t s = b;
        for (;;) {
            uint32_t v = t->valid[s];
            for (int i = 0; i < t->cap; ++i) {
                if ((v >> i) & 1u) {
                    if (c < cap) {
                        if (keys) keys[c] = t->keys[s * 32 + i];
                        if (vals) vals[c] = t->is_map ? t->vals[s * 32 + i] : 0;
                    }
                    c++;
                }
            }
            if (!t->next[s]) break;
            s = t->next[s];
        }
    }
    return c;
}

/* ------------------------------------------------------------------------ */
/* Serial-order checker                                                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    uint64_t key;
    uint64_t idx;
} kidx;

static int cmp_kidx(const void* a, const void* b) {
    const kidx* x = (const kidx*)a;
    const kidx* y = (const kidx*)b;
    if (x->key != y->key) return x->key < y->key ? -1 : 1;
    return x->idx < y->idx ? -1 : (x->idx > y->idx);
}

/* A map state: present flag + value (absent is always (0, 0)). */
typedef struct {
    uint64_t p, v;
} st;

static int cmp_st(const void* a, const void* b) {
    const st* x = (const st*)a;
    const st* y = (const st*)b;
    if (x->p != y->p) return x->p < y->p ? -1 : 1;
    return x->v < y->v ? -1 : (x->v > y->v);
}

static int64_t st_id(const st* nodes, int64_t m, st s) {
    int64_t lo = 0, hi = m - 1;
    while (lo <= hi) {
        int64_t mid = (lo + hi) / 2;
        int c = cmp_st(&nodes[mid], &s);
        if (c == 0) return mid;
        if (c < 0) lo = mid + 1;
        else hi = mid - 1;
    }
    return -1;
}

static int64_t uf_find(int64_t* par, int64_t x) {
    while (par[x] != x) {
        par[x] = par[par[x]];
        x = par[x];
    }
    return x;
}

/* Move the table's state of key to (present, v). */
static void adopt(orc_table* t, uint64_t key, int present, uint64_t v) {
    uint64_t slab, last, prev, room;
    int slot;
    int hit = find(t, key, &slab, &slot, &last, &prev, &room);
    if (present) {
        if (hit) {
            if (t->is_map) t->vals[slab * 32 + slot] = v;
        } else {
            insert_absent(t, last, room, key, v);
        }
    } else if (hit) {
        remove_at(t, slab, slot, prev);
    }
}

/* Map key group check (ops g[0..m-1] all on one key).  Mutations are edges
 * observed-state -> effect-state; a serial order exists iff those edges form
 * one Eulerian trail from the pre-batch state S0 to the end state F, and every
 * get observed S0 or some mutation's effect. */
static int check_map_key(orc_table* t, const kidx* g, uint64_t m, const uint8_t* ops,
                         const uint64_t* vals, const uint8_t* found, const uint64_t* got,
                         const uint8_t* post_found, const uint64_t* post_val, st* end_out) {
    uint64_t key = g[0].key;
    st s0 = {0, 0};
    uint64_t v0;
    if (orc_lookup(t, key, &v0)) {
        s0.p = 1;
        s0.v = v0;
    }
    uint64_t nmut = 0;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t i = g[j].idx;
        if (!found[i] && got[i] != 0) return 0; /* absent must report 0 */
        if (ops[i] != 0) nmut++;
    }
    st* nodes = (st*)malloc(sizeof(st) * (2 * nmut + 2));
    int64_t nn = 0;
    nodes[nn++] = s0;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t i = g[j].idx;
        if (ops[i] == 0) continue;
        st r = {found[i] ? 1u : 0u, found[i] ? got[i] : 0};
        st e = {ops[i] == 1 ? 1u : 0u, ops[i] == 1 ? (vals[i] & t->val_mask) : 0};
        nodes[nn++] = r;
        nodes[nn++] = e;
    }
    qsort(nodes, nn, sizeof(st), cmp_st);
    int64_t u = 0;
    for (int64_t j = 0; j < nn; ++j)
        if (u == 0 || cmp_st(&nodes[u - 1], &nodes[j]) != 0) nodes[u++] = nodes[j];
    int64_t* bal = (int64_t*)calloc(u, sizeof(int64_t));
    int64_t* par = (int64_t*)malloc(sizeof(int64_t) * u);
    char* touched = (char*)calloc(u, 1);
    char* is_effect = (char*)calloc(u, 1);
    for (int64_t j = 0; j < u; ++j) par[j] = j;
    int64_t id0 = st_id(nodes, u, s0);
    is_effect[id0] = 1;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t i = g[j].idx;
        if (ops[i] == 0) continue;
        st r = {found[i] ? 1u : 0u, found[i] ? got[i] : 0};
        st e = {ops[i] == 1 ? 1u : 0u, ops[i] == 1 ? (vals[i] & t->val_mask) : 0};
        int64_t a = st_id(nodes, u, r), b = st_id(nodes, u, e);
        bal[a]++;
        bal[b]--;
        touched[a] = touched[b] = 1;
        is_effect[b] = 1;
        int64_t ra = uf_find(par, a), rb = uf_find(par, b);
        if (ra != rb) par[ra] = rb;
    }
    int ok = 1;
    /* End state: the node with deficit 1, else S0 (the trail is a circuit). */
    int64_t endid = id0;
    for (int64_t j = 0; j < u; ++j) {
        if (j == id0) {
            if (bal[j] != 0 && bal[j] != 1) ok = 0;
        } else if (bal[j] == -1) {
            if (endid != id0) ok = 0;
            endid = j;
        } else if (bal[j] != 0) {
            ok = 0;
        }
    }
    if (ok && bal[id0] == 1 && endid == id0) ok = 0;
    if (ok && bal[id0] == 0 && endid != id0) ok = 0;
    if (ok) {
        int64_t root = uf_find(par, id0);
        for (int64_t j = 0; j < u; ++j)
            if (touched[j] && uf_find(par, j) != root) ok = 0;
    }
    /* Gets observe S0 or some mutation's effect. */
    for (uint64_t j = 0; ok && j < m; ++j) {
        uint64_t i = g[j].idx;
        if (ops[i] != 0) continue;
        st r = {found[i] ? 1u : 0u, found[i] ? got[i] : 0};
        int64_t id = st_id(nodes, u, r);
        if (id < 0 || !is_effect[id]) ok = 0;
    }
    st end = nodes[endid];
    if (ok && post_found) {
        uint64_t i = g[0].idx;
        st pf = {post_found[i] ? 1u : 0u, post_found[i] ? post_val[i] : 0};
        if (cmp_st(&pf, &end) != 0) ok = 0;
    }
    *end_out = end;
    free(nodes);
    free(bal);
    free(par);
    free(touched);
    free(is_effect);
    return ok;
}

/* Set key group check.  States are just present/absent and insert always
 * reports true, so the check counts P->A transitions (removes that reported
 * true) and A->P transitions (inserts issued while absent); see the derivation
 * in the oracle's test file. */
static int set_feasible(int s0, int F, int64_t I, int64_t rt, int64_t rf, int64_t ct,
                        int64_t cf) {
    int64_t a;
    int pphase, aphase;
    if (!s0) {
        /* A->P and P->A transitions alternate starting with A->P. */
        a = rt + (F ? 1 : 0);
        if (a > I) return 0;
        if (I - a > 0 && a < 1) return 0; /* inserts on a present key need a P phase */
        pphase = a >= 1;
        aphase = 1;
    } else {
        /* P->A and A->P transitions alternate starting with P->A. */
        a = rt - (F ? 0 : 1);
        if (a < 0 || a > I) return 0;
        pphase = 1;
        aphase = rt >= 1;
        if (rf > 0 && !aphase) return 0;
    }
    if (ct > 0 && !pphase) return 0;
    if (cf > 0 && !aphase) return 0;
    return 1;
}

/* Set key group check.  States are just present/absent and insert always
 * reports true, so a serial order exists iff the count of P->A transitions
 * (removes that reported true) and A->P transitions (inserts issued while
 * absent) can alternate from S0 to the end state F, with a P phase for every
 * contains=true and an A phase for every remove/contains that saw absent. */
static int check_set_key(orc_table* t, const kidx* g, uint64_t m, const uint8_t* ops,
                         const uint8_t* found, const uint8_t* post_found, int* end_out) {
    uint64_t key = g[0].key;
    int s0 = orc_lookup(t, key, NULL);
    int64_t I = 0, rt = 0, rf = 0, ct = 0, cf = 0;
    for (uint64_t j = 0; j < m; ++j) {
        uint64_t i = g[j].idx;
        if (ops[i] == 1) {
            if (!found[i]) return 0;
            I++;
        } else if (ops[i] == 2) {
            if (found[i]) rt++;
            else rf++;
        } else {
            if (found[i]) ct++;
            else cf++;
        }
    }
    if (post_found) {
        int F = post_found[g[0].idx] ? 1 : 0;
        *end_out = F;
        return set_feasible(s0, F, I, rt, rf, ct, cf);
    }
    for (int F = 0; F < 2; ++F) {
        if (set_feasible(s0, F, I, rt, rf, ct, cf)) {
            *end_out = F;
            return 1;
        }
    }
    *end_out = s0;
    return 0;
}

int64_t orc_check_batch(orc_table* t, uint64_t n, const uint8_t* ops, const uint64_t* keys,
                        const uint64_t* vals, const uint8_t* found, const uint64_t* got,
                        const uint8_t* post_found, const uint64_t* post_val,
                        uint64_t* first_bad) {
    kidx* a = (kidx*)malloc(sizeof(kidx) * (n ? n : 1));
    for (uint64_t i = 0; i < n; ++i) {
        a[i].key = keys[i] & t->key_mask;
        a[i].idx = i;
    }
    qsort(a, n, sizeof(kidx), cmp_kidx);
    int64_t bad = 0;
    uint64_t fb = UINT64_MAX;
    uint64_t j = 0;
    while (j < n) {
        uint64_t k = j;
        while (k < n && a[k].key == a[j].key) k++;
        int ok;
        if (t->is_map) {
            st end;
            ok = check_map_key(t, a + j, k - j, ops, vals, found, got, post_found, post_val,
                               &end);
            if (ok) adopt(t, a[j].key, (int)end.p, end.v);
        } else {
            int end;
            ok = check_set_key(t, a + j, k - j, ops, found, post_found, &end);
            if (ok) adopt(t, a[j].key, end, 0);
        }
        if (!ok) {
            bad++;
            if (a[j].idx < fb) fb = a[j].idx;
        }
        j = k;
    }
    free(a);
    if (first_bad) *first_bad = fb;
    return bad;
}
