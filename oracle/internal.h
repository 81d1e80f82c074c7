/* Test infrastructure (oracle): helpers shared between the oracle's own files. */
#ifndef TESSERA_ORACLE_INTERNAL_H
#define TESSERA_ORACLE_INTERNAL_H
#include <stddef.h>
#include <stdint.h>

/* BLAKE3 chunk state and incremental hasher with the chaining-value stack. */
typedef struct {
    uint32_t cv[8];
    uint64_t chunk_counter;
    uint8_t block[64];
    uint32_t block_len, blocks_compressed, flags;
} chunk_t;
typedef struct {
    chunk_t chunk;
    uint32_t key[8];
    uint32_t stack[54][8];
    int depth;
    uint32_t flags;
} hasher_t;

/* Incremental H(domain, .): one BLAKE3 derive-key-material hasher. */
typedef struct orc_hasher { hasher_t h; } orc_hasher;
void orc_hasher_init(orc_hasher *x, int domain);
void orc_hasher_update(orc_hasher *x, const void *in, size_t len);
void orc_hasher_finalize(const orc_hasher *x, uint8_t out[32]);

/* Run fn(arg, lo, hi) over [0, count) split into `threads` contiguous ranges on pthreads. */
void orc_parallel_for(uint64_t count, int threads, void (*fn)(void *, uint64_t, uint64_t),
                      void *arg);
/* The same, fn also given its range's index t in [0, threads). */
void orc_parallel_for_t(uint64_t count, int threads,
                        void (*fn)(void *, int, uint64_t, uint64_t), void *arg);

static inline void orc_put_u64(uint8_t *p, uint64_t v) {
    for (int i = 0; i < 8; i++) p[i] = (uint8_t)(v >> (8 * i));
}
static inline uint64_t orc_get_u64(const uint8_t *p) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)p[i] << (8 * i);
    return v;
}
#endif
