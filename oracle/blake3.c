/*
 * Test infrastructure (oracle): BLAKE3, written from the published BLAKE3 specification
 * (O'Connor, Aumasson, Neves, Wilcox-O'Hearn, 2020), all three modes, any input length
 * (chunk tree with the chaining-value stack) and extendable output.
 *
 * The reference takes BLAKE3 from the crate `blake3` 1.8.7 (Cargo.lock:173-176), in
 * derive-key mode with a precomputed context key per domain
 * (crates/tessera-hash/src/domain.rs:129-135,152-181). Pinned by the 35 official vectors and by
 * the domain digests of spec/vectors/hash/hash.json.
 */
#include <pthread.h>
#include <string.h>
#include "oracle.h"
#include "internal.h"

static const uint32_t IV[8] = {0x6A09E667u, 0xBB67AE85u, 0x3C6EF372u, 0xA54FF53Au,
                               0x510E527Fu, 0x9B05688Cu, 0x1F83D9ABu, 0x5BE0CD19u};
static const uint8_t PERM[16] = {2, 6, 3, 10, 7, 0, 4, 13, 1, 11, 12, 5, 9, 14, 15, 8};

enum { CHUNK_START = 1, CHUNK_END = 2, PARENT = 4, ROOT = 8, KEYED_HASH = 16,
       DERIVE_KEY_CONTEXT = 32, DERIVE_KEY_MATERIAL = 64 };

static inline uint32_t rotr(uint32_t x, int n) { return (x >> n) | (x << (32 - n)); }

static void g(uint32_t *s, int a, int b, int c, int d, uint32_t mx, uint32_t my) {
    s[a] = s[a] + s[b] + mx; s[d] = rotr(s[d] ^ s[a], 16);
    s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 12);
    s[a] = s[a] + s[b] + my; s[d] = rotr(s[d] ^ s[a], 8);
    s[c] = s[c] + s[d];      s[b] = rotr(s[b] ^ s[c], 7);
}

/* The compression function: 16 output words (the first 8 are the chaining value). */
static void compress(const uint32_t cv[8], const uint8_t block[64], uint64_t counter,
                     uint32_t block_len, uint32_t flags, uint32_t out[16]) {
    uint32_t m[16], s[16];
    for (int i = 0; i < 16; i++)
        m[i] = (uint32_t)block[4 * i] | ((uint32_t)block[4 * i + 1] << 8) |
               ((uint32_t)block[4 * i + 2] << 16) | ((uint32_t)block[4 * i + 3] << 24);
    for (int i = 0; i < 8; i++) s[i] = cv[i];
    s[8] = IV[0]; s[9] = IV[1]; s[10] = IV[2]; s[11] = IV[3];
    s[12] = (uint32_t)counter; s[13] = (uint32_t)(counter >> 32);
    s[14] = block_len; s[15] = flags;
    for (int r = 0; r < 7; r++) {
        g(s, 0, 4, 8, 12, m[0], m[1]);   g(s, 1, 5, 9, 13, m[2], m[3]);
        g(s, 2, 6, 10, 14, m[4], m[5]);  g(s, 3, 7, 11, 15, m[6], m[7]);
        g(s, 0, 5, 10, 15, m[8], m[9]);  g(s, 1, 6, 11, 12, m[10], m[11]);
        g(s, 2, 7, 8, 13, m[12], m[13]); g(s, 3, 4, 9, 14, m[14], m[15]);
        if (r < 6) {
            uint32_t t[16];
            for (int i = 0; i < 16; i++) t[i] = m[PERM[i]];
            memcpy(m, t, sizeof m);
        }
    }
    for (int i = 0; i < 8; i++) {
        out[i] = s[i] ^ s[i + 8];
        out[i + 8] = s[i + 8] ^ cv[i];
    }
}

/* What the last compression of the root node needs, kept so XOF output can be replayed. */
typedef struct {
    uint32_t cv[8];
    uint8_t block[64];
    uint32_t block_len, flags;
    uint64_t counter;
} output_t;


static void chunk_init(chunk_t *c, const uint32_t key[8], uint64_t counter, uint32_t flags) {
    memcpy(c->cv, key, 32);
    c->chunk_counter = counter;
    memset(c->block, 0, 64);
    c->block_len = 0;
    c->blocks_compressed = 0;
    c->flags = flags;
}
static size_t chunk_len(const chunk_t *c) { return 64 * (size_t)c->blocks_compressed + c->block_len; }
static uint32_t chunk_start_flag(const chunk_t *c) { return c->blocks_compressed == 0 ? CHUNK_START : 0; }

static void chunk_update(chunk_t *c, const uint8_t *in, size_t len) {
    while (len) {
        if (c->block_len == 64) {
            uint32_t out[16];
            compress(c->cv, c->block, c->chunk_counter, 64, c->flags | chunk_start_flag(c), out);
            memcpy(c->cv, out, 32);
            c->blocks_compressed++;
            memset(c->block, 0, 64);
            c->block_len = 0;
        }
        size_t take = 64 - c->block_len;
        if (take > len) take = len;
        memcpy(c->block + c->block_len, in, take);
        c->block_len += (uint32_t)take;
        in += take;
        len -= take;
    }
}

static output_t chunk_output(const chunk_t *c) {
    output_t o;
    memcpy(o.cv, c->cv, 32);
    memcpy(o.block, c->block, 64);
    o.block_len = c->block_len;
    o.counter = c->chunk_counter;
    o.flags = c->flags | chunk_start_flag(c) | CHUNK_END;
    return o;
}

static void output_cv(const output_t *o, uint32_t cv[8]) {
    uint32_t out[16];
    compress(o->cv, o->block, o->counter, o->block_len, o->flags, out);
    memcpy(cv, out, 32);
}

static output_t parent_output(const uint32_t l[8], const uint32_t r[8], const uint32_t key[8],
                              uint32_t flags) {
    output_t o;
    memcpy(o.cv, key, 32);
    for (int i = 0; i < 8; i++) {
        for (int b = 0; b < 4; b++) {
            o.block[4 * i + b] = (uint8_t)(l[i] >> (8 * b));
            o.block[32 + 4 * i + b] = (uint8_t)(r[i] >> (8 * b));
        }
    }
    o.block_len = 64;
    o.counter = 0;
    o.flags = PARENT | flags;
    return o;
}

static void root_bytes(const output_t *o, uint8_t *out, size_t out_len) {
    uint64_t counter = 0;
    while (out_len) {
        uint32_t w[16];
        compress(o->cv, o->block, counter, o->block_len, o->flags | ROOT, w);
        for (int i = 0; i < 16 && out_len; i++)
            for (int b = 0; b < 4 && out_len; b++, out_len--) *out++ = (uint8_t)(w[i] >> (8 * b));
        counter++;
    }
}


static void hasher_init(hasher_t *h, const uint32_t key[8], uint32_t flags) {
    memcpy(h->key, key, 32);
    h->depth = 0;
    h->flags = flags;
    chunk_init(&h->chunk, key, 0, flags);
}

static void add_chunk_cv(hasher_t *h, uint32_t cv[8], uint64_t total) {
    while ((total & 1) == 0) {
        h->depth--;
        output_t p = parent_output(h->stack[h->depth], cv, h->key, h->flags);
        output_cv(&p, cv);
        total >>= 1;
    }
    memcpy(h->stack[h->depth++], cv, 32);
}

static void hasher_update(hasher_t *h, const uint8_t *in, size_t len) {
    while (len) {
        if (chunk_len(&h->chunk) == 1024) {
            output_t o = chunk_output(&h->chunk);
            uint32_t cv[8];
            output_cv(&o, cv);
            uint64_t total = h->chunk.chunk_counter + 1;
            add_chunk_cv(h, cv, total);
            chunk_init(&h->chunk, h->key, total, h->flags);
        }
        size_t take = 1024 - chunk_len(&h->chunk);
        if (take > len) take = len;
        chunk_update(&h->chunk, in, take);
        in += take;
        len -= take;
    }
}

static void hasher_finalize(const hasher_t *h, uint8_t *out, size_t out_len) {
    output_t o = chunk_output(&h->chunk);
    for (int d = h->depth - 1; d >= 0; d--) {
        uint32_t cv[8];
        output_cv(&o, cv);
        o = parent_output(h->stack[d], cv, h->key, h->flags);
    }
    root_bytes(&o, out, out_len);
}

static void words_of(const uint8_t k[32], uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)k[4 * i] | ((uint32_t)k[4 * i + 1] << 8) | ((uint32_t)k[4 * i + 2] << 16) |
               ((uint32_t)k[4 * i + 3] << 24);
}

void orc_blake3(int mode, const uint8_t *key, size_t key_len, const uint8_t *in, size_t len,
                uint8_t *out, size_t out_len) {
    hasher_t h;
    if (mode == 0) {
        hasher_init(&h, IV, 0);
    } else if (mode == 1) {
        uint32_t w[8];
        words_of(key, w);
        hasher_init(&h, w, KEYED_HASH);
    } else {
        hasher_t ctx;
        uint8_t ck[32];
        uint32_t w[8];
        hasher_init(&ctx, IV, DERIVE_KEY_CONTEXT);
        hasher_update(&ctx, key, key_len);
        hasher_finalize(&ctx, ck, 32);
        words_of(ck, w);
        hasher_init(&h, w, DERIVE_KEY_MATERIAL);
    }
    hasher_update(&h, in, len);
    hasher_finalize(&h, out, out_len);
}

/* Domain names in declaration order (crates/tessera-hash/src/domain.rs:41-81). */
static const char *const DOMAINS[ORC_DOM_COUNT] = {
    "tessera/v1/envelope",          "tessera/v1/commitment-id",   "tessera/v1/column-init",
    "tessera/v1/column-step",       "tessera/v1/merkle-node",     "tessera/v1/column-merge",
    "tessera/v1/weights-root",      "tessera/v1/transcript-init", "tessera/v1/transcript-absorb",
    "tessera/v1/transcript-squeeze","tessera/v1/transcript-ratchet", "tessera/v1/pow",
    "tessera/v1/coin-commit",       "tessera/v1/coin-toss",       "tessera/v1/table",
    "tessera/v1/audit-sample",      "tessera/v1/audit-select",    "tessera/v1/claim-list",
    "tessera/v1/statement",
};

int orc_domain_count(void) { return ORC_DOM_COUNT; }
const char *orc_domain_context(int d) { return (d >= 0 && d < ORC_DOM_COUNT) ? DOMAINS[d] : 0; }

/* Context keys, derived once (domain.rs:129-135): hash_derive_key_context of each string. */
static uint32_t KEYS[ORC_DOM_COUNT][8];
static pthread_once_t KEYS_ONCE = PTHREAD_ONCE_INIT;
static void derive_keys(void) {
    for (int d = 0; d < ORC_DOM_COUNT; d++) {
        hasher_t ctx;
        uint8_t key[32];
        hasher_init(&ctx, IV, DERIVE_KEY_CONTEXT);
        hasher_update(&ctx, (const uint8_t *)DOMAINS[d], strlen(DOMAINS[d]));
        hasher_finalize(&ctx, key, 32);
        words_of(key, KEYS[d]);
    }
}

void orc_domain_key(int d, uint8_t key[32]) {
    pthread_once(&KEYS_ONCE, derive_keys);
    for (int i = 0; i < 8; i++)
        for (int b = 0; b < 4; b++) key[4 * i + b] = (uint8_t)(KEYS[d][i] >> (8 * b));
}

/* Incremental H(domain, .) (domain.rs:138-169). */
void orc_hasher_init(orc_hasher *x, int d) {
    pthread_once(&KEYS_ONCE, derive_keys);
    hasher_init(&x->h, KEYS[d], DERIVE_KEY_MATERIAL);
}
void orc_hasher_update(orc_hasher *x, const void *in, size_t len) { hasher_update(&x->h, in, len); }
void orc_hasher_finalize(const orc_hasher *x, uint8_t out[32]) { hasher_finalize(&x->h, out, 32); }

/* H(domain, bytes) = BLAKE3.derive_key(context, bytes) (domain.rs:177-181). */
void orc_hash(int d, const uint8_t *in, size_t len, uint8_t out[32]) {
    orc_hasher x;
    orc_hasher_init(&x, d);
    orc_hasher_update(&x, in, len);
    orc_hasher_finalize(&x, out);
}
