// The Fiat-Shamir transcript on the host side of the library, and a full BLAKE3 hasher for it.
//
// Byte conventions follow crates/tessera-hash/src/transcript.rs: state = H(transcript-init,
// TCE(context)); absorb: state = H(transcript-absorb, state || u64LE(|label|) || label ||
// TCE(value)); squeeze: out = H(transcript-squeeze, state || Ascii(label) || u64LE(counter)),
// counter += 1, state = H(transcript-ratchet, state || out). Challenges: Fp = first of the four
// LE words below p; Fp2 = the first two words, squeezing again unless both are below p;
// indices = words masked to log_n bits, four per squeeze.
#include <stdint.h>
#include <string.h>
#include <memory>
#include <string>
#include "blake3.h"
#include "common.h"
#include "field.h"
#include "transcript.h"

namespace tsr {

static void load_words(const uint8_t *b, uint32_t *w, int n) {
    for (int i = 0; i < n; i++)
        w[i] = (uint32_t)b[4 * i] | ((uint32_t)b[4 * i + 1] << 8) | ((uint32_t)b[4 * i + 2] << 16) |
               ((uint32_t)b[4 * i + 3] << 24);
}

static void store_words(const uint32_t *w, uint8_t *b, int n) {
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 4; j++) b[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
}

void Hasher::init(const uint32_t k[8]) {
    memcpy(key, k, 32);
    memcpy(cv, k, 32);
    depth = 0;
    chunk_counter = 0;
    block_len = 0;
    blocks_compressed = 0;
    memset(block, 0, 64);
}

void Hasher::compress_block(uint32_t flags_extra) {
    uint32_t w[16];
    load_words(block, w, 16);
    uint32_t flags = b3::DERIVE_KEY_MATERIAL | flags_extra;
    if (blocks_compressed == 0) flags |= b3::CHUNK_START;
    b3::compress(cv, w, chunk_counter, 64, flags, cv);
    blocks_compressed++;
    block_len = 0;
    memset(block, 0, 64);
}

void Hasher::parent(const uint32_t l[8], const uint32_t r[8], uint32_t flags_extra,
                    uint32_t out[8]) const {
    uint32_t w[16];
    memcpy(w, l, 32);
    memcpy(w + 8, r, 32);
    b3::compress(key, w, 0, 64, b3::PARENT | b3::DERIVE_KEY_MATERIAL | flags_extra, out);
}

void Hasher::update(const void *data, size_t len) {
    const uint8_t *in = (const uint8_t *)data;
    while (len) {
        if (64 * (size_t)blocks_compressed + block_len == 1024) {
            // the chunk is complete: its chaining value joins the stack
            uint32_t w[16], chunk_cv[8];
            load_words(block, w, 16);
            uint32_t flags = b3::DERIVE_KEY_MATERIAL | b3::CHUNK_END |
                             (blocks_compressed == 0 ? b3::CHUNK_START : 0);
            b3::compress(cv, w, chunk_counter, block_len, flags, chunk_cv);
            uint64_t total = chunk_counter + 1;
            while ((total & 1) == 0) {
                depth--;
                parent(stack[depth], chunk_cv, 0, chunk_cv);
                total >>= 1;
            }
            memcpy(stack[depth++], chunk_cv, 32);
            chunk_counter++;
            memcpy(cv, key, 32);
            blocks_compressed = 0;
            block_len = 0;
            memset(block, 0, 64);
        }
        if (block_len == 64) compress_block(0);
        size_t take = 64 - block_len;
        size_t room = 1024 - (64 * (size_t)blocks_compressed + block_len);
        if (take > room) take = room;
        if (take > len) take = len;
        memcpy(block + block_len, in, take);
        block_len += (uint32_t)take;
        in += take;
        len -= take;
    }
}

void Hasher::finalize(uint8_t out[32]) const {
    uint32_t w[16], res[8];
    load_words(block, w, 16);
    uint32_t flags = b3::DERIVE_KEY_MATERIAL | b3::CHUNK_END |
                     (blocks_compressed == 0 ? b3::CHUNK_START : 0);
    if (depth == 0) {
        b3::compress(cv, w, chunk_counter, block_len, flags | b3::ROOT, res);
        store_words(res, out, 8);
        return;
    }
    uint32_t node[8];
    b3::compress(cv, w, chunk_counter, block_len, flags, node);
    for (int d = depth - 1; d >= 0; d--) {
        parent(stack[d], node, d == 0 ? b3::ROOT : 0, node);
    }
    store_words(node, out, 8);
}

// Context keys for host-only use (no context needed): derived once, like each Ctx's.
static const uint32_t (*host_keys())[8] {
    static uint32_t keys[DOM_COUNT][8];
    static bool ready = [] {
        derive_domain_keys(keys);
        return true;
    }();
    (void)ready;
    return keys;
}

static const uint32_t *key_for(const Ctx *ctx, int d) { return ctx ? ctx->keys[d] : host_keys()[d]; }

void hash_domain(const Ctx *ctx, int d, const void *data, size_t len, uint8_t out[32]) {
    Hasher h;
    h.init(key_for(ctx, d));
    h.update(data, len);
    h.finalize(out);
}

static void put_u64(uint8_t *b, uint64_t v) {
    for (int i = 0; i < 8; i++) b[i] = (uint8_t)(v >> (8 * i));
}
static uint64_t get_u64(const uint8_t *b) {
    uint64_t v = 0;
    for (int i = 0; i < 8; i++) v |= (uint64_t)b[i] << (8 * i);
    return v;
}

bool label_ok(const char *label) {
    for (const char *c = label; *c; c++)
        if ((unsigned char)*c < 0x20 || (unsigned char)*c > 0x7E) return false;
    return true;
}

void Transcript::init(const Ctx *c, const uint8_t *context_tce, size_t len) {
    ctx = c;
    hash_domain(ctx, DOM_TRANSCRIPT_INIT, context_tce, len, state);
    counter = 0;
}

void Transcript::begin(Hasher &h, int domain, const char *label) const {
    uint8_t lb[8];
    size_t n = strlen(label);
    h.init(key_for(ctx, domain));
    h.update(state, 32);
    put_u64(lb, n);
    h.update(lb, 8);
    h.update(label, n);
}

void Transcript::absorb(const char *label, const void *tce, size_t len) {
    Hasher h;
    begin(h, DOM_TRANSCRIPT_ABSORB, label);
    h.update(tce, len);
    h.finalize(state);
}

void Transcript::absorb_fp2s(const char *label, const uint64_t *v, size_t count) {
    Hasher h;
    uint8_t b[8];
    begin(h, DOM_TRANSCRIPT_ABSORB, label);
    put_u64(b, count);
    h.update(b, 8);
    for (size_t i = 0; i < 2 * count; i++) {
        put_u64(b, v[i]);
        h.update(b, 8);
    }
    h.finalize(state);
}

void Transcript::squeeze(const char *label, uint8_t out[32]) {
    Hasher h;
    uint8_t b[8], in[64];
    begin(h, DOM_TRANSCRIPT_SQUEEZE, label);
    put_u64(b, counter);
    h.update(b, 8);
    h.finalize(out);
    counter++;
    memcpy(in, state, 32);
    memcpy(in + 32, out, 32);
    hash_domain(ctx, DOM_TRANSCRIPT_RATCHET, in, 64, state);
}

uint64_t Transcript::challenge_fp(const char *label, uint32_t *rejected) {
    uint32_t rej = 0;
    for (;;) {
        uint8_t out[32];
        squeeze(label, out);
        for (int i = 0; i < 4; i++) {
            uint64_t w = get_u64(out + 8 * i);
            if (w < gl::P) {
                if (rejected) *rejected = rej;
                return w;
            }
            rej++;
        }
    }
}

gl::fp2 Transcript::challenge_fp2(const char *label, uint32_t *rejected) {
    uint32_t rej = 0;
    for (;;) {
        uint8_t out[32];
        squeeze(label, out);
        uint64_t c0 = get_u64(out), c1 = get_u64(out + 8);
        if (c0 < gl::P && c1 < gl::P) {
            if (rejected) *rejected = rej;
            return gl::fp2{c0, c1};
        }
        rej++;
    }
}

void Transcript::challenge_indices(const char *label, uint32_t log_n, uint32_t count,
                                   uint64_t *out) {
    uint64_t mask = log_n == 0 ? 0 : (~0ull >> (64 - log_n));
    uint32_t have = 0;
    while (have < count) {
        uint8_t o[32];
        squeeze(label, o);
        for (int i = 0; i < 4 && have < count; i++) out[have++] = get_u64(o + 8 * i) & mask;
    }
}

static uint32_t leading_zero_bits(const uint8_t d[32]) {
    uint32_t z = 0;
    for (int i = 0; i < 32; i++) {
        if (d[i]) return z + (uint32_t)__builtin_clz((unsigned)d[i]) - 24;
        z += 8;
    }
    return z;
}

bool Transcript::pow_meets(const char *label, uint8_t bits, uint64_t nonce) const {
    Hasher h;
    uint8_t nb[8], d[32];
    begin(h, DOM_POW, label);
    put_u64(nb, nonce);
    h.update(nb, 8);
    h.finalize(d);
    return leading_zero_bits(d) >= bits;
}

uint64_t Transcript::grind(const char *label, uint8_t bits) {
    uint64_t nonce = 0;
    while (!pow_meets(label, bits, nonce)) nonce++;
    uint8_t nb[8];
    put_u64(nb, nonce);
    absorb(label, nb, 8);
    return nonce;
}

bool Transcript::check_grind(const char *label, uint8_t bits, uint64_t nonce) {
    bool ok = pow_meets(label, bits, nonce);
    if (ok) {
        uint8_t nb[8];
        put_u64(nb, nonce);
        absorb(label, nb, 8);
    }
    return ok;
}

}  // namespace tsr

using namespace tsr;


#define T_CHECK(cond, status, msg)                          \
    do {                                                    \
        if (!(cond)) return Status::fail(status, msg).code; \
    } while (0)

extern "C" {

tsr_status tsr_transcript_create(tsr_ctx *ctx, const uint8_t *context_tce, uint64_t len,
                                 tsr_transcript **out) {
    T_CHECK(out && (len == 0 || context_tce), TSR_BAD_ARG, "null argument");
    std::unique_ptr<tsr_transcript> t(new tsr_transcript());
    t->t.init((const Ctx *)ctx, context_tce, len);
    *out = t.release();
    return TSR_OK;
}

tsr_status tsr_transcript_from_state(tsr_ctx *ctx, const uint8_t state[32], uint64_t counter,
                                     tsr_transcript **out) {
    T_CHECK(out && state, TSR_BAD_ARG, "null argument");
    std::unique_ptr<tsr_transcript> t(new tsr_transcript());
    t->t.ctx = (const Ctx *)ctx;
    memcpy(t->t.state, state, 32);
    t->t.counter = counter;
    *out = t.release();
    return TSR_OK;
}

void tsr_transcript_free(tsr_transcript *t) { delete t; }

tsr_status tsr_transcript_state(const tsr_transcript *t, uint8_t state[32], uint64_t *counter) {
    T_CHECK(t, TSR_BAD_ARG, "null transcript");
    if (state) memcpy(state, t->t.state, 32);
    if (counter) *counter = t->t.counter;
    return TSR_OK;
}

tsr_status tsr_transcript_absorb(tsr_transcript *t, const char *label, const uint8_t *tce,
                                 uint64_t len) {
    T_CHECK(t && label && (len == 0 || tce), TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    t->t.absorb(label, tce, len);
    return TSR_OK;
}

tsr_status tsr_transcript_squeeze(tsr_transcript *t, const char *label, uint8_t out[32]) {
    T_CHECK(t && label && out, TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    t->t.squeeze(label, out);
    return TSR_OK;
}

tsr_status tsr_transcript_challenge_fp(tsr_transcript *t, const char *label, uint64_t *out) {
    T_CHECK(t && label && out, TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    *out = t->t.challenge_fp(label, nullptr);
    return TSR_OK;
}

tsr_status tsr_transcript_challenge_fp2(tsr_transcript *t, const char *label, uint64_t out[2]) {
    T_CHECK(t && label && out, TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    gl::fp2 v = t->t.challenge_fp2(label, nullptr);
    out[0] = v.c0;
    out[1] = v.c1;
    return TSR_OK;
}

tsr_status tsr_transcript_challenge_indices(tsr_transcript *t, const char *label,
                                            uint32_t log_n, uint32_t count, uint64_t *out) {
    T_CHECK(t && label && (count == 0 || out), TSR_BAD_ARG, "null argument");
    T_CHECK(log_n <= 64, TSR_BAD_SHAPE, "an index domain of 2^log_n exceeds u64");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    t->t.challenge_indices(label, log_n, count, out);
    return TSR_OK;
}

tsr_status tsr_transcript_grind(tsr_transcript *t, const char *label, uint8_t bits,
                                uint64_t *nonce) {
    T_CHECK(t && label && nonce, TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    *nonce = t->t.grind(label, bits);
    return TSR_OK;
}

tsr_status tsr_transcript_check_grind(tsr_transcript *t, const char *label, uint8_t bits,
                                      uint64_t nonce, int *valid) {
    T_CHECK(t && label && valid, TSR_BAD_ARG, "null argument");
    T_CHECK(label_ok(label), TSR_BAD_ARG, "a transcript label is printable ASCII");
    *valid = t->t.check_grind(label, bits, nonce) ? 1 : 0;
    return TSR_OK;
}

tsr_status tsr_hash(tsr_ctx *ctx, int domain, const uint8_t *data, uint64_t len,
                    uint8_t out[32]) {
    T_CHECK(out && (len == 0 || data), TSR_BAD_ARG, "null argument");
    T_CHECK(domain >= 0 && domain < DOM_COUNT, TSR_BAD_ARG, "unknown domain");
    hash_domain((const Ctx *)ctx, domain, data, len, out);
    return TSR_OK;
}

}  // extern "C"
