/*
 * Test infrastructure (oracle): the Fiat-Shamir transcript.
 * Restates crates/tessera-hash/src/transcript.rs: state = H(transcript-init, TCE(context));
 * absorb = H(transcript-absorb, state || u64LE(len label) || label || TCE(value));
 * squeeze = H(transcript-squeeze, state || Ascii(label) || u64LE(counter)), then
 * state = H(transcript-ratchet, state || out).
 */
#include <string.h>
#include "oracle.h"
#include "internal.h"

static void begin(const orc_transcript *t, orc_hasher *h, int domain, const char *label) {
    uint8_t len[8];
    size_t n = strlen(label);
    orc_hasher_init(h, domain);
    orc_hasher_update(h, t->state, 32);
    orc_put_u64(len, n); /* label_prefix (transcript.rs:294-300) */
    orc_hasher_update(h, len, 8);
    orc_hasher_update(h, label, n);
}

/* Transcript::with_hash (transcript.rs:143-153). */
void orc_transcript_init(orc_transcript *t, const uint8_t *context_tce, size_t len) {
    orc_hash(ORC_DOM_TRANSCRIPT_INIT, context_tce, len, t->state);
    t->counter = 0;
}

/* absorb (transcript.rs:168-176): the TCE bytes are given whole. */
void orc_transcript_absorb(orc_transcript *t, const char *label, const uint8_t *tce, size_t len) {
    orc_hasher h;
    begin(t, &h, ORC_DOM_TRANSCRIPT_ABSORB, label);
    orc_hasher_update(&h, tce, len);
    orc_hasher_finalize(&h, t->state);
}

/* squeeze (transcript.rs:178-193). */
void orc_transcript_squeeze(orc_transcript *t, const char *label, uint8_t out[32]) {
    orc_hasher h;
    uint8_t ctr[8], in[64];
    begin(t, &h, ORC_DOM_TRANSCRIPT_SQUEEZE, label);
    orc_put_u64(ctr, t->counter);
    orc_hasher_update(&h, ctr, 8);
    orc_hasher_finalize(&h, out);
    t->counter++;
    memcpy(in, t->state, 32);
    memcpy(in + 32, out, 32);
    orc_hash(ORC_DOM_TRANSCRIPT_RATCHET, in, 64, t->state);
}

/* challenge_fp (transcript.rs:195-206): first of four LE words below p. */
uint64_t orc_transcript_challenge_fp(orc_transcript *t, const char *label, int *rejected) {
    int rej = 0;
    for (;;) {
        uint8_t out[32];
        orc_transcript_squeeze(t, label, out);
        for (int i = 0; i < 4; i++) {
            uint64_t w = orc_get_u64(out + 8 * i);
            if (w < ORC_P) {
                if (rejected) *rejected = rej;
                return w;
            }
            rej++;
        }
    }
}

/* challenge_fp2 (transcript.rs:208-221): first two words, squeeze again unless both < p. */
orc_fp2 orc_transcript_challenge_fp2(orc_transcript *t, const char *label, int *rejected) {
    int rej = 0;
    for (;;) {
        uint8_t out[32];
        orc_transcript_squeeze(t, label, out);
        uint64_t c0 = orc_get_u64(out), c1 = orc_get_u64(out + 8);
        if (c0 < ORC_P && c1 < ORC_P) {
            orc_fp2 r = {c0, c1};
            if (rejected) *rejected = rej;
            return r;
        }
        rej++;
    }
}

void orc_transcript_challenge_fp2_v(orc_transcript *t, const char *label, uint64_t out[2]) {
    orc_fp2 r = orc_transcript_challenge_fp2(t, label, NULL);
    out[0] = r.c0;
    out[1] = r.c1;
}

/* challenge_indices (transcript.rs:223-239). */
void orc_transcript_challenge_indices(orc_transcript *t, const char *label, uint32_t log_n,
                                      uint32_t count, uint64_t *out) {
    uint64_t mask = log_n == 0 ? 0 : (~0ULL >> (64 - log_n));
    uint32_t have = 0;
    while (have < count) {
        uint8_t o[32];
        orc_transcript_squeeze(t, label, o);
        for (int i = 0; i < 4 && have < count; i++) out[have++] = orc_get_u64(o + 8 * i) & mask;
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

static int pow_meets(const orc_transcript *t, const char *label, uint8_t bits, uint64_t nonce) {
    orc_hasher h;
    uint8_t nb[8], d[32];
    begin(t, &h, ORC_DOM_POW, label);
    orc_put_u64(nb, nonce);
    orc_hasher_update(&h, nb, 8);
    orc_hasher_finalize(&h, d);
    return leading_zero_bits(d) >= bits;
}

/* grind (transcript.rs:245-253) and check_grind (:265-272). */
uint64_t orc_transcript_grind(orc_transcript *t, const char *label, uint8_t bits) {
    uint64_t nonce = 0;
    while (!pow_meets(t, label, bits, nonce)) nonce++;
    uint8_t nb[8];
    orc_put_u64(nb, nonce);
    orc_transcript_absorb(t, label, nb, 8);
    return nonce;
}
int orc_transcript_check_grind(orc_transcript *t, const char *label, uint8_t bits,
                               uint64_t nonce) {
    int ok = pow_meets(t, label, bits, nonce);
    if (ok) {
        uint8_t nb[8];
        orc_put_u64(nb, nonce);
        orc_transcript_absorb(t, label, nb, 8);
    }
    return ok;
}

/* A Vec<Fp2> message as TCE: u64LE length, then c0, c1 LE of each element
 * (crates/tessera-sumcheck/src/transcript.rs:15-44, tessera-codec tce.rs). */
void orc_transcript_absorb_fp2s(orc_transcript *t, const char *label, const uint64_t *v,
                                size_t count) {
    orc_hasher h;
    uint8_t b[8];
    begin(t, &h, ORC_DOM_TRANSCRIPT_ABSORB, label);
    orc_put_u64(b, count);
    orc_hasher_update(&h, b, 8);
    for (size_t i = 0; i < 2 * count; i++) {
        orc_put_u64(b, v[i]);
        orc_hasher_update(&h, b, 8);
    }
    orc_hasher_finalize(&h, t->state);
}
