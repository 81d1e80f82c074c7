// The Fiat-Shamir transcript on the device: one lane runs the transcript steps of a round loop
// so that a whole sumcheck needs no host round trip. Byte conventions are those of
// crates/tessera-hash/src/transcript.rs (absorb: state || u64LE(|label|) || label || TCE(value);
// squeeze: state || Ascii(label) || u64LE(counter), then the ratchet state || out). Every input
// a round makes is one BLAKE3 chunk (< 1024 bytes): labels are at most 64 bytes and messages at
// most 11 Fp2.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "blake3.h"
#include "field.h"

namespace tsr {

struct DevKeys {
    uint32_t absorb[8], squeeze[8], ratchet[8];
};

struct DevTranscript {
    uint32_t state[8];
    uint64_t counter;
};

constexpr int LABEL_SLOT = 64;  // bytes per label in a label table

// Hash inputs are assembled in a caller-provided LDS buffer of DEV_BUF_WORDS words (the
// byte-addressed writes and the word reads of the message blocks then stay in LDS instead of
// private scratch memory).
constexpr int DEV_BUF_WORDS = 96;  // 384 bytes: state, label, count and 11 Fp2 fit

// H(key, buf[0..len)) for len < 1024: one chunk, blocks of 64 bytes, the last one ROOT. The
// bytes of buf from len up to the next multiple of 64 are zeroed here.
__device__ inline void dev_hash_chunk(const uint32_t key[8], uint32_t *buf, uint32_t len,
                                      uint32_t out[8]) {
    uint8_t *bytes = (uint8_t *)buf;
    const uint32_t nblocks = len == 0 ? 1 : (len + 63) / 64;
    for (uint32_t at = len; at < 64 * nblocks; at++) bytes[at] = 0;
    uint32_t cv[8];
    for (int i = 0; i < 8; i++) cv[i] = key[i];
    for (uint32_t b = 0; b < nblocks; b++) {
        uint32_t m[16];
#pragma unroll
        for (int i = 0; i < 16; i++) m[i] = buf[16 * b + i];
        uint32_t flags = b3::DERIVE_KEY_MATERIAL;
        if (b == 0) flags |= b3::CHUNK_START;
        uint32_t blen = 64;
        if (b == nblocks - 1) {
            flags |= b3::CHUNK_END | b3::ROOT;
            blen = len - 64 * b;
        }
        b3::compress(cv, m, 0, blen, flags, cv);
    }
    for (int i = 0; i < 8; i++) out[i] = cv[i];
}

__device__ inline uint32_t put_words(uint32_t *buf, uint32_t at, const uint32_t *w, int n) {
    uint8_t *b = (uint8_t *)buf;
    for (int i = 0; i < n; i++)
        for (int j = 0; j < 4; j++) b[at++] = (uint8_t)(w[i] >> (8 * j));
    return at;
}
__device__ inline uint32_t put_u64(uint32_t *buf, uint32_t at, uint64_t v) {
    uint8_t *b = (uint8_t *)buf;
    for (int j = 0; j < 8; j++) b[at++] = (uint8_t)(v >> (8 * j));
    return at;
}
__device__ inline uint32_t put_label(uint32_t *buf, uint32_t at, const uint8_t *label,
                                     uint32_t len) {
    at = put_u64(buf, at, len);
    uint8_t *b = (uint8_t *)buf;
    for (uint32_t i = 0; i < len; i++) b[at++] = label[i];
    return at;
}

// absorb(label, Vec<Fp2>): TCE = u64LE(count) then c0 || c1 of each element (count <= 11).
__device__ inline void dev_absorb_fp2s(DevTranscript &t, const DevKeys &k, const uint8_t *label,
                                       uint32_t label_len, const gl::fp2 *v, uint32_t count,
                                       uint32_t *buf) {
    uint32_t at = put_words(buf, 0, t.state, 8);
    at = put_label(buf, at, label, label_len);
    at = put_u64(buf, at, count);
    for (uint32_t i = 0; i < count; i++) {
        at = put_u64(buf, at, v[i].c0);
        at = put_u64(buf, at, v[i].c1);
    }
    dev_hash_chunk(k.absorb, buf, at, t.state);
}

__device__ inline void dev_squeeze(DevTranscript &t, const DevKeys &k, const uint8_t *label,
                                   uint32_t label_len, uint32_t out[8], uint32_t *buf) {
    uint32_t at = put_words(buf, 0, t.state, 8);
    at = put_label(buf, at, label, label_len);
    at = put_u64(buf, at, t.counter);
    dev_hash_chunk(k.squeeze, buf, at, out);
    t.counter++;
    for (int i = 0; i < 8; i++) {
        buf[i] = t.state[i];
        buf[8 + i] = out[i];
    }
    dev_hash_chunk(k.ratchet, buf, 64, t.state);
}

// challenge_fp2: the first two LE words of a squeeze, squeezing again unless both are below p.
__device__ inline gl::fp2 dev_challenge_fp2(DevTranscript &t, const DevKeys &k,
                                            const uint8_t *label, uint32_t label_len,
                                            uint32_t *buf) {
    for (;;) {
        uint32_t o[8];
        dev_squeeze(t, k, label, label_len, o, buf);
        uint64_t c0 = (uint64_t)o[0] | ((uint64_t)o[1] << 32);
        uint64_t c1 = (uint64_t)o[2] | ((uint64_t)o[3] << 32);
        if (c0 < gl::P && c1 < gl::P) return gl::fp2{c0, c1};
    }
}

}  // namespace tsr
