// The Fiat-Shamir transcript on the device: one wave runs the transcript steps of a round loop
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

// Hash inputs are assembled in an LDS buffer of DEV_BUF_WORDS words.
constexpr int DEV_BUF_WORDS = 96;  // 384 bytes: state, label, count and 11 Fp2 fit

// One wave runs the transcript: every lane of the wave calls these functions with the same
// arguments (uniform control flow). Lane q = lane mod 4 of each quad holds column q of the
// BLAKE3 state (words q, 4 + q, 8 + q, 12 + q); the diagonal step rotates rows 1-3 across the
// quad with DPP quad permutations, so a compression is 7 x 2 G functions deep instead of 7 x 8.
// Every quad computes the same values. Inputs are assembled in LDS by all 64 lanes at once.

// Orders this wave's LDS accesses (LDS operations of one wave complete in issue order).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Lane q of each quad reads lane (q + k) mod 4: k = 1, 2, 3.
template <int K>
__device__ __forceinline__ uint32_t quad_rot(uint32_t v) {
    constexpr int ctrl = K == 1 ? 0x39 : K == 2 ? 0x4E : 0x93;  // quad_perm selectors
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, ctrl, 0xF, 0xF, false);
}

// Nibble i of B3_SCHED[r] is the index of the i-th message word of round r (the message
// permutation 2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8 applied r times).
constexpr uint64_t B3_SCHED[7] = {0xfedcba9876543210ull, 0x8fe95cb1d407a362ull,
                                  0x18fb0956e72dca43ull, 0x61852b04fd3e9c7aull,
                                  0x461035278eafb9dcull, 0x7462a03d1fc85be9ull,
                                  0xd743c2ae689105fbull};

__device__ __forceinline__ void quad_g(uint32_t &a, uint32_t &b, uint32_t &c, uint32_t &d,
                                       uint32_t mx, uint32_t my) {
    a = a + b + mx;
    d = b3::rotr(d ^ a, 16);
    c = c + d;
    b = b3::rotr(b ^ c, 12);
    a = a + b + my;
    d = b3::rotr(d ^ a, 8);
    c = c + d;
    b = b3::rotr(b ^ c, 7);
}

// The compression of the 16 LDS words at blk (counter 0): (a, b) hold chaining value words
// (q, 4 + q) and receive the output's.
__device__ __forceinline__ void quad_compress(uint32_t &a, uint32_t &b, const uint32_t *blk,
                                              uint32_t blen, uint32_t flags) {
    const uint32_t q = threadIdx.x & 3;
    uint32_t c = q == 0 ? B3_IV0 : q == 1 ? B3_IV1 : q == 2 ? B3_IV2 : B3_IV3;
    uint32_t d = q == 2 ? blen : q == 3 ? flags : 0u;
    uint32_t m[28];
#pragma unroll
    for (int r = 0; r < 7; r++) {
        m[4 * r + 0] = blk[(B3_SCHED[r] >> (8 * q)) & 15];
        m[4 * r + 1] = blk[(B3_SCHED[r] >> (8 * q + 4)) & 15];
        m[4 * r + 2] = blk[(B3_SCHED[r] >> (32 + 8 * q)) & 15];
        m[4 * r + 3] = blk[(B3_SCHED[r] >> (36 + 8 * q)) & 15];
    }
#pragma unroll
    for (int r = 0; r < 7; r++) {
        quad_g(a, b, c, d, m[4 * r], m[4 * r + 1]);
        b = quad_rot<1>(b);
        c = quad_rot<2>(c);
        d = quad_rot<3>(d);
        quad_g(a, b, c, d, m[4 * r + 2], m[4 * r + 3]);
        b = quad_rot<3>(b);
        c = quad_rot<2>(c);
        d = quad_rot<1>(d);
    }
    a ^= c;
    b ^= d;
}

// H(key, buf[0..len)) for len < 1024: one chunk of derive-key material blocks, the last one
// ROOT; the bytes of buf from len to the next multiple of 64 are zero. key (8 words), buf and
// out (8 words) are in LDS.
__device__ inline void wave_hash_chunk(const uint32_t *key, const uint32_t *buf, uint32_t len,
                                       uint32_t *out) {
    const uint32_t q = threadIdx.x & 3;
    const uint32_t nb = len == 0 ? 1 : (len + 63) / 64;
    uint32_t a = key[q];
    uint32_t b = key[4 + q];
    for (uint32_t i = 0; i < nb; i++) {
        uint32_t flags = b3::DERIVE_KEY_MATERIAL | (i == 0 ? b3::CHUNK_START : 0u);
        uint32_t blen = 64;
        if (i == nb - 1) {
            flags |= b3::CHUNK_END | b3::ROOT;
            blen = len - 64 * i;
        }
        quad_compress(a, b, buf + 16 * i, blen, flags);
    }
    if ((threadIdx.x & 63) < 4) {
        out[q] = a;
        out[4 + q] = b;
    }
    wave_sync();
}

// buf = st (8 words) || u64LE(L) || label[0..L) || u64LE(x) || tail[0..tail_len), zero to the
// next multiple of 64 bytes; returns the length. The absorb input with x = count and the TCE of
// the values as tail, or the squeeze input with x = counter and no tail. Word-level: label (L <= 64,
// in a 4-byte aligned slot) and tail (4-byte aligned, tail_len a multiple of 4) are read a word at
// a time, the label's bytes past L masked off, and Y = u64LE(x) || tail, which starts at byte
// 40 + L, placed by a funnel shift of two adjacent Y words.
__device__ inline uint32_t wave_put(uint32_t *buf, const uint32_t *st, const uint8_t *label,
                                    uint32_t L, uint64_t x, const uint8_t *tail, uint32_t tail_len) {
    const uint32_t len = 48 + L + tail_len;
    const uint32_t words = 16 * ((len + 63) / 64);
    const uint32_t *lw = (const uint32_t *)label, *tw = (const uint32_t *)tail;
    const int ny = 2 + (int)(tail_len / 4);  // words of Y
    auto yword = [&](int j) -> uint32_t {
        if (j < 0 || j >= ny) return 0u;
        if (j == 0) return (uint32_t)x;
        if (j == 1) return (uint32_t)(x >> 32);
        return tw[j - 2];
    };
    for (uint32_t w = threadIdx.x & 63; w < words; w += 64) {
        uint32_t v;
        if (w < 8) {
            v = st[w];
        } else if (w < 10) {
            v = w == 8 ? L : 0u;  // u64LE(L), L < 2^32
        } else {
            const uint32_t o = 4 * (w - 10);  // the word's first byte after the length field
            uint32_t lab = 0;
            if (o < L) {
                lab = lw[o >> 2];
                const uint32_t keep = L - o;
                if (keep < 4) lab &= (1u << (8 * keep)) - 1u;
            }
            // Y bytes t .. t + 3 (t < 0: the word's first -t bytes are the label's)
            const int t = (int)o - (int)L;
            uint32_t y = 0;
            if (t > -4) {
                const int j = t >> 2, sh = t & 3;  // floor division
                const uint64_t pair = ((uint64_t)yword(j + 1) << 32) | yword(j);
                y = (uint32_t)(pair >> (8 * sh));
            }
            v = lab | y;
        }
        buf[w] = v;
    }
    wave_sync();
    return len;
}

// The three transcript keys in LDS, as wave_hash_chunk reads them.
struct WaveKeys {
    uint32_t absorb[8], squeeze[8], ratchet[8];
};

// absorb(label, Vec<Fp2>) on the state st (LDS, updated): TCE = u64LE(count) then c0 || c1 of
// each element (count <= 11; vals in LDS).
__device__ inline void wave_absorb_fp2s(uint32_t *st, const WaveKeys &k, const uint8_t *label,
                                        uint32_t label_len, const gl::fp2 *vals, uint32_t count,
                                        uint32_t *buf) {
    const uint32_t len = wave_put(buf, st, label, label_len, count, (const uint8_t *)vals, 16 * count);
    wave_hash_chunk(k.absorb, buf, len, st);
}

// challenge_fp2: squeeze (state || Ascii(label) || u64LE(counter), then the ratchet
// state || out), taking the first two LE words pairs of the output, again unless both are
// below p.
__device__ inline gl::fp2 wave_challenge_fp2(uint32_t *st, uint64_t &counter, const WaveKeys &k,
                                             const uint8_t *label, uint32_t label_len,
                                             uint32_t *buf, uint32_t *out) {
    for (;;) {
        const uint32_t len = wave_put(buf, st, label, label_len, counter, nullptr, 0);
        wave_hash_chunk(k.squeeze, buf, len, out);
        counter++;
        const uint32_t lane = threadIdx.x & 63;
        if (lane < 16) buf[lane] = lane < 8 ? st[lane] : out[lane - 8];
        wave_sync();
        wave_hash_chunk(k.ratchet, buf, 64, st);
        const uint64_t c0 = (uint64_t)out[0] | ((uint64_t)out[1] << 32);
        const uint64_t c1 = (uint64_t)out[2] | ((uint64_t)out[3] << 32);
        if (c0 < gl::P && c1 < gl::P) return gl::fp2{c0, c1};
    }
}

}  // namespace tsr
