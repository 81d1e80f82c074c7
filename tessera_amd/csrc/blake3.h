// BLAKE3 for gfx950 kernels and the library's host side, written from the BLAKE3
// specification. The reference hashes every domain in derive-key mode with a precomputed
// context key (crates/tessera-hash/src/domain.rs:129-181, crate blake3 1.8.7); the inputs the
// kernels hash are all single-chunk (<= 1024 bytes): column-init (40 B), column-step
// (32 + 8*count B, count <= 124) and merkle-node (64 B).
//
// Digests are kept as 8 little-endian u32 words, the byte order of the reference's [u8; 32].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace b3 {

enum : uint32_t {
    CHUNK_START = 1,
    CHUNK_END = 2,
    PARENT = 4,
    ROOT = 8,
    KEYED_HASH = 16,
    DERIVE_KEY_CONTEXT = 32,
    DERIVE_KEY_MATERIAL = 64,
};

#define B3_IV0 0x6A09E667u
#define B3_IV1 0xBB67AE85u
#define B3_IV2 0x3C6EF372u
#define B3_IV3 0xA54FF53Au
#define B3_IV4 0x510E527Fu
#define B3_IV5 0x9B05688Cu
#define B3_IV6 0x1F83D9ABu
#define B3_IV7 0x5BE0CD19u

__host__ __device__ __forceinline__ uint32_t rotr(uint32_t x, int n) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_alignbit(x, x, n);
#else
    return (x >> n) | (x << (32 - n));
#endif
}

#define B3_G(a, b, c, d, mx, my)           \
    do {                                   \
        a = a + b + (mx);                  \
        d = rotr(d ^ a, 16);               \
        c = c + d;                         \
        b = rotr(b ^ c, 12);               \
        a = a + b + (my);                  \
        d = rotr(d ^ a, 8);                \
        c = c + d;                         \
        b = rotr(b ^ c, 7);                \
    } while (0)

// One round on state s with message words m (already permuted for this round).
#define B3_ROUND(s, m)                                          \
    do {                                                        \
        B3_G(s[0], s[4], s[8], s[12], m[0], m[1]);              \
        B3_G(s[1], s[5], s[9], s[13], m[2], m[3]);              \
        B3_G(s[2], s[6], s[10], s[14], m[4], m[5]);             \
        B3_G(s[3], s[7], s[11], s[15], m[6], m[7]);             \
        B3_G(s[0], s[5], s[10], s[15], m[8], m[9]);             \
        B3_G(s[1], s[6], s[11], s[12], m[10], m[11]);           \
        B3_G(s[2], s[7], s[8], s[13], m[12], m[13]);            \
        B3_G(s[3], s[4], s[9], s[14], m[14], m[15]);            \
    } while (0)

// The compression function; writes the 8-word chaining value (first half of the output) into
// cv_out, and, when full != nullptr, all 16 output words (for extendable output).
__host__ __device__ __forceinline__ void compress(const uint32_t cv[8], const uint32_t block[16],
                                                  uint64_t counter, uint32_t block_len,
                                                  uint32_t flags, uint32_t cv_out[8],
                                                  uint32_t *full = nullptr) {
    uint32_t s[16] = {cv[0], cv[1], cv[2], cv[3], cv[4], cv[5], cv[6], cv[7],
                      B3_IV0, B3_IV1, B3_IV2, B3_IV3, (uint32_t)counter,
                      (uint32_t)(counter >> 32), block_len, flags};
    uint32_t m[16];
#pragma unroll
    for (int i = 0; i < 16; i++) m[i] = block[i];
#pragma unroll
    for (int r = 0; r < 7; r++) {
        B3_ROUND(s, m);
        if (r < 6) {
            // message permutation 2,6,3,10,7,0,4,13,1,11,12,5,9,14,15,8
            uint32_t t[16] = {m[2], m[6], m[3], m[10], m[7], m[0], m[4], m[13],
                              m[1], m[11], m[12], m[5], m[9], m[14], m[15], m[8]};
#pragma unroll
            for (int i = 0; i < 16; i++) m[i] = t[i];
        }
    }
    if (full) {
#pragma unroll
        for (int i = 0; i < 8; i++) {
            full[i] = s[i] ^ s[i + 8];
            full[i + 8] = s[i + 8] ^ cv[i];
        }
    }
#pragma unroll
    for (int i = 0; i < 8; i++) cv_out[i] = s[i] ^ s[i + 8];
}

// H(domain, left || right) for a 64-byte input: one block, root of a one-chunk tree.
__host__ __device__ __forceinline__ void hash64(const uint32_t key[8], const uint32_t l[8],
                                                const uint32_t r[8], uint32_t out[8]) {
    uint32_t blk[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        blk[i] = l[i];
        blk[8 + i] = r[i];
    }
    compress(key, blk, 0, 64, CHUNK_START | CHUNK_END | ROOT | DERIVE_KEY_MATERIAL, out);
}

// H(domain, id || u64LE(j)) for the 40-byte column-init input.
__host__ __device__ __forceinline__ void hash_id_index(const uint32_t key[8], const uint32_t id[8],
                                                       uint64_t j, uint32_t out[8]) {
    uint32_t blk[16];
#pragma unroll
    for (int i = 0; i < 8; i++) blk[i] = id[i];
    blk[8] = (uint32_t)j;
    blk[9] = (uint32_t)(j >> 32);
#pragma unroll
    for (int i = 10; i < 16; i++) blk[i] = 0;
    compress(key, blk, 0, 40, CHUNK_START | CHUNK_END | ROOT | DERIVE_KEY_MATERIAL, out);
}

}  // namespace b3
