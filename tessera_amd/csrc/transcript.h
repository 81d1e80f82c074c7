// Host-side BLAKE3 (any length) and the Fiat-Shamir transcript (see transcript.hip).
#pragma once
#include <stddef.h>
#include <stdint.h>
#include "common.h"
#include "field.h"

namespace tsr {

// Incremental BLAKE3 in derive-key-material mode under a domain's context key.
struct Hasher {
    uint32_t key[8], cv[8];
    uint32_t stack[54][8];
    int depth;
    uint64_t chunk_counter;
    uint8_t block[64];
    uint32_t block_len, blocks_compressed;

    void init(const uint32_t key[8]);
    void update(const void *data, size_t len);
    void finalize(uint8_t out[32]) const;

   private:
    void compress_block(uint32_t flags_extra);
    void parent(const uint32_t l[8], const uint32_t r[8], uint32_t flags_extra,
                uint32_t out[8]) const;
};

void hash_domain(const Ctx *ctx, int d, const void *data, size_t len, uint8_t out[32]);
bool label_ok(const char *label);

struct Transcript {
    const Ctx *ctx = nullptr;
    uint8_t state[32];
    uint64_t counter = 0;

    void init(const Ctx *ctx, const uint8_t *context_tce, size_t len);
    void absorb(const char *label, const void *tce, size_t len);
    // A Vec<Fp2> message: u64LE length, then c0 || c1 of each element, 8 LE bytes each.
    void absorb_fp2s(const char *label, const uint64_t *v, size_t count);
    void squeeze(const char *label, uint8_t out[32]);
    uint64_t challenge_fp(const char *label, uint32_t *rejected);
    gl::fp2 challenge_fp2(const char *label, uint32_t *rejected);
    void challenge_indices(const char *label, uint32_t log_n, uint32_t count, uint64_t *out);
    uint64_t grind(const char *label, uint8_t bits);
    bool check_grind(const char *label, uint8_t bits, uint64_t nonce);

   private:
    void begin(Hasher &h, int domain, const char *label) const;
    bool pow_meets(const char *label, uint8_t bits, uint64_t nonce) const;
};

}  // namespace tsr

// The C ABI's opaque transcript handle.
struct tsr_transcript {
    tsr::Transcript t;
};
