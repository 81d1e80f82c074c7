// Column-chain and Merkle launches (see commit.hip). Digests are 8 u32 words; a tree over n
// leaves is stored leaves-first, 2n-1 digests, level l starting at sum_{i<l} n >> i.
#pragma once
#include <vector>
#include "common.h"

namespace tsr {

Status merkle_build(Ctx *ctx, uint32_t *nodes, uint64_t n, hipStream_t stream);
Status chain_leaves(Ctx *ctx, const uint64_t *cw, uint64_t n, uint64_t rows, uint32_t block,
                    const uint8_t id[32], uint32_t *leaves, hipStream_t stream);
Status chains_init(Ctx *ctx, const uint8_t id[32], uint64_t n, uint32_t *chains,
                   hipStream_t stream);
Status chains_absorb(Ctx *ctx, const uint64_t *rows, uint64_t n, uint32_t count,
                     uint32_t *chains, hipStream_t stream);
Status gather_columns(const uint64_t *cw, uint64_t n, uint64_t rows, const uint64_t *idx,
                      uint64_t q, uint64_t *out, hipStream_t stream);
Status gather_digests(const uint32_t *nodes, const uint64_t *idx, uint64_t q, uint32_t *out,
                      hipStream_t stream);
std::vector<uint64_t> multiproof_positions(uint64_t n, std::vector<uint64_t> known);

}  // namespace tsr
