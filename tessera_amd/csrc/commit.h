// Column-chain and Merkle launches (see commit.hip). Digests are 8 u32 words; a tree over n
// leaves is stored leaves-first, 2n-1 digests, level l starting at sum_{i<l} n >> i.
#pragma once
#include <vector>
#include "common.h"
#include "field.h"

namespace tsr {

// Takes the device address of the no-allocation scratch array on `device` (current), once per
// device; tsr_ctx_create calls it.
void noalloc_setup(int device);
// Every level above the n leaves at nodes[0 .. 8n) (leaves first, root last). `timed`: record
// the launches when the context profiles (false: no event is taken, nothing is allocated).
Status merkle_build(Ctx *ctx, uint32_t *nodes, uint64_t n, hipStream_t stream, bool timed = true);
Status chain_leaves(Ctx *ctx, const uint64_t *cw, uint64_t n, uint64_t rows, uint32_t block,
                    const uint8_t id[32], uint32_t *leaves, hipStream_t stream);
// leaf_j = H(column-merge, d_j^(1) || ... || d_j^(S)) from S segments' chains (segment-major).
Status merge_leaves(Ctx *ctx, const uint32_t *digests, uint32_t segments, uint64_t n,
                    uint32_t *leaves, hipStream_t stream);
// The opening's row coefficients ((groups + 1) x N Fp2) and combined rows (opening.hip).
Status row_coefficients(Ctx *ctx, uint32_t row_vars, const uint64_t *rows, uint64_t N,
                        const gl::fp2 *points, const gl::fp2 *gammas, const uint32_t *first,
                        uint32_t groups, const gl::fp2 *xi, gl::fp2 *out, hipStream_t s);
Status combine_rows(Ctx *ctx, const gl::fp2 *coef, uint32_t count, const uint64_t *m, uint64_t N,
                    uint64_t width, gl::fp2 *out, hipStream_t s);
// Zeroes n words (a fill kernel, usable where a memset node is not wanted).
void zero_words(uint64_t *p, uint64_t n, hipStream_t s);
Status chains_init(Ctx *ctx, const uint8_t id[32], uint64_t n, uint32_t *chains,
                   hipStream_t stream);
Status chains_absorb(Ctx *ctx, const uint64_t *rows, uint64_t n, uint32_t count,
                     uint32_t *chains, hipStream_t stream);
Status gather_columns(const uint64_t *cw, uint64_t n, uint64_t rows, const uint64_t *idx,
                      uint64_t q, uint64_t *out, hipStream_t stream);
Status gather_digests(const uint32_t *nodes, const uint64_t *idx, uint64_t q, uint32_t *out,
                      hipStream_t stream);
std::vector<uint64_t> multiproof_positions(uint64_t n, std::vector<uint64_t> known);

// One matrix of a batch, as the multi-matrix kernels see it.
struct MatDesc {
    const uint64_t *cw;  // rows x n codewords
    uint64_t n, rows;
    uint32_t block;
    uint32_t id[8];
    uint32_t *nodes;     // (2n-1) x 8 words, leaves first
};
// A block of 256 consecutive columns (chains) or parents (a Merkle level) of matrix `mat`.
struct BlockRef {
    uint32_t mat, pad_;
    uint64_t first;
};
void launch_chain_leaves_multi(Ctx *ctx, const MatDesc *mats, const BlockRef *refs,
                               uint32_t nblocks, hipStream_t s);
void launch_merkle_level_multi(Ctx *ctx, const MatDesc *mats, const BlockRef *refs,
                               uint32_t nblocks, uint32_t level, hipStream_t s);
void launch_gather_roots(const MatDesc *mats, uint32_t count, uint32_t *roots, hipStream_t s);

}  // namespace tsr

// A commitment handle (tsr_commit_unmasked, tsr_commit, tsr_commit_from_encoded): the codewords
// and the tree stay on the device for its openings. `ctx` is the context whose stream, pool and
// pinned block its calls use; each buffer goes back to the pool it came from.
struct tsr_commitment {
    tsr::Ctx *ctx = nullptr;
    tsr_commitment_params params{};
    uint64_t count = 0, n = 0;
    tsr::DevBuf<uint64_t> codewords;
    tsr::DevBuf<uint32_t> nodes;  // (2n-1) x 8 words
    uint8_t root[32];
    // A masked (advice) commitment keeps its rows (N x k) and masks (N x P) for its openings;
    // they and the codewords are secret-class and zeroed when it is freed.
    tsr::DevBuf<uint64_t> rows, masks;
    bool secret = false;
};
