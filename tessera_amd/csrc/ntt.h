// The row encoder's plan and tables (see ntt.hip).
#pragma once
#include <utility>
#include <vector>
#include "common.h"

namespace tsr {

struct Encoder {
    Ctx *ctx = nullptr;
    uint32_t log_k = 0, log_r = 0, tk = 0, tile_log = 13;
    uint64_t P = 0;
    uint64_t shift_pow_k = 0;
    // (lo, nb) stage-bit ranges of the strided passes, in launch order.
    std::vector<std::pair<uint32_t, uint32_t>> ct_passes, gs_passes;
    DevBuf<uint64_t> zk_inv, zn_fwd, spread_tab, pow7_tab;
    DevBuf<uint64_t> vk_inv, vn_fwd;  // the V powers of zk_inv and zn_fwd for the radix-8 rounds

    // tile_log: log2 of the LDS tile (u64 elements) every pass works on, at most 13 (64 KiB);
    // max_nb: the most stage bits of one strided pass (0: 10 with tile_log 13, else tile_log - 4)
    Status init(Ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P, uint32_t tile_log = 13,
                uint32_t max_nb = 0);
    // u64 elements of scratch encode() needs for `count` rows.
    uint64_t scratch_elems(uint64_t count) const;
    // All pointers are device pointers; launches on `stream`, does not synchronize.
    Status encode(const uint64_t *rows, const uint64_t *masks, uint64_t count,
                  uint64_t *codewords, uint64_t *scratch, hipStream_t stream) const;
    // Unmasked rows addressed through device pointer tables (one entry per row): source, a
    // k-element scratch row, and the n-symbol codeword row. For batches of matrices.
    Status encode_rowmap(const uint64_t *const *src_rows, uint64_t *const *tmp_rows,
                         uint64_t *const *cw_rows, uint64_t count, hipStream_t stream) const;

   private:
    void launch(const uint64_t *src, uint64_t src_stride, const uint64_t *const *src_rows,
                uint64_t *tmp, uint64_t *const *tmp_rows, uint64_t *cw, uint64_t *const *cw_rows,
                const uint64_t *masks, uint64_t rc, hipStream_t stream) const;
};

// The context's cached encoder for (log_k, r, P): tables are built and uploaded once.
Status get_encoder(Ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P, Encoder **out,
                   uint32_t tile_log = 13, uint32_t max_nb = 0);

}  // namespace tsr
