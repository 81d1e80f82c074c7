// The reference's device sinks as C objects: a block sink (BlockSink: encode a block of rows with
// masks, absorb the codewords into every column chain, crates/tessera-pcs/src/streaming.rs:24-38)
// and a matrix sink (MatrixSink: the same without masks, also returning each row's symbols at the
// columns the sink was opened with, crates/tessera-pcs/src/weights/regenerated.rs:24-62). Both
// keep their rows, masks and codewords in their own device buffers, which are zeroed when the
// sink is freed (the rows are secret-class, as the reference's sinks clear theirs on drop).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "commit.h"
#include "common.h"
#include "ntt.h"

struct tsr_block_sink {
    tsr::Ctx *ctx = nullptr;
    uint32_t log_k = 0, log_r = 0;
    uint64_t P = 0, block = 0, k = 0, n = 0;
    bool closed = false, finished = false;
    tsr::DevBuf<uint64_t> rows, masks, codewords, queries, out;
    tsr::DevBuf<uint32_t> chains;
    uint64_t q = 0;
};

namespace tsr {

// out[r * q + i] = cw[r * n + idx[i]]: each row's symbols at the queried columns, row by row.
__global__ void k_gather_row_symbols(const uint64_t *__restrict__ cw, uint64_t n, uint64_t rows,
                                     const uint64_t *__restrict__ idx, uint64_t q,
                                     uint64_t *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= q * rows) return;
    const uint64_t r = t / q, i = t % q;
    out[t] = cw[r * n + idx[i]];
}

__global__ void k_zero_words(uint64_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static void zero(DevBuf<uint64_t> &b, hipStream_t s) {
    if (b.ptr && b.len)
        hipLaunchKernelGGL(k_zero_words, dim3((unsigned)((b.len + 255) / 256)), dim3(256), 0, s, b.ptr, b.len);
}

static Status sink_create(Ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P, uint64_t block,
                          const uint8_t id[32], tsr_block_sink **out) {
    if ((uint64_t)log_k + log_r > 32 || log_r == 0)
        return Status::fail(TSR_BAD_SHAPE, "TooLong or rate: n = 2^(log_k + r) with 1 <= r, at most 2^32");
    if (P > (1ull << log_k)) return Status::fail(TSR_BAD_SHAPE, "MaskTooLarge: mask_coeffs exceeds k");
    if (block == 0 || block > 124) return Status::fail(TSR_BAD_SHAPE, "BlockSize: a block holds 1 to 124 rows");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    std::unique_ptr<tsr_block_sink> s(new tsr_block_sink());
    s->ctx = ctx;
    s->log_k = log_k;
    s->log_r = log_r;
    s->P = P;
    s->block = block;
    s->k = 1ull << log_k;
    s->n = s->k << log_r;
    Encoder *enc = nullptr;
    Status st;
    if ((st = get_encoder(ctx, log_k, log_r, P, &enc))) return st;
    if ((st = s->rows.alloc(ctx, block * s->k)) || (st = s->masks.alloc(ctx, std::max<uint64_t>(1, block * P))) ||
        (st = s->codewords.alloc(ctx, block * s->n)) || (st = s->chains.alloc(ctx, 8 * s->n)))
        return st;
    if ((st = chains_init(ctx, id, s->n, s->chains.ptr, ctx->stream))) return st;
    *out = s.release();
    return Status::ok();
}

// Encodes `count` rows (and masks) into the sink's codewords and absorbs them into the chains.
static Status sink_absorb(tsr_block_sink *s, const uint64_t *rows, const uint64_t *masks,
                          uint64_t count) {
    Ctx *ctx = s->ctx;
    if (s->finished) return Status::fail(TSR_BAD_SHAPE, "the sink's chains are finished");
    if (s->closed) return Status::fail(TSR_BAD_SHAPE, "AfterFinalBlock: a block followed the short final block");
    if (count == 0 || count > s->block)
        return Status::fail(TSR_BAD_SHAPE, "a block of 1 to " + std::to_string(s->block) + " whole rows");
    if (!rows || (s->P && !masks)) return Status::fail(TSR_BAD_ARG, "null rows or masks");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t st_ = ctx->stream;
    TSR_HIP(hipMemcpyAsync(s->rows.ptr, rows, 8 * count * s->k, hipMemcpyDefault, st_));
    if (s->P) TSR_HIP(hipMemcpyAsync(s->masks.ptr, masks, 8 * count * s->P, hipMemcpyDefault, st_));
    Encoder *enc = nullptr;
    Status st;
    if ((st = get_encoder(ctx, s->log_k, s->log_r, s->P, &enc))) return st;
    if ((st = ctx->scratch.reserve(ctx, enc->scratch_elems(count)))) return st;
    if ((st = enc->encode(s->rows.ptr, s->P ? s->masks.ptr : nullptr, count, s->codewords.ptr,
                          ctx->scratch.ptr, st_)))
        return st;
    if ((st = chains_absorb(ctx, s->codewords.ptr, s->n, (uint32_t)count, s->chains.ptr, st_))) return st;
    s->closed = count < s->block;
    return Status::ok();
}

static Status sink_finish(tsr_block_sink *s, uint8_t *digests) {
    if (s->finished) return Status::fail(TSR_BAD_SHAPE, "the sink's chains are finished");
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    TSR_HIP(hipMemcpyAsync(digests, s->chains.ptr, 32 * s->n, hipMemcpyDefault, s->ctx->stream));
    TSR_HIP(hipStreamSynchronize(s->ctx->stream));
    s->finished = true;
    return Status::ok();
}

static void sink_free(tsr_block_sink *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    zero(s->rows, s->ctx->stream);
    zero(s->masks, s->ctx->stream);
    zero(s->codewords, s->ctx->stream);
    (void)hipStreamSynchronize(s->ctx->stream);
    delete s;
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_block_sink_create(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                                 uint64_t mask_coeffs, uint64_t block, const uint8_t commitment_id[32],
                                 tsr_block_sink **out) {
    if (!ctx || !commitment_id || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    return sink_create((Ctx *)ctx, log_k, log2_inv_rate, mask_coeffs, block, commitment_id, out).code;
}

tsr_status tsr_block_sink_absorb(tsr_block_sink *s, const uint64_t *rows, uint64_t count,
                                 const uint64_t *masks) {
    if (!s) return Status::fail(TSR_BAD_ARG, "null sink").code;
    Status st = sink_absorb(s, rows, masks, count);
    if (!st) {
        hipError_t e = hipStreamSynchronize(s->ctx->stream);
        if (e != hipSuccess) st = hip_fail(e, "block sink");
    }
    return st.code;
}

tsr_status tsr_block_sink_finish(tsr_block_sink *s, uint8_t *digests) {
    if (!s || !digests) return Status::fail(TSR_BAD_ARG, "null argument").code;
    return sink_finish(s, digests).code;
}

void tsr_block_sink_free(tsr_block_sink *s) { sink_free(s); }

tsr_status tsr_matrix_sink_create(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                                  uint64_t block, const uint8_t commitment_id[32], uint64_t columns,
                                  const uint64_t *queries, uint64_t q, tsr_block_sink **out) {
    if (!ctx || !commitment_id || !out || (q && !queries))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    if ((uint64_t)log_k + log2_inv_rate > 32 || log2_inv_rate == 0 ||
        columns != (1ull << (log_k + log2_inv_rate)) || block == 0 || block > 124)
        return Status::fail(TSR_DECLINED, "the device encodes columns = 2^(log_k + r) codewords in "
                                          "blocks of 1 to 124 rows; the host encodes this matrix").code;
    for (uint64_t i = 0; i < q; i++) {
        if (queries[i] >= columns) return Status::fail(TSR_BAD_SHAPE, "QueryRange: a query beyond the columns").code;
        if (i && queries[i] <= queries[i - 1])
            return Status::fail(TSR_BAD_SHAPE, "queries must be distinct and increasing").code;
    }
    tsr_block_sink *s = nullptr;
    Status st = sink_create((Ctx *)ctx, log_k, log2_inv_rate, 0, block, commitment_id, &s);
    if (st) return st.code;
    s->q = q;
    if ((st = s->queries.alloc(s->ctx, std::max<uint64_t>(1, q))) ||
        (st = s->out.alloc(s->ctx, std::max<uint64_t>(1, block * q)))) {
        sink_free(s);
        return st.code;
    }
    if (q) {
        hipError_t e = hipMemcpyAsync(s->queries.ptr, queries, 8 * q, hipMemcpyHostToDevice, s->ctx->stream);
        if (e != hipSuccess) {
            st = hip_fail(e, "queries");
            sink_free(s);
            return st.code;
        }
    }
    *out = s;
    return TSR_OK;
}

tsr_status tsr_matrix_sink_absorb(tsr_block_sink *s, const uint64_t *rows, uint64_t count,
                                  uint64_t *symbols_out) {
    if (!s) return Status::fail(TSR_BAD_ARG, "null sink").code;
    if (s->q && !symbols_out) return Status::fail(TSR_BAD_ARG, "null output").code;
    if (s->P) return Status::fail(TSR_BAD_ARG, "a block sink with masks is not a matrix sink").code;
    Status st = sink_absorb(s, rows, nullptr, count);
    if (st) return st.code;
    hipStream_t stream = s->ctx->stream;
    if (s->q) {
        const uint64_t total = count * s->q;
        hipLaunchKernelGGL(k_gather_row_symbols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                           s->codewords.ptr, s->n, count, s->queries.ptr, s->q, s->out.ptr);
        if ((st = check_launch("matrix sink symbols"))) return st.code;
        hipError_t e = hipMemcpyAsync(symbols_out, s->out.ptr, 8 * total, hipMemcpyDefault, stream);
        if (e != hipSuccess) return hip_fail(e, "symbols").code;
    }
    hipError_t e = hipStreamSynchronize(stream);
    if (e != hipSuccess) return hip_fail(e, "matrix sink").code;
    return TSR_OK;
}

tsr_status tsr_matrix_sink_finish(tsr_block_sink *s, uint8_t *digests) {
    return tsr_block_sink_finish(s, digests);
}

void tsr_matrix_sink_free(tsr_block_sink *s) { sink_free(s); }

}  // extern "C"
