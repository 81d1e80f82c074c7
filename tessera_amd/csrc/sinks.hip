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
    TSR_HIP(tsr::stream_sync(s->ctx->stream));
    s->finished = true;
    return Status::ok();
}

static void sink_free(tsr_block_sink *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    zero(s->rows, s->ctx->stream);
    zero(s->masks, s->ctx->stream);
    zero(s->codewords, s->ctx->stream);
    (void)tsr::stream_sync(s->ctx->stream);
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
        hipError_t e = tsr::stream_sync(s->ctx->stream);
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
    hipError_t e = tsr::stream_sync(stream);
    if (e != hipSuccess) return hip_fail(e, "matrix sink").code;
    return TSR_OK;
}

tsr_status tsr_matrix_sink_finish(tsr_block_sink *s, uint8_t *digests) {
    return tsr_block_sink_finish(s, digests);
}

void tsr_matrix_sink_free(tsr_block_sink *s) { sink_free(s); }

}  // extern "C"

// ---- ColumnSink (crates/tessera-pcs/src/streaming/opening.rs:845-890): P5's column answers from
// re-encoded blocks of rows, and P4's combined rows summed on the device as the blocks pass.
struct tsr_column_sink {
    tsr::Ctx *ctx = nullptr;
    uint32_t log_k = 0, log_r = 0;
    uint64_t P = 0, block = 0, k = 0, n = 0, m = 0;
    uint32_t groups = 0;  // of the combined rows being summed (0: none yet)
    tsr::DevBuf<uint64_t> rows, masks, codewords, columns, out;
    tsr::DevBuf<gl::fp2> coef, sums, part;  // a block's coefficients, the running sums, a block's sums
};

namespace tsr {

__global__ void k_add_fp2(gl::fp2 *__restrict__ acc, const gl::fp2 *__restrict__ x, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) acc[i] = gl::add(acc[i], x[i]);
}

static void column_sink_clear(tsr_column_sink *s) {
    hipStream_t st = s->ctx->stream;
    zero(s->rows, st);
    zero(s->masks, st);
    zero(s->codewords, st);
    zero(s->out, st);
    if (s->sums.ptr) zero_words((uint64_t *)s->sums.ptr, 2 * s->sums.len, st);
    if (s->part.ptr) zero_words((uint64_t *)s->part.ptr, 2 * s->part.len, st);
}

// One block: optionally the combined rows' sums (before the encode, as the reference's device
// sink orders them), the encode, the symbols at the sink's columns (row by row) into `symbols`.
static Status column_sink_run(tsr_column_sink *s, const uint64_t *rows, const uint64_t *masks,
                              uint64_t count, const uint64_t *coefficients, uint32_t groups,
                              uint64_t *symbols) {
    Ctx *ctx = s->ctx;
    if (count == 0 || count > s->block)
        return Status::fail(TSR_BAD_SHAPE, "a block of 1 to " + std::to_string(s->block) + " whole rows");
    if (!rows || (s->P && !masks) || !symbols || (groups && !coefficients))
        return Status::fail(TSR_BAD_ARG, "null argument");
    if (groups && s->groups && groups != s->groups)
        return Status::fail(TSR_BAD_SHAPE, "one group count per opening");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    hipStream_t st_ = ctx->stream;
    Status st;
    TSR_HIP(hipMemcpyAsync(s->rows.ptr, rows, 8 * count * s->k, hipMemcpyDefault, st_));
    if (s->P) TSR_HIP(hipMemcpyAsync(s->masks.ptr, masks, 8 * count * s->P, hipMemcpyDefault, st_));
    if (groups) {
        for (uint64_t i = 0; i < 2ull * groups * count; i++)
            if (coefficients[i] >= gl::P) return Status::fail(TSR_BAD_ARG, "a coefficient is not canonical");
        if (!s->groups) {
            if ((st = s->sums.alloc(ctx, (uint64_t)groups * s->k)) || (st = s->part.alloc(ctx, (uint64_t)groups * s->k)) ||
                (st = s->coef.alloc(ctx, (uint64_t)groups * s->block)))
                return st;
            zero_words((uint64_t *)s->sums.ptr, 2 * s->sums.len, st_);
            s->groups = groups;
        }
        // coefficients[g][r] of the block's rows, group-major: combine_rows reads coef[g * N + r]
        TSR_HIP(hipMemcpyAsync(s->coef.ptr, coefficients, 16ull * groups * count, hipMemcpyDefault, st_));
        if ((st = combine_rows(ctx, s->coef.ptr, groups, s->rows.ptr, count, s->k, s->part.ptr, st_))) return st;
        const uint64_t len = (uint64_t)groups * s->k;
        hipLaunchKernelGGL(k_add_fp2, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st_, s->sums.ptr,
                           s->part.ptr, len);
    }
    Encoder *enc = nullptr;
    if ((st = get_encoder(ctx, s->log_k, s->log_r, s->P, &enc))) return st;
    if ((st = ctx->scratch.reserve(ctx, enc->scratch_elems(count)))) return st;
    if ((st = enc->encode(s->rows.ptr, s->P ? s->masks.ptr : nullptr, count, s->codewords.ptr,
                          ctx->scratch.ptr, st_)))
        return st;
    const uint64_t total = count * s->m;
    hipLaunchKernelGGL(k_gather_row_symbols, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st_,
                       s->codewords.ptr, s->n, count, s->columns.ptr, s->m, s->out.ptr);
    if ((st = check_launch("column sink"))) return st;
    TSR_HIP(hipMemcpyAsync(symbols, s->out.ptr, 8 * total, hipMemcpyDefault, st_));
    TSR_HIP(tsr::stream_sync(st_));
    return Status::ok();
}

}  // namespace tsr

extern "C" {

tsr_status tsr_column_sink_create(tsr_ctx *ctx_, uint32_t log_k, uint32_t log2_inv_rate,
                                  uint64_t mask_coeffs, uint64_t block, const uint64_t *columns,
                                  uint64_t m, tsr_column_sink **out) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !out || (m && !columns)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if ((uint64_t)log_k + log2_inv_rate > 32 || log2_inv_rate == 0)
        return Status::fail(TSR_BAD_SHAPE, "TooLong or rate: n = 2^(log_k + r) with 1 <= r, at most 2^32").code;
    if (mask_coeffs > (1ull << log_k)) return Status::fail(TSR_BAD_SHAPE, "MaskTooLarge: mask_coeffs exceeds k").code;
    if (block == 0 || block > 65535) return Status::fail(TSR_BAD_SHAPE, "a block of 1 to 65535 rows").code;
    if (m == 0) return Status::fail(TSR_BAD_SHAPE, "NoQueries").code;
    const uint64_t n = 1ull << (log_k + log2_inv_rate);
    for (uint64_t i = 0; i < m; i++) {
        if (columns[i] >= n) return Status::fail(TSR_BAD_SHAPE, "QueryRange: a column beyond the codeword").code;
        if (i && columns[i] <= columns[i - 1])
            return Status::fail(TSR_BAD_SHAPE, "columns must be distinct and increasing").code;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    std::unique_ptr<tsr_column_sink> s(new tsr_column_sink());
    s->ctx = ctx;
    s->log_k = log_k;
    s->log_r = log2_inv_rate;
    s->P = mask_coeffs;
    s->block = block;
    s->k = 1ull << log_k;
    s->n = n;
    s->m = m;
    Encoder *enc = nullptr;
    Status st;
    if ((st = get_encoder(ctx, log_k, log2_inv_rate, mask_coeffs, &enc))) return st.code;
    if ((st = s->rows.alloc(ctx, block * s->k)) || (st = s->masks.alloc(ctx, std::max<uint64_t>(1, block * mask_coeffs))) ||
        (st = s->codewords.alloc(ctx, block * n)) || (st = s->columns.alloc(ctx, m)) ||
        (st = s->out.alloc(ctx, block * m)))
        return st.code;
    e = hipMemcpyAsync(s->columns.ptr, columns, 8 * m, hipMemcpyDefault, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "columns").code;
    e = tsr::stream_sync(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "columns").code;
    *out = s.release();
    return TSR_OK;
}

tsr_status tsr_column_sink_select(tsr_column_sink *s, const uint64_t *rows, const uint64_t *masks,
                                  uint64_t count, uint64_t *symbols_out) {
    if (!s) return Status::fail(TSR_BAD_ARG, "null sink").code;
    return column_sink_run(s, rows, masks, count, nullptr, 0, symbols_out).code;
}

int tsr_column_sink_combines(const tsr_column_sink *s) { return s ? 1 : 0; }

tsr_status tsr_column_sink_select_combining(tsr_column_sink *s, const uint64_t *rows,
                                            const uint64_t *masks, uint64_t count,
                                            const uint64_t *coefficients, uint32_t groups,
                                            uint64_t *symbols_out) {
    if (!s) return Status::fail(TSR_BAD_ARG, "null sink").code;
    if (groups == 0) return column_sink_run(s, rows, masks, count, nullptr, 0, symbols_out).code;
    return column_sink_run(s, rows, masks, count, coefficients, groups, symbols_out).code;
}

tsr_status tsr_column_sink_combined(tsr_column_sink *s, uint64_t *out, uint32_t groups) {
    if (!s) return Status::fail(TSR_BAD_ARG, "null sink").code;
    if (!s->groups) return TSR_OK;  // nothing was combined
    if (groups != s->groups || !out) return Status::fail(TSR_BAD_SHAPE, "one combined row per group").code;
    hipError_t e = hipSetDevice(s->ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    const uint64_t words = 2ull * groups * s->k;
    std::vector<uint64_t> h(words);
    e = hipMemcpyAsync(h.data(), s->sums.ptr, 8 * words, hipMemcpyDeviceToHost, s->ctx->stream);
    if (e == hipSuccess) e = tsr::stream_sync(s->ctx->stream);
    // the sums are secret-class: cleared on the device whatever happened, then on the host
    zero_words((uint64_t *)s->sums.ptr, words, s->ctx->stream);
    (void)tsr::stream_sync(s->ctx->stream);
    s->groups = 0;
    if (e != hipSuccess) {
        memset(h.data(), 0, 8 * words);
        return hip_fail(e, "combined rows").code;
    }
    for (uint64_t i = 0; i < words; i += 2) {
        gl::fp2 a{out[i], out[i + 1]}, b{h[i], h[i + 1]};
        gl::fp2 c = gl::add(a, b);
        out[i] = c.c0;
        out[i + 1] = c.c1;
    }
    volatile uint64_t *vh = h.data();
    for (uint64_t i = 0; i < words; i++) vh[i] = 0;
    return TSR_OK;
}

void tsr_column_sink_free(tsr_column_sink *s) {
    if (!s) return;
    (void)hipSetDevice(s->ctx->device);
    column_sink_clear(s);
    (void)tsr::stream_sync(s->ctx->stream);
    delete s;
}

}  // extern "C"
