/*
 * tessera_mi355x — the C ABI of the MI355X (gfx950) prover back end for Tessera's
 * data-parallel core: Goldilocks Reed-Solomon row encoding (NTT/LDE), BLAKE3 column chains and
 * Merkle commitment, and the sumcheck / LogUp evaluate-and-fold rounds.
 *
 * Plain C: pointers, explicit sizes and status codes; no HIP or torch types. Field elements are
 * canonical Goldilocks u64 (little-endian, in [0, p), p = 2^64 - 2^32 + 1); an Fp2 element is
 * two u64 (c0, c1) with X^2 = 7; a digest is 32 bytes. Unless a function says otherwise, element
 * pointers may be host memory or device memory of the context's GPU; the library detects which.
 *
 * Every entry point checks its shapes before it launches anything or moves a transcript, and
 * reports failure through a status code, with the reason in tsr_last_error() (thread-local).
 * TSR_DECLINED means the caller should compute the result on the CPU instead; nothing was
 * changed. A context is used by one thread at a time; separate contexts may run concurrently.
 *
 * Each declaration cites the reference interface it replaces (AbdelStark/tessera, file:line).
 */
#ifndef TESSERA_MI355X_H
#define TESSERA_MI355X_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    TSR_OK = 0,
    TSR_DECLINED = 1,  /* caller falls back to the CPU path; no side effects */
    TSR_BAD_SHAPE = 2, /* a length or parameter check failed; no side effects */
    TSR_DEVICE = 3,    /* a HIP runtime error */
    TSR_OOM = 4,       /* device memory exhausted or above the context's budget */
    TSR_BAD_ARG = 5,   /* null pointer or unusable argument */
    TSR_CANCELLED = 6  /* a served request ended by tsr_server_cancel */
} tsr_status;

typedef struct tsr_ctx tsr_ctx;
typedef struct tsr_commitment tsr_commitment;
typedef struct tsr_chains tsr_chains;
typedef struct tsr_transcript tsr_transcript;

/* The reason for the last non-OK status on this thread (empty when none). */
const char *tsr_last_error(void);
/* The library's version string. */
const char *tsr_version(void);

/* ---- context: one GPU, one stream (replaces Device::new, crates/tessera-gpu/src/device.rs:125-146) */
tsr_status tsr_ctx_create(int device, tsr_ctx **out);
void tsr_ctx_destroy(tsr_ctx *ctx);
/* Wait for all work of this context (Device::synchronize, device.rs:220). */
tsr_status tsr_ctx_synchronize(tsr_ctx *ctx);
/* Cap the context's device memory (bytes, 0 = no cap); larger requests return TSR_OOM. */
tsr_status tsr_ctx_set_budget(tsr_ctx *ctx, uint64_t bytes);
/* Device memory currently held by the context (bytes). */
uint64_t tsr_ctx_memory(const tsr_ctx *ctx);
/* Device buffers owned by the context, for callers that keep inputs resident in HBM
 * (Device::upload / zeroed / download, device.rs:149-247). */
tsr_status tsr_alloc(tsr_ctx *ctx, uint64_t bytes, void **out);
tsr_status tsr_free(tsr_ctx *ctx, void *ptr);
tsr_status tsr_copy(tsr_ctx *ctx, void *dst, const void *src, uint64_t bytes);
/* Fill dst[i] = SplitMix64 stream value i, words >= p rejected, exactly uniform (the generator of
 * crates/tessera-gpu/tests/commit.rs:19-30 with rejection). Device-side bench input generator. */
tsr_status tsr_fill_uniform(tsr_ctx *ctx, uint64_t *dst, uint64_t count, uint64_t seed);

/* ---- NTT plans (NttPlan::{new, size, ntt, intt, coset_ntt, coset_intt},
 *      crates/tessera-pcs/src/ntt.rs:397-516) ---- */
typedef struct tsr_ntt_plan tsr_ntt_plan;
/* NttPlan::new(size): the forward and inverse twiddles of size n on the device. NotPowerOfTwo
 * (size 0 or not a power of two), TooLong (above 2^32) (BAD_SHAPE). */
tsr_status tsr_ntt_plan_create(tsr_ctx *ctx, uint64_t size, tsr_ntt_plan **out);
uint64_t tsr_ntt_plan_size(const tsr_ntt_plan *plan);
/* In place, natural order on both sides, over len / n rows of n canonical values (host or device
 * memory; the reference transforms one row, len = n). ntt: values[j] = f(w_n^j) for
 * f = sum_i values[i] X^i; intt its inverse; coset_ntt: values[j] = f(shift w_n^j); coset_intt
 * its inverse. Errors before any value moves: TransformLength (len not a non-zero multiple of
 * n), then ZeroShift (a zero shift) (BAD_SHAPE); a shift >= p is BAD_ARG. */
tsr_status tsr_ntt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len);
tsr_status tsr_intt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len);
tsr_status tsr_coset_ntt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len, uint64_t shift);
tsr_status tsr_coset_intt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len, uint64_t shift);
void tsr_ntt_plan_free(tsr_ntt_plan *plan);

/* ---- row encoding (RowEncoder, crates/tessera-pcs/src/encoding.rs:126-290;
 *      Device::encoder / encode_rows, crates/tessera-gpu/src/commit.rs:178,277) ---- */
/* Encode `count` rows of k = 2^log_k values (row-major) into codewords of n = 2^(log_k+r)
 * symbols: symbol j = g(7 * w_n^j), g = f + (X^k - 1) m, with `mask_coeffs` = P mask
 * coefficients per row in `masks` (count x P, m_0 first; NULL when P = 0).
 * Errors (in this order): r = 0 or log_k + r > 32 (BAD_SHAPE), P > k (BAD_SHAPE). */
tsr_status tsr_encode_rows(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                           uint64_t mask_coeffs, const uint64_t *rows, const uint64_t *masks,
                           uint64_t count, uint64_t *codewords);

/* ---- column chains (ColumnChains, crates/tessera-hash/src/columns.rs:72-148;
 *      Device::column_chains / absorb_block / download_chains, commit.rs:418,460,598) ---- */
/* d_j[0] = H(column-init, commitment_id || u64LE(j)) for `columns` columns, blocks of `block` rows. */
tsr_status tsr_chains_create(tsr_ctx *ctx, const uint8_t commitment_id[32], uint64_t columns,
                             uint64_t block, tsr_chains **out);
/* Absorb `count` rows (row-major, `columns` symbols each): 1 <= count <= block; a block shorter
 * than `block` must be the last (ColumnError::{BlockSize, AfterFinalBlock}). */
tsr_status tsr_chains_absorb(tsr_chains *chains, const uint64_t *rows, uint64_t count);
/* The chains' current digests, 32 bytes per column (ColumnChains::finish). */
tsr_status tsr_chains_read(tsr_chains *chains, uint8_t *out);
void tsr_chains_free(tsr_chains *chains);

/* ---- device sinks (BlockSink, crates/tessera-pcs/src/streaming.rs:24-38; MatrixSink and
 *      WeightEncoding::matrix, crates/tessera-pcs/src/weights/regenerated.rs:24-62; the reference's
 *      implementations AdviceSink / WeightSink, crates/tessera-prover/src/gpu.rs:161-196,432-667).
 *      A sink owns its rows, masks and codewords on the device and zeroes them when freed. ---- */
typedef struct tsr_block_sink tsr_block_sink;
typedef struct tsr_block_sink tsr_matrix_sink;
/* A block sink for rows of k = 2^log_k values with P = mask_coeffs masks per row, at rate
 * 2^-log2_inv_rate, absorbing blocks of up to `block` (1..124) rows into the chains of the
 * n = 2^(log_k + r) codeword columns of `commitment_id`. */
tsr_status tsr_block_sink_create(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                                 uint64_t mask_coeffs, uint64_t block,
                                 const uint8_t commitment_id[32], tsr_block_sink **out);
/* BlockSink::absorb: `count` whole rows (1..block; a shorter block must be the last) and their
 * masks (count x P, m_0 first) are encoded and absorbed into every column chain. */
tsr_status tsr_block_sink_absorb(tsr_block_sink *sink, const uint64_t *rows, uint64_t count,
                                 const uint64_t *masks);
/* BlockSink::finish: the n column chain values, 32 bytes each, in column order. */
tsr_status tsr_block_sink_finish(tsr_block_sink *sink, uint8_t *digests);
void tsr_block_sink_free(tsr_block_sink *sink);
/* WeightEncoding::matrix: a sink for unmasked rows of 2^log_k values at rate 2^-r into the chains of
 * `columns` columns, answering the q `queries` (distinct, increasing). DECLINED (the host encodes
 * the matrix) unless columns = 2^(log_k + r) and block is 1..124. */
tsr_status tsr_matrix_sink_create(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                                  uint64_t block, const uint8_t commitment_id[32],
                                  uint64_t columns, const uint64_t *queries, uint64_t q,
                                  tsr_matrix_sink **out);
/* MatrixSink::absorb: encodes and absorbs `count` whole rows and writes each row's symbols at the
 * queried columns, row by row (count x q values). */
tsr_status tsr_matrix_sink_absorb(tsr_matrix_sink *sink, const uint64_t *rows, uint64_t count,
                                  uint64_t *symbols_out);
/* MatrixSink::finish: the column chain values, 32 bytes each. */
tsr_status tsr_matrix_sink_finish(tsr_matrix_sink *sink, uint8_t *digests);
void tsr_matrix_sink_free(tsr_matrix_sink *sink);

/* ---- ColumnSink (crates/tessera-pcs/src/streaming/opening.rs:845-890; the reference's device
 *      sink crates/tessera-prover/src/gpu.rs:195-360): P5's answers re-encoded block by block,
 *      P4's combined rows summed on the device as the blocks pass ---- */
typedef struct tsr_column_sink tsr_column_sink;
/* A sink for codewords of RowEncoder(2^log_k, r, mask_coeffs) in blocks of up to `block` rows,
 * answering the m distinct increasing `columns`. NoQueries (m = 0), QueryRange (a column >= n),
 * TooLong / MaskTooLarge (BAD_SHAPE). */
tsr_status tsr_column_sink_create(tsr_ctx *ctx, uint32_t log_k, uint32_t log2_inv_rate,
                                  uint64_t mask_coeffs, uint64_t block, const uint64_t *columns,
                                  uint64_t m, tsr_column_sink **out);
/* ColumnSink::select: encodes `count` whole rows (count x k) with their masks (count x P, m_0
 * first) and writes each row's symbols at the sink's columns, row by row (count x m). */
tsr_status tsr_column_sink_select(tsr_column_sink *s, const uint64_t *rows, const uint64_t *masks,
                                  uint64_t count, uint64_t *symbols_out);
/* ColumnSink::combines: 1 (the sums of select_combining stay on the device). */
int tsr_column_sink_combines(const tsr_column_sink *s);
/* ColumnSink::select_combining: as select, and adds sum_r coefficients[g][r] rows[r] to the
 * sink's g-th sum (coefficients: groups x count Fp2, group-major; the same group count for
 * every block of an opening). */
tsr_status tsr_column_sink_select_combining(tsr_column_sink *s, const uint64_t *rows,
                                            const uint64_t *masks, uint64_t count,
                                            const uint64_t *coefficients, uint32_t groups,
                                            uint64_t *symbols_out);
/* ColumnSink::combined: adds the g-th sum into out[g] (groups x k Fp2, host memory) and clears
 * the sums; OK with nothing added when no block was combined. The host copy is cleared too. */
tsr_status tsr_column_sink_combined(tsr_column_sink *s, uint64_t *out, uint32_t groups);
/* Rows, masks, codewords and sums are cleared before the memory is released. */
void tsr_column_sink_free(tsr_column_sink *s);

/* ---- Merkle (MerkleTree::new / root, crates/tessera-hash/src/merkle.rs:93-119) ---- */
/* root of the binary BLAKE3 tree over n = 2^m leaves (32 bytes each). */
tsr_status tsr_merkle_root(tsr_ctx *ctx, const uint8_t *leaves, uint64_t n, uint8_t root[32]);
/* tsr_merkle_root with no allocation of any kind. No counterpart in the reference: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc). The tree is built in a fixed array of the library's
 * own (4 MiB per GPU, held until the process ends), so n is at most 2^16 leaves (DECLINED above
 * that); leaves and root may be host or device memory. Every call on one GPU shares that array:
 * calls that overlap on one GPU race. */
tsr_status tsr_merkle_root_noalloc(tsr_ctx *ctx, const uint8_t *leaves, uint64_t n,
                                   uint8_t root[32]);

/* ---- the unmasked reference commitment (ReferenceCommitment::commit_unmasked / root /
 *      leaves / answer, crates/tessera-pcs/src/commitment.rs:343-598) ---- */
typedef struct {
    uint32_t log2_k;         /* k = 2^log2_k values per row */
    uint32_t log2_inv_rate;  /* r >= 1: n = 2^r k symbols per codeword */
    uint64_t mask_coeffs;    /* P: must be 0 for commit_unmasked */
    uint64_t row_block;      /* B >= 1 rows per chain step */
    uint8_t commitment_id[32];
} tsr_commitment_params;

/* Encode `count` rows, chain each column over them in blocks of B (leaf = last digest; the
 * leaves of a commitment of no row are d_j[0]), build the Merkle tree. Codewords, leaves and the
 * tree stay resident on the device for answer(). Errors: r = 0 or log_k + r > 32, P != 0,
 * B = 0 (BAD_SHAPE), before any work. */
tsr_status tsr_commit_unmasked(tsr_ctx *ctx, const tsr_commitment_params *params,
                               const uint64_t *rows, uint64_t count, tsr_commitment **out);
/* ReferenceCommitment::commit (commitment.rs:318-326): the masked advice commitment of a layout
 * of `segments` segments holding segment_rows[s] non-zero rows each (RowLayout, commitment.rs:
 * 186-262; the layout's row indices are the caller's to check, only the counts enter). `rows`
 * holds the N = sum(segment_rows) rows in layout order (N x k) and `masks` their P = mask_coeffs
 * coefficients each, m_0 first, as drawn from the MaskSource (N x P). Every segment's columns
 * are chained in blocks of B over that segment's rows (an empty segment keeps d_j[0]), leaf_j =
 * H(column-merge, d_j^(1) || ... || d_j^(S)), and the tree is built over the n leaves.
 * Errors, in this order and before any work: r = 0 or log_k + r > 32, P = 0 (NoMask), P > k
 * (MaskTooLarge), no segment (NoSegments), B = 0 (BAD_SHAPE); B > 124 (DECLINED). */
tsr_status tsr_commit(tsr_ctx *ctx, const tsr_commitment_params *params, const uint64_t *segment_rows,
                      uint32_t segments, const uint64_t *rows, const uint64_t *masks,
                      tsr_commitment **out);
/* ReferenceCommitment::from_encoded (commitment.rs:418-450): steps 2 to 4 from N x n codewords
 * already encoded (host or device; copied and kept resident): merged != 0 gives the masked
 * commitment's column-merge leaves over the segments, merged = 0 the chained leaves of one
 * segment (then P must be 0). Same errors as tsr_commit / tsr_commit_unmasked. */
tsr_status tsr_commit_from_encoded(tsr_ctx *ctx, const tsr_commitment_params *params,
                                   const uint64_t *segment_rows, uint32_t segments, int merged,
                                   const uint64_t *codewords, tsr_commitment **out);
/* RowLayout::row_coefficients (crates/tessera-pcs/src/opening.rs:116-158): for each claim group g
 * (group_sizes[g] claims, listed group after group, each a row point of row_vars Fp2 in
 * row_points and a gamma), c_g(row_i) = Sum_claims gamma eq(row_point, row_i) over the N layout
 * rows `rows` (row indices, layout order), then the proximity row eq(xi_point, row_i).
 * out: (groups + 1) x N Fp2 (host or device). BAD_SHAPE for row_vars > 64. */
tsr_status tsr_row_coefficients(tsr_ctx *ctx, uint32_t row_vars, const uint64_t *rows, uint64_t N,
                                const uint64_t *row_points, const uint64_t *gammas,
                                const uint32_t *group_sizes, uint32_t groups, const uint64_t *xi_point,
                                uint64_t *out);
/* ReferenceCommitment::open's combine (commitment.rs:489-524, opening.rs:241-330) over a masked
 * commitment's resident rows and masks: for each of `count` coefficient rows (count x N Fp2,
 * e.g. tsr_row_coefficients' output), u = Sum_row c(row) A_row (k Fp2) and M = Sum_row c(row)
 * m_row (P Fp2). BAD_SHAPE for an unmasked commitment, which keeps no rows. */
tsr_status tsr_commitment_open(const tsr_commitment *c, const uint64_t *coefficients, uint32_t count,
                               uint64_t *u_out, uint64_t *mask_out);
tsr_status tsr_commitment_root(const tsr_commitment *c, uint8_t root[32]);
/* n leaves, 32 bytes each, column order. */
tsr_status tsr_commitment_leaves(const tsr_commitment *c, uint8_t *out);
/* The device address of the codewords (count x n u64, row-major) and their shape. */
tsr_status tsr_commitment_codewords(const tsr_commitment *c, const uint64_t **dev_codewords,
                                    uint64_t *count, uint64_t *n);
/* answer(queries): the sorted distinct indices, the column of each (one symbol per row) and the
 * canonical multiproof. Output capacities are checked first: `distinct_cap` indices (and
 * distinct_cap x count symbols), `siblings_cap` digests; the sizes needed are always written.
 * Errors: q = 0 (NoQueries, BAD_SHAPE), a query >= n (QueryRange, BAD_SHAPE), capacity short
 * (BAD_SHAPE). */
tsr_status tsr_commitment_answer(const tsr_commitment *c, const uint64_t *queries, uint64_t q,
                                 uint64_t *distinct_out, uint64_t distinct_cap,
                                 uint64_t *columns_out, uint8_t *siblings_out,
                                 uint64_t siblings_cap, uint64_t *n_distinct,
                                 uint64_t *n_siblings);
void tsr_commitment_free(tsr_commitment *c);

/* ---- batched commitments: many unmasked matrices in one planned launch sequence. No single
 *      reference function: each matrix's root is ReferenceCommitment::commit_unmasked's
 *      (crates/tessera-pcs/src/commitment.rs:343-446); the batch replaces calling it per matrix. */
typedef struct tsr_batch tsr_batch;
typedef struct {
    uint32_t log2_k, log2_inv_rate; /* k = 2^log2_k, n = 2^(log2_k + r) */
    uint64_t row_block;             /* B, 1..124 */
    uint8_t commitment_id[32];
    const uint64_t *rows;           /* count x k values, host or device */
    uint64_t count;                 /* rows of the matrix (may be 0) */
} tsr_commit_desc;
enum {
    TSR_BATCH_AUTOTUNE = 1, /* time each shape group's plans (LDS tile 2^13 or 2^12, strided passes
                               of at most the default or fewer stage bits) on its rows; the choice
                               is kept in the context per (log2 k, r, log2 rows) for later batches */
    TSR_BATCH_GRAPH = 2,    /* capture the launch sequence once into a hipGraph and replay it */
    TSR_BATCH_SERIAL = 4,   /* every shape group on the context's stream, one after another (by
                               default each group's encode is a parallel branch) */
    TSR_BATCH_PLAN = 8      /* every group uses the plan in bits 8..15 (tile_log, 12 or 13) and
                               16..23 (most stage bits per strided pass, 0: the default) of flags */
};
/* Plans a batch: device buffers for every matrix's codewords and tree, the shape groups, the
 * block tables. Host rows are uploaded once; device rows are used in place on every run. */
tsr_status tsr_commit_batch_create(tsr_ctx *ctx, const tsr_commit_desc *descs, uint32_t count,
                                   uint32_t flags, tsr_batch **out);
/* Encodes, chains and Merkle-commits every matrix (synchronously). */
tsr_status tsr_commit_batch_run(tsr_batch *b);
/* The count roots, 32 bytes each, in descriptor order, of the last run. */
tsr_status tsr_commit_batch_roots(const tsr_batch *b, uint8_t *roots);
/* The plan per shape group, five words each, `cap` groups at most: log2_k, r, tile_log, the most
 * stage bits of a strided pass (0: the default), and the tuned plan's encode time in
 * microseconds (0 when not timed by this batch: no autotune, or a plan from the context's cache). */
tsr_status tsr_commit_batch_tuning(const tsr_batch *b, uint32_t *plans, uint32_t cap,
                                   uint32_t *n_groups);
void tsr_commit_batch_free(tsr_batch *b);

/* ---- served proving with continuous batching. No reference counterpart: the reference proves
 *      one statement per call and has no serving loop. Each request: a synthetic int8
 *      transformer (4 layers, d_model 256, 4 heads, d_ff 1024, vocab 4096) decodes greedily and
 *      writes its witness; the witness is committed (rows of 2^13, rate 1/4, B = 32, unmasked);
 *      a degree-3 sumcheck Sum eq(r,x) A(x) B(x) over two witness tables follows on the request's
 *      own transcript. Admitted requests decode together (one launch advances each by up to
 *      `chunk` positions; requests join and leave between launches) and are proved on prover
 *      streams while decoding goes on: a prover commits the decoded requests waiting for it (up
 *      to `group`) in one launch set, then proves each on its own transcript. ---- */
#define TSR_MAX_TOKENS 512
typedef struct tsr_server tsr_server;
typedef struct {
    uint32_t slots;       /* requests in flight at once, decoding or proving (1..1024) */
    uint32_t provers;     /* prover threads, one stream each (0 = 4) */
    uint32_t chunk;       /* positions a decoding request advances per decode launch (0 = 32) */
    uint32_t group;       /* most decoded requests a prover commits in one launch set (0 = 1:
                           * each request on its own) */
    uint64_t hbm_budget;  /* admission budget in bytes for the requests in flight (0 = none) */
    uint64_t model_seed;  /* weights: SplitMix64 words mod 255, minus 127 */
    uint32_t resident;    /* 1: a proved request's witness codewords and Merkle tree stay in device
                           * memory, holding their bytes of hbm_budget, until tsr_server_release;
                           * tsr_server_open answers column openings from them */
} tsr_server_config;
typedef struct {
    uint64_t id;           /* also the proof's session id */
    uint32_t n_prompt;     /* >= 1 prompt tokens, each < 4096 */
    const uint32_t *prompt;
    uint32_t n_output;     /* tokens generated greedily; n_prompt + n_output <= 512 */
} tsr_request;
typedef struct {
    uint64_t id;
    tsr_status status;
    uint32_t slot;                  /* the prover that proved it */
    uint32_t group;                 /* requests whose witnesses were committed with it (1: alone) */
    uint8_t root[32];               /* the witness commitment */
    uint8_t transcript_state[32];   /* after the sumcheck's final evaluations */
    uint64_t transcript_counter;
    uint64_t final_a[2];            /* A~(point), the first final evaluation */
    uint64_t witness_rows, sumcheck_vars;
    uint32_t n_tokens;
    uint32_t tokens[TSR_MAX_TOKENS]; /* the prompt, then the generated tokens */
    double t_submit, t_admit, t_done; /* seconds on the server's clock */
} tsr_result;
tsr_status tsr_server_create(int device, const tsr_server_config *cfg, tsr_server **out);
/* Queues a request (copied); returns at once. The ids of requests in flight are distinct
 * (BAD_ARG for a repeat). */
tsr_status tsr_server_submit(tsr_server *srv, const tsr_request *req);
/* Cancels a request that is waiting for admission, decoding, or decoded and waiting for a
 * prover: its result arrives with status TSR_CANCELLED, and its slot, budget and device memory
 * are released (a decoding request at the end of the decode launch in progress). DECLINED when
 * the request is being proved, is done or was never submitted. The other requests' results are
 * unchanged. */
tsr_status tsr_server_cancel(tsr_server *srv, uint64_t id);
/* Takes one finished result, waiting up to timeout_ms (DECLINED when none arrived). */
tsr_status tsr_server_wait(tsr_server *srv, tsr_result *out, uint32_t timeout_ms);
/* The most requests ever in flight together, and the footprint in flight now. */
tsr_status tsr_server_stats(tsr_server *srv, uint32_t *max_active, uint64_t *in_flight_bytes);
/* The plain per-request call: proves one request synchronously on the calling thread, on the
 * server's own context (concurrent calls take turns). */
tsr_status tsr_server_prove(tsr_server *srv, const tsr_request *req, tsr_result *out);
/* Opens columns of a resident request's witness commitment: as tsr_commitment_answer (sorted
 * distinct columns, each column's symbols of every witness row, the canonical multiproof), on
 * the server's own context. DECLINED when the request has no resident commitment. */
tsr_status tsr_server_open(tsr_server *srv, uint64_t id, const uint64_t *queries, uint64_t q,
                           uint64_t *distinct_out, uint64_t distinct_cap, uint64_t *columns_out,
                           uint8_t *siblings_out, uint64_t siblings_cap, uint64_t *n_distinct,
                           uint64_t *n_siblings);
/* Frees a resident commitment and gives its bytes back to the budget (DECLINED: none). */
tsr_status tsr_server_release(tsr_server *srv, uint64_t id);
/* The resident commitments and the device bytes they hold. */
tsr_status tsr_server_resident(tsr_server *srv, uint32_t *count, uint64_t *bytes);
/* The admission footprint of a request of `tokens` tokens, in bytes. */
uint64_t tsr_server_footprint(uint32_t tokens);
/* Seconds since the server started (the clock of the results' timestamps). */
double tsr_server_clock(tsr_server *srv);
void tsr_server_destroy(tsr_server *srv);

/* ---- hashing and the Fiat-Shamir transcript (crates/tessera-hash/src/domain.rs:177-181,
 *      crates/tessera-hash/src/transcript.rs:40-290) ---- */
/* H(domain, bytes) = BLAKE3.derive_key("tessera/v1/<name>", bytes); `domain` indexes the
 * reference's 19 domains in declaration order (domain.rs:41-81), e.g. 2 = column-init.
 * The hash and the transcript run on the host; `ctx` may be NULL for them. */
tsr_status tsr_hash(tsr_ctx *ctx, int domain, const uint8_t *data, uint64_t len, uint8_t out[32]);
/* Transcript::new(context): state = H(transcript-init, TCE(SessionContext)). */
tsr_status tsr_transcript_create(tsr_ctx *ctx, const uint8_t *context_tce, uint64_t len,
                                 tsr_transcript **out);
/* A transcript continuing from a state and squeeze counter another party's Transcript reached
 * (the fields the reference's Transcript holds, crates/tessera-hash/src/transcript.rs:92-96): the hand-over of a session
 * between a host prover and this library. */
tsr_status tsr_transcript_from_state(tsr_ctx *ctx, const uint8_t state[32], uint64_t counter,
                                     tsr_transcript **out);
void tsr_transcript_free(tsr_transcript *t);
tsr_status tsr_transcript_state(const tsr_transcript *t, uint8_t state[32], uint64_t *counter);
/* absorb(label, value): `tce` is the value's TCE encoding. Labels are printable ASCII. */
tsr_status tsr_transcript_absorb(tsr_transcript *t, const char *label, const uint8_t *tce,
                                 uint64_t len);
tsr_status tsr_transcript_squeeze(tsr_transcript *t, const char *label, uint8_t out[32]);
tsr_status tsr_transcript_challenge_fp(tsr_transcript *t, const char *label, uint64_t *out);
tsr_status tsr_transcript_challenge_fp2(tsr_transcript *t, const char *label, uint64_t out[2]);
tsr_status tsr_transcript_challenge_indices(tsr_transcript *t, const char *label,
                                            uint32_t log_n, uint32_t count, uint64_t *out);
tsr_status tsr_transcript_grind(tsr_transcript *t, const char *label, uint8_t bits,
                                uint64_t *nonce);
tsr_status tsr_transcript_check_grind(tsr_transcript *t, const char *label, uint8_t bits,
                                      uint64_t nonce, int *valid);

/* ---- the batched product sumcheck (prove_batch, crates/tessera-sumcheck/src/prover.rs:1122-1145;
 *      Table / Instance::new, prover.rs:78-305; SumcheckProof, proof.rs:16-24) ---- */
typedef struct {
    const uint64_t *data; /* 2^vars entries: u64 each (Fp, Table::base) or 2 u64 (Fp2, Table::ext);
                             for ext = TSR_TABLE_EQ the point of eq(point, .): vars Fp2
                             coordinates, coordinate i for bit i of the index */
    uint32_t ext;         /* 1: Fp2 entries; TSR_TABLE_EQ: the table eq_table(point)
                             (crates/tessera-poly/src/eq.rs:23) given by its point instead of its
                             2^vars entries, the same proof either way */
    uint32_t is_public;   /* 1: a public factor, left out of the final evaluations (Table::public) */
} tsr_table;
#define TSR_TABLE_EQ 2u
/* One term of a sum-of-products instance (Term, crates/tessera-sumcheck/src/terms.rs:10-40):
 * weight * the product of the tables at `factors`, positions in the instance's table list,
 * which may repeat (a squared table). */
typedef struct {
    uint64_t weight[2];       /* Fp2 */
    const uint32_t *factors;
    uint32_t nfactors;        /* 1 ..= the batch's degree */
} tsr_term;
typedef struct {
    uint32_t vars;            /* n_i: every table of the instance holds 2^n_i entries */
    uint32_t factors;         /* the number of tables */
    const tsr_table *tables;
    const tsr_term *terms;    /* NULL: the product of every table (Instance::new); else the sum
                                 of the terms (Instance::sum_of_products, prover.rs:300-312) */
    uint32_t nterms;
} tsr_instance;
/* Proves one batched product sumcheck of `degree` under the labels of `prefix` (e.g. "stage/0")
 * on transcript `t`: gamma from <prefix>/gamma, instance i weighted gamma^i and front-padded to
 * n_max = max n_i, round j's message p_j(0), p_j(2), ..., p_j(d) absorbed under
 * <prefix>/round/{j}, r_j squeezed under <prefix>/r/{j}, the final evaluations (public factors
 * left out) absorbed under <prefix>/evals. Outputs: gamma (1 Fp2), rounds (n_max x degree Fp2),
 * challenges (n_max Fp2), finals (n_finals Fp2; capacity finals_cap). The rounds and their
 * transcript run on the device. An instance is a product of its tables or a sum of products
 * (terms). Errors, before the transcript moves, in the order of check_instance
 * (prover.rs:739-771): degree outside 1..=10, no instance, then per instance NoTables,
 * TooManyTables (a product of more tables than the degree) or NoTerms (terms given, none),
 * TermFactors (a term of 0 or more than `degree` factors), TermFactor (a factor naming no
 * table) (BAD_SHAPE). DECLINED above the device's limits: 10 tables, 16 terms per instance.
 * A public TSR_TABLE_EQ table in a product of three tables at degree 3 (the Sum eq(r, x) a b of a
 * claim reduction, prover.rs:186-192) is never built: the rounds weight each pair by the eq of
 * the unbound variables and fold the bound ones into a running factor. In every other use the
 * table is built on the device from its point. */
tsr_status tsr_sumcheck_prove(tsr_ctx *ctx, tsr_transcript *t, const char *prefix,
                              uint32_t degree, const tsr_instance *instances, uint32_t count,
                              uint64_t *gamma_out, uint64_t *rounds_out, uint64_t *challenges_out,
                              uint64_t *finals_out, uint64_t finals_cap, uint64_t *n_finals);

/* ---- LogUp-GKR (prove_logup, crates/tessera-logup/src/prover.rs:361-469; LogUpInstance and
 *      multiplicities, crates/tessera-logup/src/instance.rs:19-102,225-253) ---- */
typedef struct {
    uint64_t rows;                  /* T >= 1 table rows */
    const uint64_t *keys;           /* T keys (Fp) */
    const uint64_t *values;         /* T values, or NULL for a table without values */
    uint64_t lookups;               /* N >= 1 lookups */
    const uint64_t *lookup_keys;    /* N keys */
    const uint64_t *lookup_values;  /* N values; NULL exactly when `values` is NULL */
    const uint64_t *multiplicities; /* 2^ceil(log2 T) entries (Fp) */
} tsr_logup_instance;
/* A product forest (Forest::new, crates/tessera-sumcheck/src/forest.rs:30-60): 2^s stacked
 * product trees of height `height` over `len` base leaves (len = 2^(s + height)), node z of level
 * t the product of nodes 2z, 2z+1 of level t+1. */
typedef struct {
    const uint64_t *leaves; /* len leaves (Fp), host or device */
    uint64_t len;           /* a power of two */
    uint32_t height;        /* 1 <= height <= log2(len) */
} tsr_forest;
/* Proves the LogUp part of a leaf stage: alpha (logup/alpha), beta (logup/beta), each
 * instance's lookup tree (1, alpha - (key + beta value)) and table tree (-m_j, alpha - t_j),
 * both padded with row 0, their roots absorbed under logup/roots, then the levels tree/{l} of
 * the trees and the product forests in `schedule` order (2 count + nforests tree ids, 2k =
 * lookup tree k, 2k+1 = table tree k, 2 count + f = forest f; NULL = TreeId::paired followed
 * by the forests), then the values under logup/values. A level squeezes the root challenges
 * tree/{l}/root/{c} of the forests that join it (each takes the last s as its point), gamma,
 * and lambda only while a fraction tree is active. Outputs: alpha, beta, roots (4 Fp2 per
 * instance: n_a, d_a, n_t, d_t), values (one Fp2 per instance with values), claims (4 Fp2 per
 * instance: the lookup and table trees' leaf numerator and denominator), forest_claims (one
 * Fp2 per forest: its value on the stacked leaves at its last point), and, when `log` is given,
 * per level: s_max, the root challenges (s_max Fp2), has_lambda, n_rounds, gamma, lambda (0
 * without fraction trees), rounds (n x 3 Fp2), challenges, n_evals, evals (N0 N1 D0 D1 per
 * fraction tree, V0 V1 per forest), rho, transcript state (4 u64), counter. Errors before the
 * transcript moves, in this order: NoRows, NoLookups, ValuesMismatch, a missing multiplicity
 * column, a forest's LeafCount or Height, a schedule that does not name each tree once
 * (BAD_SHAPE). crates/tessera-logup/src/prover.rs:307-469. */
tsr_status tsr_logup_prove(tsr_ctx *ctx, tsr_transcript *t, const tsr_logup_instance *instances,
                           uint32_t count, const tsr_forest *forests, uint32_t nforests,
                           const uint32_t *schedule, uint64_t alpha[2], uint64_t beta[2],
                           uint64_t *roots, uint64_t *values, uint64_t *n_values,
                           uint64_t *claims, uint64_t *forest_claims, uint64_t *log,
                           uint64_t log_cap, uint64_t *log_len);
/* prove_trees (crates/tessera-sumcheck/src/trees.rs:761-776) over product forests alone, in
 * order, under tree/{l}: the level loop of tsr_logup_prove without fraction trees (no lambda).
 * forest_claims: one Fp2 per forest; `log` as tsr_logup_prove's. LeafCount or Height (BAD_SHAPE)
 * before the transcript moves. */
tsr_status tsr_prove_trees(tsr_ctx *ctx, tsr_transcript *t, const tsr_forest *forests,
                           uint32_t nforests, uint64_t *forest_claims, uint64_t *log,
                           uint64_t log_cap, uint64_t *log_len);
/* multiplicities(table, lookups): out[j] = lookups of row j over 2^ceil(log2 rows) entries, row 0
 * also counting the lookup tree's 2^ceil(log2 N) - N padding leaves. LookupRow (BAD_SHAPE) when a
 * lookup names no row, `out` untouched. lookup_rows and out may each be host or device memory; a
 * device `out` is written on the device. A table of no row has no leaf: with no lookup that is
 * OK and nothing is written (`out` may be NULL), with any lookup it is LookupRow
 * (crates/tessera-logup/src/instance.rs:225-253). The device copies of the rows and counts are
 * cleared before the library reuses their memory. */
tsr_status tsr_logup_multiplicities(tsr_ctx *ctx, uint64_t rows, const uint64_t *lookup_rows,
                                    uint64_t lookups, uint64_t *out);

/* ---- the weight commitment (RFC-0040: WeightLayout::new / check, WeightCommitment::commit,
 *      WeightCommitmentPublic; crates/tessera-pcs/src/weights/{layout.rs:55-260,466-560,
 *      767-870, commit.rs:61-135, public.rs:22-48}, descriptor.rs:118-160) ---- */
typedef struct {
    uint8_t log2_k;          /* k_W = 2^log2_k (WeightLayoutParams) */
    uint8_t log2_inv_rate;   /* r >= 1 */
    uint32_t mask_coeffs;    /* P_W (carried in the descriptor and the public artifact) */
    uint32_t blinding_rows;  /* B_W (likewise) */
} tsr_weight_layout_params;
typedef struct {
    uint32_t out, inputs;    /* WeightShape: output channels (rows) x input channels */
} tsr_weight_shape;
typedef struct {             /* WeightTensorRef */
    uint32_t matrix;         /* WeightMatrix: Embed 0, Q, K, V, O, Gate, Up, Down, LmHead 8 */
    uint32_t layer;          /* 0 for Embed and LmHead */
    const int8_t *data;      /* out x inputs INT8 entries, row-major; host or device memory */
    uint64_t len;
} tsr_weight_tensor;
typedef struct {             /* ChannelParamsRef */
    uint32_t matrix, layer;
    uint32_t kind;           /* 0: class indices (u8), 1: multipliers (u16) */
    uint64_t count;          /* one per output channel */
    const void *values;      /* host memory */
} tsr_channel_params;
typedef struct tsr_weight_commitment tsr_weight_commitment;
/* Commits a model's weights: the layout of `layers` layers and the nine weight types' shapes
 * (channel_weights: n_channel_weights (matrix, layer) pairs, the weights that have a
 * ChannelParams), then each of the ten matrices' committed rows, laid out from the INT8 entries
 * on the device, committed unmasked at rate 2^-r with chains of `row_block` rows from
 * commitment_id_T = H("commitment-id", TCE(CommitmentDescriptor::WeightMatrix)); the public
 * artifact TCE(layout) || the ten roots || u32 epoch and R_W = H("weights-root", it). Errors
 * before any work, in the reference's order: EmptyWeight, InverseRate, TooLong, WeightPlace,
 * WeightTwice (of the layout's channel list), then WeightPlace, WeightLength, WeightTwice,
 * WeightMissing of the tensors and of the channel parameters (BAD_SHAPE). */
tsr_status tsr_weight_commit(tsr_ctx *ctx, const tsr_weight_layout_params *params, uint32_t layers,
                             const tsr_weight_shape shapes[9], const uint32_t *channel_weights,
                             uint32_t n_channel_weights, uint32_t row_block, uint32_t epoch,
                             const tsr_weight_tensor *tensors, uint32_t ntensors,
                             const tsr_channel_params *channels, uint32_t nchannels,
                             tsr_weight_commitment **out);
/* TCE(WeightCommitmentPublic), 334 bytes (public.rs:22-31). */
tsr_status tsr_weight_commitment_public(const tsr_weight_commitment *w, uint8_t out[334]);
/* R_W = H("weights-root", TCE(WeightCommitmentPublic)) (public.rs:33-36). */
tsr_status tsr_weight_commitment_weights_root(const tsr_weight_commitment *w, uint8_t out[32]);
/* The .twr artifact: the public part sealed in its envelope (schema 0x0301, version 1), 386
 * bytes (WeightCommitmentPublic::write, public.rs:48). */
tsr_status tsr_weight_commitment_artifact(const tsr_weight_commitment *w, uint8_t out[386]);
/* WeightCommitment::matrix (commit.rs:148): matrix T's commitment (borrowed; for root, leaves,
 * answer), its log2 k_T, row bits and committed rows m_T. */
tsr_status tsr_weight_commitment_matrix(const tsr_weight_commitment *w, uint32_t matrix,
                                        const tsr_commitment **commitment, uint32_t *log2_k,
                                        uint32_t *row_vars, uint64_t *rows);
void tsr_weight_commitment_free(tsr_weight_commitment *w);
/* Envelope::seal (crates/tessera-codec/src/envelope.rs; RFC-0003 'Envelope'): b"TSRA", u16 1,
 * u16 schema_id, u16 schema_version, u16 0, u64 payload length, the payload, then
 * H("envelope", every preceding byte): len + 52 bytes into `out` (capacity `cap`). */
tsr_status tsr_envelope_seal(tsr_ctx *ctx, uint16_t schema_id, uint16_t schema_version,
                             const uint8_t *payload, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *out_len);

/* ---- wire formats: TCE (RFC-0003; crates/tessera-codec/src/tce.rs:306-441) of SumcheckProof
 *      (crates/tessera-sumcheck/src/proof.rs:9-24) and the commitment descriptors
 *      (crates/tessera-pcs/src/descriptor.rs:10-160). Host code: no context or device needed. ---- */
typedef struct {
    uint64_t max_total_bytes; /* DecodeLimits (RFC-0003 'Limits') */
    uint64_t max_vec_len;
    uint32_t max_depth;       /* above 256 counts as 256 */
} tsr_decode_limits;
enum {
    TSR_CODEC_OK = 0, TSR_CODEC_MAGIC = 1, TSR_CODEC_VERSION = 2, TSR_CODEC_NONCANONICAL = 3,
    TSR_CODEC_TRUNCATED = 4, TSR_CODEC_DIGEST = 5, TSR_CODEC_LIMIT = 6
};
/* TCE(SumcheckProof { rounds, final_evals }): u64 n_rounds, each round u64 degree then its
 * degree Fp2 (p(0), p(2), ..., p(d)), then u64 n_finals and the finals. `rounds` as the prover
 * writes them (n_rounds x degree Fp2). The size needed is written to out_len; BAD_SHAPE when
 * `cap` is short, BAD_ARG for a non-canonical value. */
tsr_status tsr_sumcheck_proof_encode(const uint64_t *rounds, uint32_t n_rounds, uint32_t degree,
                                     const uint64_t *finals, uint64_t n_finals, uint8_t *out,
                                     uint64_t cap, uint64_t *out_len);
/* Decodes a SumcheckProof under `limits` with RFC-0003's checks and offsets: round j's length in
 * round_lens[j], the round values one after another in `values` (Fp2), the finals in `finals`.
 * A malformed input is BAD_SHAPE with codec_error (TSR_CODEC_*) and error_offset set. */
tsr_status tsr_sumcheck_proof_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                                     uint32_t *n_rounds, uint32_t *round_lens, uint32_t round_lens_cap,
                                     uint64_t *values, uint64_t values_cap, uint64_t *n_values,
                                     uint64_t *finals, uint64_t finals_cap, uint64_t *n_finals,
                                     int *codec_error, uint64_t *error_offset);
/* One opened weight matrix of a Round3Response (MatrixColumns,
 * crates/tessera-protocol/src/round3.rs:150-175). */
typedef struct {
    uint8_t matrix;           /* WeightMatrix id 0..9; strictly increasing across the entries */
    uint64_t columns;         /* distinct queried columns, in increasing order */
    uint64_t symbols;         /* symbols per column */
    const uint64_t *values;   /* columns x symbols, column after column */
    uint64_t siblings;        /* the canonical multiproof's siblings */
    const uint8_t *proof;     /* siblings x 32 bytes */
} tsr_matrix_columns;
/* TCE(Round3Response { advice_columns, advice_proof, weight_columns }) (round3.rs:100-175,
 * schema 0x0513): Vec<Vec<Fp>> of the advice's answered columns (advice_values: columns x
 * symbols, column after column, as tsr_commitment_answer gives them), the advice multiproof
 * (Vec<Digest>), then Vec<MatrixColumns>. Non-canonical symbols or unordered matrix ids are
 * BAD_ARG; a short `out` is BAD_SHAPE with *out_len the size needed. */
tsr_status tsr_round3_encode(const uint64_t *advice_values, uint64_t advice_columns,
                             uint64_t advice_symbols, const uint8_t *advice_proof,
                             uint64_t advice_siblings, const tsr_matrix_columns *weights,
                             uint32_t n_weights, uint8_t *out, uint64_t cap, uint64_t *out_len);
/* Decodes and checks a Round3Response with the decoder's checks in order (limits, depth with
 * structs, sequences and digests one level each, lengths against the bytes left, canonical
 * symbols, matrix ids below 10 and strictly increasing, no trailing byte): BAD_SHAPE with
 * codec_error and error_offset on a malformed input, else the counts of advice columns and
 * opened matrices. */
tsr_status tsr_round3_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                             uint64_t *advice_columns, uint32_t *n_matrices, int *codec_error,
                             uint64_t *error_offset);
/* TCE(ColumnQueries { advice: Vec<u32>, weights: Vec<MatrixQueries { matrix, indices }> })
 * (round3.rs:19-95, schema 0x0512): the queries in draw order; matrix m's counts[m] indices
 * follow one another in `indices`. */
tsr_status tsr_column_queries_encode(const uint32_t *advice, uint64_t n_advice, const uint8_t *matrices,
                                     const uint64_t *counts, const uint32_t *indices, uint32_t n_matrices,
                                     uint8_t *out, uint64_t cap, uint64_t *out_len);
/* Decodes the verifier's ColumnQueries (the checks as tsr_round3_decode's) into the advice
 * indices, the matrix ids with their index counts, and every matrix's indices one after another. */
tsr_status tsr_column_queries_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                                     uint32_t *advice, uint64_t advice_cap, uint64_t *n_advice,
                                     uint8_t *matrices, uint64_t *counts, uint32_t matrices_cap,
                                     uint32_t *n_matrices, uint32_t *indices, uint64_t indices_cap,
                                     uint64_t *n_indices, int *codec_error, uint64_t *error_offset);
typedef struct {             /* CommitmentDescriptor::Advice */
    uint8_t statement_digest[32];
    uint32_t unit;
    uint8_t log2_k, log2_inv_rate;  /* AdviceLayoutParams */
    uint32_t mask_coeffs, row_block, segments;
    uint64_t rows;           /* the unit's non-zero rows, every segment's */
} tsr_advice_descriptor;
/* TCE(CommitmentDescriptor::Advice) (60 bytes, may be NULL) and the commitment_id
 * H("commitment-id", it) the masked commitment's chains start from (RFC-0031). */
tsr_status tsr_advice_commitment_id(const tsr_advice_descriptor *d, uint8_t tce_out[60],
                                    uint8_t commitment_id[32]);
/* TCE(CommitmentDescriptor::WeightMatrix) (31 bytes, may be NULL) and its commitment_id. */
tsr_status tsr_weight_matrix_commitment_id(const tsr_weight_layout_params *layout, uint32_t matrix,
                                           uint8_t log2_k, uint32_t row_block, uint64_t rows,
                                           uint32_t epoch, uint8_t tce_out[31],
                                           uint8_t commitment_id[32]);

/* ---- the advice mask stream (Masks as a MaskSource, crates/tessera-prover/src/cpu/random.rs:
 *      123-229,275-320; the tessera/v1/advice-masks domain, crates/tessera-hash/src/domain.rs:
 *      90-127) ---- */
typedef struct tsr_mask_stream tsr_mask_stream;
/* A mask stream over the 32-byte secret per-proof seed: the BLAKE3 derive-key extended output of
 * the seed under "tessera/v1/advice-masks", read as u64 little-endian words, a word >= p discarded
 * whole (MaskSource::draw). The seed is copied; the copy is cleared by tsr_mask_stream_free. */
tsr_status tsr_mask_stream_create(tsr_ctx *ctx, const uint8_t seed[32], tsr_mask_stream **out);
/* The next `count` draws of the stream into `out` (host or device memory), generated on the
 * device: the masks a commitment takes (N rows x P coefficients, m_0 first, row after row, in
 * layout order; commitment.rs:385-391). Two draws in a row equal one draw of their sum. */
tsr_status tsr_mask_stream_draw(tsr_mask_stream *m, uint64_t count, uint64_t *out);
/* The number of stream words consumed so far (draws plus discarded words). */
uint64_t tsr_mask_stream_position(const tsr_mask_stream *m);
void tsr_mask_stream_free(tsr_mask_stream *m);

/* ---- the level offload (LevelOffload::load, DeviceLevel::{sums, bind, firsts},
 *      crates/tessera-sumcheck/src/offload.rs:17-79; the reference's GpuLevel,
 *      crates/tessera-gpu/src/level.rs:52-307) ---- */
typedef struct tsr_level tsr_level;
enum {
    TSR_LEVEL_FRACTION = 0, /* N0 D1 + N1 D0 + lambda D0 D1 over [N0, N1, D0, D1] */
    TSR_LEVEL_PRODUCT = 1   /* V0 V1 over [V0, V1] */
};
/* LevelOffload::load: `n_tables` tables (4 for a fraction level, 2 for a product level) of 2^m
 * Fp2 entries each (host or device; copied), 1 <= m <= 32. */
tsr_status tsr_level_load(tsr_ctx *ctx, int shape, const uint64_t lambda[2],
                          const uint64_t *const *tables, uint32_t n_tables, uint32_t m,
                          tsr_level **out);
/* DeviceLevel::sums: Sum_p w(p) P(line_x(p)) for x = 0, 2, 3 (3 Fp2 in `out`) over the pairs p of
 * the current tables, w(p) = low[p mod 2^low_bits] high[p >> low_bits] (2^low_bits and high_len
 * Fp2), line_x(p) = t[2p] + x (t[2p+1] - t[2p]). BAD_SHAPE unless the weights cover the pairs. */
tsr_status tsr_level_sums(tsr_level *level, const uint64_t *low, uint32_t low_bits,
                          const uint64_t *high, uint64_t high_len, uint64_t out[6]);
/* DeviceLevel::bind: t'[i] = t[2i] + r (t[2i+1] - t[2i]) for every table. */
tsr_status tsr_level_bind(tsr_level *level, const uint64_t r[2]);
/* DeviceLevel::firsts: each table's first entry (n_tables Fp2). */
tsr_status tsr_level_firsts(tsr_level *level, uint64_t *out);
void tsr_level_free(tsr_level *level);

/* eq(r, .) over n variables (crates/tessera-poly/src/eq.rs:23: entry x = prod_i (r_i x_i +
 * (1 - r_i)(1 - x_i)), variables LSB-first): `point` n Fp2 on the host, `out` 2^n Fp2 (device or
 * host). BAD_SHAPE for n > 30. */
tsr_status tsr_eq_table(tsr_ctx *ctx, const uint64_t *point, uint32_t n, uint64_t *out);

/* ---- per-kernel timing with HIP events on the context's stream (the reference's CUDA-event
 *      harness, docs/perf/v0.1/cc-h100.md:30-40) ---- */
/* enable = 1 starts recording an event pair around every launch; 0 stops; 2 clears. */
tsr_status tsr_ctx_profile(tsr_ctx *ctx, int enable);
/* Total milliseconds and launch count of the named kernel since the last clear. */
tsr_status tsr_ctx_kernel_time(tsr_ctx *ctx, const char *kernel, double *total_ms,
                               uint64_t *launches);
/* The most workgroups one sumcheck round launch uses (1..1024; 0 restores the default, 1024).
 * Results do not depend on it. No reference counterpart: a tuning knob of this back end (fewer
 * workgroups give each thread more pairs of a round, which changes the kernels' code paths at
 * smaller sizes). BAD_ARG above 1024. */
tsr_status tsr_ctx_set_round_blocks(tsr_ctx *ctx, uint32_t blocks);
/* The device's register-resident Goldilocks butterfly rate (u + v, (u - v) w per butterfly, no
 * memory traffic): the VALU ceiling the NTT kernels are measured against. No reference
 * counterpart (the reference quotes butterflies per second, docs/perf/v0.1/cc-h100.md). */
tsr_status tsr_butterfly_peak(tsr_ctx *ctx, double *butterflies_per_s);

#ifdef __cplusplus
}
#endif
#endif
