/*
 * Test infrastructure: the CPU oracle. NOT part of the product.
 *
 * A plain-C restatement of the reference CPU algorithms of AbdelStark/tessera for the
 * data-parallel prover core (Goldilocks field, NTT / Reed-Solomon row encoding, BLAKE3
 * derive-key domains, column chains, Merkle trees and multiproofs, the Fiat-Shamir transcript,
 * the batched product sumcheck, fraction trees and LogUp). Each function cites the reference
 * file:line it follows. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load this library, and only as the checker or the timed CPU baseline; the product
 * library (tessera_amd/) never links or calls it.
 *
 * Pinning (what fixes each part, besides the reference source it restates):
 *   field      spec/vectors/field/{constants,goldilocks,goldilocks_ext2}.json + closed forms
 *   blake3     35 official BLAKE3 vectors (spec/vectors/blake3/test_vectors.json) and the
 *              19 domain digests of spec/vectors/hash/hash.json
 *   ntt/encode spec/vectors/encoding/encoding.json (ntt + rate 1/2 encodings) and the
 *              rate-1/4 encodings of spec/vectors/weights/weights.json, plus O(n^2) direct
 *              evaluation and intt(ntt(x)) = x in tests
 *   chains     spec/vectors/hash/hash.json "columns"
 *   merkle     spec/vectors/hash/hash.json "merkle" (roots and canonical multiproofs)
 *   transcript spec/vectors/transcript/transcript.json (6 sessions, every step)
 *   sumcheck   spec/vectors/sumcheck/sumcheck.json "sessions" (messages, challenges,
 *              final evaluations, transcript states) and verifier invariants
 *   trees      spec/vectors/sumcheck/sumcheck.json "tree_batches" and "logup_batches"
 *
 * Everything is exact modular integer arithmetic: p = 2^64 - 2^32 + 1, Fp2 = Fp[X]/(X^2 - 7).
 */
#ifndef TESSERA_ORACLE_H
#define TESSERA_ORACLE_H

#include <stddef.h>
#include <stdint.h>
#include "fp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_P 0xFFFFFFFF00000001ULL

/* ---- field (crates/tessera-field/src/goldilocks.rs, ext2.rs) ---- */
uint64_t orc_fp_add(uint64_t a, uint64_t b);
uint64_t orc_fp_sub(uint64_t a, uint64_t b);
uint64_t orc_fp_mul(uint64_t a, uint64_t b);
uint64_t orc_fp_pow(uint64_t a, uint64_t e);
uint64_t orc_fp_inv(uint64_t a);
uint64_t orc_fp_root(uint32_t log_n);
uint64_t orc_fp_reduce128(uint64_t lo, uint64_t hi);
orc_fp2 orc_fp2_add(orc_fp2 a, orc_fp2 b);
orc_fp2 orc_fp2_sub(orc_fp2 a, orc_fp2 b);
orc_fp2 orc_fp2_mul(orc_fp2 a, orc_fp2 b);
orc_fp2 orc_fp2_mul_base(orc_fp2 a, uint64_t b);
orc_fp2 orc_fp2_inv(orc_fp2 a);
void orc_fp2_mul_v(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]);
void orc_fp_mul_vec(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n);
void orc_fp2_inv_v(const uint64_t a[2], uint64_t out[2]);

/* ---- BLAKE3 (spec; crate blake3 1.8.7 as called by crates/tessera-hash/src/domain.rs) ---- */
/* Full BLAKE3 hash in one of the three modes: mode 0 = hash, 1 = keyed (key 32 bytes),
 * 2 = derive_key (key = context string, key_len its length). out_len bytes of XOF output. */
void orc_blake3(int mode, const uint8_t *key, size_t key_len, const uint8_t *in, size_t len,
                uint8_t *out, size_t out_len);
/* Context key of "tessera/v1/<name>" for domain index d (domain.rs:41-81 order). */
int orc_domain_count(void);
const char *orc_domain_context(int d);
void orc_domain_key(int d, uint8_t key[32]);
/* H(domain, bytes): derive_key(context, bytes) (domain.rs:177-181). */
void orc_hash(int d, const uint8_t *in, size_t len, uint8_t out[32]);

enum {
    ORC_DOM_ENVELOPE = 0, ORC_DOM_COMMITMENT_ID, ORC_DOM_COLUMN_INIT, ORC_DOM_COLUMN_STEP,
    ORC_DOM_MERKLE_NODE, ORC_DOM_COLUMN_MERGE, ORC_DOM_WEIGHTS_ROOT, ORC_DOM_TRANSCRIPT_INIT,
    ORC_DOM_TRANSCRIPT_ABSORB, ORC_DOM_TRANSCRIPT_SQUEEZE, ORC_DOM_TRANSCRIPT_RATCHET,
    ORC_DOM_POW, ORC_DOM_COIN_COMMIT, ORC_DOM_COIN_TOSS, ORC_DOM_TABLE, ORC_DOM_AUDIT_SAMPLE,
    ORC_DOM_AUDIT_SELECT, ORC_DOM_CLAIM_LIST, ORC_DOM_STATEMENT, ORC_DOM_COUNT
};

/* ---- NTT and row encoding (crates/tessera-pcs/src/ntt.rs, encoding.rs) ---- */
void orc_twiddles(uint32_t log_n, int inverse, uint64_t *out /* n/2 */);
void orc_ntt(uint64_t *values, uint32_t log_n);
void orc_intt(uint64_t *values, uint32_t log_n);
void orc_coset_ntt(uint64_t *values, uint32_t log_n, uint64_t shift);
void orc_coset_intt(uint64_t *values, uint32_t log_n, uint64_t shift);
/* RowEncoder::encode_rows (encoding.rs:257-290): rows (count x k), masks (count x P, m_0
 * first, row after row), codewords (count x n). threads > 1 splits rows over pthreads. */
int orc_encode_rows(uint32_t log_k, uint32_t log_inv_rate, uint64_t mask_coeffs,
                    const uint64_t *rows, const uint64_t *masks, uint64_t count,
                    uint64_t *codewords, int threads);

/* ---- column chains and Merkle (crates/tessera-hash/src/columns.rs, merkle.rs) ---- */
void orc_column_init(const uint8_t id[32], uint64_t columns, uint8_t *chains /* 32*columns */);
/* absorb `count` rows (row-major, `columns` symbols each) into every chain (columns.rs:90). */
void orc_column_absorb(uint8_t *chains, uint64_t columns, const uint64_t *rows, uint64_t count,
                       int threads);
void orc_merkle_node(const uint8_t l[32], const uint8_t r[32], uint8_t out[32]);
/* Every level, leaves first, 2n-1 digests (merkle.rs:93-112). */
int orc_merkle_tree(const uint8_t *leaves, uint64_t n, uint8_t *nodes, int threads);
/* Canonical multiproof (merkle.rs:126-169); returns sibling count or -1. */
int64_t orc_merkle_open(const uint8_t *nodes, uint64_t n, const uint64_t *indices, uint64_t q,
                        uint8_t *siblings /* capacity 32*cap */, uint64_t cap);
int orc_merkle_verify(const uint8_t root[32], uint64_t n, const uint64_t *idx,
                      const uint8_t *leaf_digests, uint64_t q, const uint8_t *siblings,
                      uint64_t n_siblings);
/* ReferenceCommitment::commit_unmasked (commitment.rs:343-446): encode, chain every column over
 * all rows in blocks of `block`, leaves = chains, Merkle tree. nodes: 2n-1 digests. */
void orc_row_coefficients(uint32_t row_vars, const uint64_t *rows, uint64_t N,
                          const orc_fp2 *row_points, const orc_fp2 *gammas,
                          const uint32_t *group_sizes, uint32_t groups, const orc_fp2 *xi_point,
                          orc_fp2 *out);
void orc_combine(const orc_fp2 *coefficients, uint32_t count, const uint64_t *matrix, uint64_t N,
                 uint64_t width, orc_fp2 *out);
int orc_commit_masked(uint32_t log_k, uint32_t log_r, uint64_t P, uint64_t block,
                      const uint8_t id[32], const uint64_t *segment_rows, uint32_t segments,
                      const uint64_t *rows, const uint64_t *masks, uint64_t *codewords,
                      uint8_t *leaves, uint8_t *nodes, int threads);
int orc_commit_unmasked(uint32_t log_k, uint32_t log_inv_rate, uint64_t block,
                        const uint8_t id[32], const uint64_t *rows, uint64_t count,
                        uint64_t *codewords, uint8_t *nodes, int threads);

/* ---- transcript (crates/tessera-hash/src/transcript.rs) ---- */
typedef struct { uint8_t state[32]; uint64_t counter; } orc_transcript;
void orc_transcript_init(orc_transcript *t, const uint8_t *context_tce, size_t len);
void orc_transcript_absorb(orc_transcript *t, const char *label, const uint8_t *tce, size_t len);
void orc_transcript_squeeze(orc_transcript *t, const char *label, uint8_t out[32]);
uint64_t orc_transcript_challenge_fp(orc_transcript *t, const char *label, int *rejected);
orc_fp2 orc_transcript_challenge_fp2(orc_transcript *t, const char *label, int *rejected);
void orc_transcript_challenge_fp2_v(orc_transcript *t, const char *label, uint64_t out[2]);
void orc_transcript_challenge_indices(orc_transcript *t, const char *label, uint32_t log_n,
                                      uint32_t count, uint64_t *out);
uint64_t orc_transcript_grind(orc_transcript *t, const char *label, uint8_t bits);
int orc_transcript_check_grind(orc_transcript *t, const char *label, uint8_t bits,
                               uint64_t nonce);
/* absorb a Vec<Fp2> as TCE: u64 length then c0 || c1 of each (sumcheck/src/transcript.rs). */
void orc_transcript_absorb_fp2s(orc_transcript *t, const char *label, const uint64_t *v,
                                size_t count);

/* ---- multilinear + batched product sumcheck (crates/tessera-poly, tessera-sumcheck) ---- */
/* One factor table: 2^vars entries, base (u64 each) or extension (2 u64 each), public or not. */
typedef struct {
    const uint64_t *data;
    int ext;      /* 1: Fp2 entries (c0,c1), 0: Fp entries */
    int is_public;/* 1: omitted from the final evaluations (Table::public) */
} orc_table;
/* One term of a sum-of-products instance (Term, crates/tessera-sumcheck/src/terms.rs:10-40):
 * weight * prod of the tables at `factors` (positions in the instance's list; may repeat). */
typedef struct {
    uint64_t weight[2];
    uint32_t nfactors;
    const uint32_t *factors;
} orc_term;
typedef struct {
    uint32_t vars;
    uint32_t factors;          /* the number of tables */
    const orc_table *tables;
    const orc_term *terms;     /* NULL: the product of every table (Instance::new) */
    uint32_t nterms;           /* with terms: Instance::sum_of_products (prover.rs:300-312) */
} orc_instance;
/* prove_batch of one stage (prover.rs:1122-1421): gamma, rounds, finals. Outputs:
 *   rounds: n_max * degree Fp2 (p_j(0), p_j(2..d)), challenges: n_max Fp2,
 *   finals: every non-public factor's evaluation, instance by instance (Fp2 each),
 *   claim: the batched claim (Fp2). Returns number of finals or -1. */
int64_t orc_prove_batch(orc_transcript *t, const char *prefix, uint32_t degree,
                        const orc_instance *instances, uint32_t count, uint64_t *gamma_out,
                        uint64_t *claim_out, uint64_t *rounds, uint64_t *challenges,
                        uint64_t *finals, int threads);
/* eq(r, x) table of 2^n Fp2 entries, x_1 = least significant bit (tessera-poly eq.rs:23). */
void orc_eq_table(const uint64_t *point /* n Fp2 */, uint32_t n, uint64_t *out);
/* evaluate a multilinear table (base or ext) at an Fp2 point (multilinear.rs:136-180). */
void orc_mle_eval(const uint64_t *table, int ext, uint32_t n, const uint64_t *point,
                  uint64_t out[2]);
/* RoundPoly::evaluate: messages p(0), p(2..d) with p(1) = claim - p(0) (round.rs:91-151). */
void orc_round_eval(const uint64_t *evals /* d Fp2 */, uint32_t degree, const uint64_t claim[2],
                    const uint64_t r[2], uint64_t out[2]);

/* ---- LogUp (crates/tessera-logup/src/{instance,prover}.rs, tessera-sumcheck trees.rs) ---- */
typedef struct {
    uint64_t rows;                 /* T >= 1 */
    const uint64_t *keys;          /* T keys */
    const uint64_t *values;        /* T values, or NULL for a table without values */
    uint64_t lookups;              /* N >= 1 */
    const uint64_t *lookup_keys;   /* N keys */
    const uint64_t *lookup_values; /* N values, NULL iff values is NULL */
    const uint64_t *lookup_rows;   /* N row indices (orc_multiplicities only) */
    const uint64_t *multiplicities;/* 2^{h_t} entries (Fp) */
} orc_logup_instance;
/* multiplicities over 2^{h_t} table leaves (instance.rs:225-253). */
void orc_multiplicities(const orc_logup_instance *inst, uint64_t *out /* 2^h_t */);
/* prove_logup (prover.rs:361-469) with fraction trees only: schedule lists 2*count tree ids
 * (2k = lookup tree k, 2k+1 = table tree k). Outputs alpha, beta, roots (4 Fp2 per instance),
 * values (one Fp2 per instance with values), leaf claims (4 Fp2 per instance: lookup num/den,
 * table num/den) and a per-level log (see logup.c). Returns 0. */
typedef struct { const uint64_t *leaves; uint64_t len; uint32_t height; } orc_forest;
int orc_prove_trees(orc_transcript *tr, const orc_forest *forests, uint32_t nforests,
                    uint64_t *forest_claims, uint64_t *log, uint64_t cap, uint64_t *log_len);
int orc_prove_logup(orc_transcript *tr, const orc_logup_instance *inst, uint32_t count,
                    const orc_forest *forests, uint32_t nforests,
                    const uint32_t *schedule, uint64_t *alpha_out, uint64_t *beta_out,
                    uint64_t *roots, uint64_t *values, uint64_t *n_values, uint64_t *claims,
                    uint64_t *forest_claims, uint64_t *log, uint64_t cap, uint64_t *log_len);

#ifdef __cplusplus
}
#endif
#endif
