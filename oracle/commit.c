/*
 * Test infrastructure (oracle): column chains, Merkle trees and canonical multiproofs, and the
 * unmasked reference commitment.
 * Restates crates/tessera-hash/src/columns.rs, crates/tessera-hash/src/merkle.rs and
 * crates/tessera-pcs/src/commitment.rs (commit_unmasked, from_encoded, answer).
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "internal.h"

/* ColumnChains::new (columns.rs:72-88): d_j[0] = H(column-init, id || u64LE(j)). */
void orc_column_init(const uint8_t id[32], uint64_t columns, uint8_t *chains) {
    uint8_t in[40];
    memcpy(in, id, 32);
    for (uint64_t j = 0; j < columns; j++) {
        orc_put_u64(in + 32, j);
        orc_hash(ORC_DOM_COLUMN_INIT, in, 40, chains + 32 * j);
    }
}

typedef struct {
    uint8_t *chains;
    uint64_t columns, count;
    const uint64_t *rows;
} absorb_t;

/* absorb_block (columns.rs:90-104): d_j <- H(column-step, d_j || u64LE(sym) of each row). */
static void absorb_range(void *arg, uint64_t lo, uint64_t hi) {
    const absorb_t *a = arg;
    uint8_t *in = malloc(32 + 8 * a->count);
    for (uint64_t j = lo; j < hi; j++) {
        memcpy(in, a->chains + 32 * j, 32);
        for (uint64_t r = 0; r < a->count; r++) orc_put_u64(in + 32 + 8 * r, a->rows[r * a->columns + j]);
        orc_hash(ORC_DOM_COLUMN_STEP, in, 32 + 8 * a->count, a->chains + 32 * j);
    }
    free(in);
}

void orc_column_absorb(uint8_t *chains, uint64_t columns, const uint64_t *rows, uint64_t count,
                       int threads) {
    absorb_t a = {chains, columns, count, rows};
    orc_parallel_for(columns, threads, absorb_range, &a);
}

/* merkle_node (merkle.rs:10-16). */
void orc_merkle_node(const uint8_t l[32], const uint8_t r[32], uint8_t out[32]) {
    uint8_t in[64];
    memcpy(in, l, 32);
    memcpy(in + 32, r, 32);
    orc_hash(ORC_DOM_MERKLE_NODE, in, 64, out);
}

typedef struct {
    const uint8_t *below;
    uint8_t *above;
} level_t;
static void level_range(void *arg, uint64_t lo, uint64_t hi) {
    const level_t *l = arg;
    for (uint64_t i = lo; i < hi; i++)
        orc_merkle_node(l->below + 64 * i, l->below + 64 * i + 32, l->above + 32 * i);
}

/* MerkleTree::new (merkle.rs:93-112): levels appended leaves first. */
int orc_merkle_tree(const uint8_t *leaves, uint64_t n, uint8_t *nodes, int threads) {
    if (n == 0 || (n & (n - 1))) return -1;
    memcpy(nodes, leaves, 32 * n);
    uint64_t start = 0, width = n;
    while (width > 1) {
        level_t l = {nodes + 32 * start, nodes + 32 * (start + width)};
        orc_parallel_for(width / 2, width >= 4096 ? threads : 1, level_range, &l);
        start += width;
        width /= 2;
    }
    return 0;
}

static int cmp_u64(const void *a, const void *b) {
    uint64_t x = *(const uint64_t *)a, y = *(const uint64_t *)b;
    return x < y ? -1 : x > y;
}

static uint64_t dedup_sorted(uint64_t *v, uint64_t q) {
    uint64_t m = 0;
    for (uint64_t i = 0; i < q; i++)
        if (m == 0 || v[m - 1] != v[i]) v[m++] = v[i];
    return m;
}

/* MerkleTree::open (merkle.rs:126-169): siblings level by level, left to right. */
int64_t orc_merkle_open(const uint8_t *nodes, uint64_t n, const uint64_t *indices, uint64_t q,
                        uint8_t *siblings, uint64_t cap) {
    if (q == 0) return -1;
    for (uint64_t i = 0; i < q; i++)
        if (indices[i] >= n) return -1;
    uint64_t *known = malloc(8 * q), *parents = malloc(8 * q);
    memcpy(known, indices, 8 * q);
    qsort(known, q, 8, cmp_u64);
    uint64_t m = dedup_sorted(known, q);
    int64_t out = 0;
    if (m == n) goto done;
    uint64_t start = 0, width = n;
    while (width > 1) {
        uint64_t pm = 0, i = 0;
        while (i < m) {
            uint64_t node = known[i];
            if ((node & 1) == 0 && i + 1 < m && known[i + 1] == node + 1) {
                i += 2;
            } else {
                if ((uint64_t)out >= cap) { out = -1; goto done; }
                memcpy(siblings + 32 * out, nodes + 32 * (start + (node ^ 1)), 32);
                out++;
                i += 1;
            }
            parents[pm++] = node / 2;
        }
        m = dedup_sorted(parents, pm);
        memcpy(known, parents, 8 * m);
        start += width;
        width /= 2;
    }
done:
    free(known);
    free(parents);
    return out;
}

/* verify_multiproof (merkle.rs:183-225). */
int orc_merkle_verify(const uint8_t root[32], uint64_t n, const uint64_t *idx,
                      const uint8_t *leaf_digests, uint64_t q, const uint8_t *siblings,
                      uint64_t n_siblings) {
    if (q == 0 || n == 0 || (n & (n - 1))) return 0;
    for (uint64_t i = 0; i + 1 < q; i++)
        if (idx[i] >= idx[i + 1]) return 0;
    if (idx[q - 1] >= n) return 0;
    uint64_t *kn = malloc(8 * q), *pn = malloc(8 * q);
    uint8_t *kd = malloc(32 * q), *pd = malloc(32 * q);
    memcpy(kn, idx, 8 * q);
    memcpy(kd, leaf_digests, 32 * q);
    uint64_t m = q, used = 0, width = n;
    int ok = 1;
    while (width > 1 && ok) {
        uint64_t pm = 0, i = 0;
        while (i < m) {
            uint64_t node = kn[i];
            const uint8_t *l, *r;
            if ((node & 1) == 0) {
                if (i + 1 < m && kn[i + 1] == node + 1) {
                    l = kd + 32 * i;
                    r = kd + 32 * (i + 1);
                    i++;
                } else {
                    if (used >= n_siblings) { ok = 0; break; }
                    l = kd + 32 * i;
                    r = siblings + 32 * used++;
                }
            } else {
                if (used >= n_siblings) { ok = 0; break; }
                l = siblings + 32 * used++;
                r = kd + 32 * i;
            }
            pn[pm] = node / 2;
            orc_merkle_node(l, r, pd + 32 * pm);
            pm++;
            i++;
        }
        memcpy(kn, pn, 8 * pm);
        memcpy(kd, pd, 32 * pm);
        m = pm;
        width /= 2;
    }
    ok = ok && used == n_siblings && m == 1 && kn[0] == 0 && memcmp(kd, root, 32) == 0;
    free(kn);
    free(pn);
    free(kd);
    free(pd);
    return ok;
}

/* ReferenceCommitment::commit_unmasked -> encode -> from_encoded (commitment.rs:343-446):
 * codewords of every row, one chain per column over the rows in blocks of B (the last block may
 * be short), leaf_j = the chain's last digest, Merkle tree over the n leaves. */
int orc_commit_unmasked(uint32_t log_k, uint32_t log_r, uint64_t block, const uint8_t id[32],
                        const uint64_t *rows, uint64_t count, uint64_t *codewords,
                        uint8_t *nodes, int threads) {
    if (block == 0 || log_r == 0 || log_k + log_r > 32) return -1;
    uint64_t n = 1ULL << (log_k + log_r);
    if (orc_encode_rows(log_k, log_r, 0, rows, NULL, count, codewords, threads)) return -1;
    uint8_t *chains = malloc(32 * n);
    orc_column_init(id, n, chains);
    for (uint64_t r = 0; r < count; r += block) {
        uint64_t c = count - r < block ? count - r : block;
        orc_column_absorb(chains, n, codewords + r * n, c, threads);
    }
    int rc = orc_merkle_tree(chains, n, nodes, threads);
    free(chains);
    return rc;
}

/* ReferenceCommitment::commit -> encode -> from_encoded (commitment.rs:318-446) for the masked
 * advice commitment: the N = sum(segment_rows) rows encoded with their P masks each (m_0
 * first, in layout order), one set of chains per segment over that segment's rows in blocks of
 * B (an empty segment keeps d_j[0]), leaf_j = H(column-merge, d_j^(1) || ... || d_j^(S)), the
 * Merkle tree over the n leaves. `leaves` receives the n leaves (32 bytes each). */
int orc_commit_masked(uint32_t log_k, uint32_t log_r, uint64_t P, uint64_t block,
                      const uint8_t id[32], const uint64_t *segment_rows, uint32_t segments,
                      const uint64_t *rows, const uint64_t *masks, uint64_t *codewords,
                      uint8_t *leaves, uint8_t *nodes, int threads) {
    if (block == 0 || log_r == 0 || log_k + log_r > 32 || P == 0 || P > (1ULL << log_k) ||
        segments == 0)
        return -1;
    uint64_t n = 1ULL << (log_k + log_r), count = 0;
    for (uint32_t s = 0; s < segments; s++) count += segment_rows[s];
    if (orc_encode_rows(log_k, log_r, P, rows, masks, count, codewords, threads)) return -1;
    uint8_t *digests = malloc(32 * n * (size_t)segments);
    uint64_t off = 0;
    for (uint32_t s = 0; s < segments; s++) {
        uint8_t *chains = digests + 32 * n * (size_t)s;
        orc_column_init(id, n, chains);
        for (uint64_t r = 0; r < segment_rows[s]; r += block) {
            uint64_t c = segment_rows[s] - r < block ? segment_rows[s] - r : block;
            orc_column_absorb(chains, n, codewords + (off + r) * n, c, threads);
        }
        off += segment_rows[s];
    }
    for (uint64_t j = 0; j < n; j++) {
        orc_hasher h;
        orc_hasher_init(&h, ORC_DOM_COLUMN_MERGE);
        for (uint32_t s = 0; s < segments; s++) orc_hasher_update(&h, digests + 32 * (n * s + j), 32);
        orc_hasher_finalize(&h, leaves + 32 * j);
    }
    free(digests);
    return orc_merkle_tree(leaves, n, nodes, threads);
}
