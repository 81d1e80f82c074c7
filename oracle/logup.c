/*
 * Test infrastructure (oracle): LogUp-GKR fraction trees and their level-by-level sumchecks.
 * Restates crates/tessera-logup/src/prover.rs:201-469 (compress, tree, lookup_tree, table_tree,
 * value_at, prove_logup), crates/tessera-logup/src/instance.rs:225-253 (multiplicities) and
 * crates/tessera-sumcheck/src/trees.rs:199-345,761-930 (FractionTree, prove_trees_in) with the
 * level prover of crates/tessera-sumcheck/src/level.rs:535-590, whose messages equal the batch
 * prover's over [eq(r_i, .) public, N0, N1, D0, D1] with the relation
 * eq (N0 D1 + N1 D0 + lambda D0 D1), degree 3, padding rounds from the claims.
 * Product forests are not restated here.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "internal.h"

typedef orc_fp2 F2;
static F2 f2(uint64_t a, uint64_t b) { F2 r = {a, b}; return r; }
static F2 add2(F2 a, F2 b) { return orc_fp2_add(a, b); }
static F2 sub2(F2 a, F2 b) { return orc_fp2_sub(a, b); }
static F2 mul2(F2 a, F2 b) { return orc_fp2_mul(a, b); }

static uint32_t ceil_log2(uint64_t x) {
    uint32_t h = 0;
    while ((1ULL << h) < x) h++;
    return h;
}

/* multiplicities (instance.rs:225-253): counts per row over 2^{ceil log2 T} leaves, row 0 also
 * counting the lookup tree's 2^{h_a} - N padding leaves. */
void orc_multiplicities(const orc_logup_instance *inst, uint64_t *out) {
    uint64_t leaves = 1ULL << ceil_log2(inst->rows);
    memset(out, 0, 8 * leaves);
    for (uint64_t i = 0; i < inst->lookups; i++) out[inst->lookup_rows[i]]++;
    out[0] += (1ULL << ceil_log2(inst->lookups)) - inst->lookups;
}

/* One fraction tree: levels[t] holds 2^t nodes (N, D) for t = 0..h, in Fp2. */
typedef struct {
    uint32_t h;
    F2 **N, **D;
    /* the current claim */
    uint32_t plen;
    F2 *point;
    F2 num, den;
    uint32_t join;
    int done; /* proven down to its leaves */
} tree_t;

static void build_tree(tree_t *t, uint32_t h, const F2 *leafN, const F2 *leafD) {
    t->h = h;
    t->N = calloc(h + 1, sizeof(F2 *));
    t->D = calloc(h + 1, sizeof(F2 *));
    uint64_t n = 1ULL << h;
    t->N[h] = malloc(sizeof(F2) * n);
    t->D[h] = malloc(sizeof(F2) * n);
    memcpy(t->N[h], leafN, sizeof(F2) * n);
    memcpy(t->D[h], leafD, sizeof(F2) * n);
    for (int l = (int)h - 1; l >= 0; l--) {
        uint64_t m = 1ULL << l;
        t->N[l] = malloc(sizeof(F2) * m);
        t->D[l] = malloc(sizeof(F2) * m);
        for (uint64_t x = 0; x < m; x++) {
            F2 nl = t->N[l + 1][2 * x], nr = t->N[l + 1][2 * x + 1];
            F2 dl = t->D[l + 1][2 * x], dr = t->D[l + 1][2 * x + 1];
            /* (n_L d_R + n_R d_L, d_L d_R) (trees.rs:225-260) */
            t->N[l][x] = add2(mul2(nl, dr), mul2(nr, dl));
            t->D[l][x] = mul2(dl, dr);
        }
    }
    t->num = t->N[0][0];
    t->den = t->D[0][0];
    t->plen = 0;
    t->point = NULL;
    t->done = 0;
}

static void free_tree(tree_t *t) {
    for (uint32_t l = 0; l <= t->h; l++) {
        free(t->N[l]);
        free(t->D[l]);
    }
    free(t->N);
    free(t->D);
    free(t->point);
}

/* compress: key + beta value, or the key alone (prover.rs:201-207). */
static F2 compress_pair(F2 beta, uint64_t key, const uint64_t *value) {
    if (!value) return f2(key, 0);
    return add2(orc_fp2_mul_base(beta, *value), f2(key, 0));
}

static void log_put(uint64_t *log, uint64_t cap, uint64_t *len, uint64_t v) {
    if (*len < cap) log[*len] = v;
    (*len)++;
}
static void log_f2(uint64_t *log, uint64_t cap, uint64_t *len, F2 v) {
    log_put(log, cap, len, v.c0);
    log_put(log, cap, len, v.c1);
}

/* One level's sumcheck over the active trees (level.rs:535-590 semantics), then rho and the next
 * claims (trees.rs:859-905). Writes the level record into the log. */
static void prove_level(orc_transcript *tr, uint32_t level, tree_t **act, uint32_t count,
                        uint32_t *lvl_of, uint64_t *log, uint64_t cap, uint64_t *len) {
    char label[128];
    snprintf(label, sizeof label, "tree/%u/gamma", level);
    F2 gamma = orc_transcript_challenge_fp2(tr, label, NULL);
    snprintf(label, sizeof label, "tree/%u/lambda", level);
    F2 lambda = orc_transcript_challenge_fp2(tr, label, NULL);
    uint32_t n_max = 0;
    for (uint32_t i = 0; i < count; i++)
        if (act[i]->plen > n_max) n_max = act[i]->plen;
    /* working tables per instance: eq, N0, N1, D0, D1 of 2^{plen} entries */
    F2 *(*tab)[5] = calloc(count, sizeof *tab);
    F2 *w = malloc(sizeof(F2) * count), *c = malloc(sizeof(F2) * count);
    F2 g = f2(1, 0);
    for (uint32_t i = 0; i < count; i++) {
        tree_t *t = act[i];
        uint32_t n = t->plen, below = lvl_of[i] + 1;
        uint64_t m = 1ULL << n;
        for (int k = 0; k < 5; k++) tab[i][k] = malloc(sizeof(F2) * m);
        orc_eq_table((const uint64_t *)t->point, n, (uint64_t *)tab[i][0]);
        for (uint64_t x = 0; x < m; x++) {
            tab[i][1][x] = t->N[below][2 * x];
            tab[i][2][x] = t->N[below][2 * x + 1];
            tab[i][3][x] = t->D[below][2 * x];
            tab[i][4][x] = t->D[below][2 * x + 1];
        }
        w[i] = g;
        c[i] = add2(t->num, mul2(lambda, t->den));
        g = mul2(g, gamma);
    }
    log_put(log, cap, len, n_max);
    log_f2(log, cap, len, gamma);
    log_f2(log, cap, len, lambda);
    F2 *chal = malloc(sizeof(F2) * (n_max ? n_max : 1));
    for (uint32_t j = 0; j < n_max; j++) {
        F2 msg[3] = {f2(0, 0), f2(0, 0), f2(0, 0)};
        for (uint32_t i = 0; i < count; i++) {
            uint32_t pad = n_max - act[i]->plen;
            if (j < pad) {
                F2 v = orc_fp2_mul_base(mul2(w[i], c[i]), 1ULL << (pad - j - 1));
                for (int z = 0; z < 3; z++) msg[z] = add2(msg[z], v);
                continue;
            }
            uint64_t pairs = 1ULL << (act[i]->plen - (j - pad) - 1);
            F2 s[3] = {f2(0, 0), f2(0, 0), f2(0, 0)};
            for (uint64_t q = 0; q < pairs; q++) {
                F2 at[5][4]; /* tables at z = 0, 2, 3 (slots 0..2) */
                for (int k = 0; k < 5; k++) {
                    F2 a = tab[i][k][2 * q], b = tab[i][k][2 * q + 1], d = sub2(b, a);
                    at[k][0] = a;
                    at[k][1] = add2(b, d);
                    at[k][2] = add2(at[k][1], d);
                }
                for (int z = 0; z < 3; z++) {
                    F2 inner = add2(add2(mul2(at[1][z], at[4][z]), mul2(at[2][z], at[3][z])),
                                    mul2(lambda, mul2(at[3][z], at[4][z])));
                    s[z] = add2(s[z], mul2(at[0][z], inner));
                }
            }
            for (int z = 0; z < 3; z++) msg[z] = add2(msg[z], mul2(w[i], s[z]));
        }
        snprintf(label, sizeof label, "tree/%u/round/%u", level, j);
        orc_transcript_absorb_fp2s(tr, label, (const uint64_t *)msg, 3);
        snprintf(label, sizeof label, "tree/%u/r/%u", level, j);
        F2 r = orc_transcript_challenge_fp2(tr, label, NULL);
        chal[j] = r;
        for (int z = 0; z < 3; z++) log_f2(log, cap, len, msg[z]);
        for (uint32_t i = 0; i < count; i++) {
            uint32_t pad = n_max - act[i]->plen;
            if (j < pad) continue;
            uint64_t half = 1ULL << (act[i]->plen - (j - pad) - 1);
            for (int k = 0; k < 5; k++)
                for (uint64_t x = 0; x < half; x++) {
                    F2 a = tab[i][k][2 * x], b = tab[i][k][2 * x + 1];
                    tab[i][k][x] = add2(a, mul2(r, sub2(b, a)));
                }
        }
    }
    for (uint32_t j = 0; j < n_max; j++) log_f2(log, cap, len, chal[j]);
    /* final evaluations: N0, N1, D0, D1 per instance */
    F2 *fin = malloc(sizeof(F2) * 4 * (count ? count : 1));
    for (uint32_t i = 0; i < count; i++)
        for (int k = 0; k < 4; k++) fin[4 * i + k] = tab[i][k + 1][0];
    snprintf(label, sizeof label, "tree/%u/evals", level);
    orc_transcript_absorb_fp2s(tr, label, (const uint64_t *)fin, 4 * count);
    log_put(log, cap, len, 4 * count);
    for (uint32_t i = 0; i < 4 * count; i++) log_f2(log, cap, len, fin[i]);
    snprintf(label, sizeof label, "tree/%u/line", level);
    F2 rho = orc_transcript_challenge_fp2(tr, label, NULL);
    log_f2(log, cap, len, rho);
    for (int b = 0; b < 4; b++) log_put(log, cap, len, orc_get_u64(tr->state + 8 * b));
    log_put(log, cap, len, tr->counter);
    /* next claims: point rho || own point (the last plen challenges) */
    for (uint32_t i = 0; i < count; i++) {
        tree_t *t = act[i];
        uint32_t n = t->plen;
        F2 *np = malloc(sizeof(F2) * (n + 1));
        np[0] = rho;
        for (uint32_t k = 0; k < n; k++) np[1 + k] = chal[n_max - n + k];
        free(t->point);
        t->point = np;
        t->plen = n + 1;
        F2 n0 = fin[4 * i], n1 = fin[4 * i + 1], d0 = fin[4 * i + 2], d1 = fin[4 * i + 3];
        t->num = add2(n0, mul2(rho, sub2(n1, n0)));
        t->den = add2(d0, mul2(rho, sub2(d1, d0)));
        for (int k = 0; k < 5; k++) free(tab[i][k]);
    }
    free(tab);
    free(w);
    free(c);
    free(chal);
    free(fin);
}

int orc_prove_logup(orc_transcript *tr, const orc_logup_instance *inst, uint32_t count,
                    const uint32_t *schedule, uint64_t *alpha_out, uint64_t *beta_out,
                    uint64_t *roots, uint64_t *values, uint64_t *n_values, uint64_t *claims,
                    uint64_t *log, uint64_t cap, uint64_t *log_len) {
    F2 alpha = orc_transcript_challenge_fp2(tr, "logup/alpha", NULL);
    F2 beta = orc_transcript_challenge_fp2(tr, "logup/beta", NULL);
    alpha_out[0] = alpha.c0;
    alpha_out[1] = alpha.c1;
    beta_out[0] = beta.c0;
    beta_out[1] = beta.c1;
    tree_t *trees = calloc(2 * count + 1, sizeof(tree_t));
    for (uint32_t k = 0; k < count; k++) {
        const orc_logup_instance *in = &inst[k];
        uint32_t ha = ceil_log2(in->lookups), ht = ceil_log2(in->rows);
        uint64_t na = 1ULL << ha, nt = 1ULL << ht;
        const uint64_t *v0 = in->values ? &in->values[0] : NULL;
        F2 pad = sub2(alpha, compress_pair(beta, in->keys[0], v0)); /* prover.rs:247-256 */
        F2 *N = malloc(sizeof(F2) * (na > nt ? na : nt)), *D = malloc(sizeof(F2) * (na > nt ? na : nt));
        for (uint64_t x = 0; x < na; x++) {
            N[x] = f2(1, 0);
            D[x] = x < in->lookups
                       ? sub2(alpha, compress_pair(beta, in->lookup_keys[x],
                                                   in->lookup_values ? &in->lookup_values[x] : NULL))
                       : pad;
        }
        build_tree(&trees[2 * k], ha, N, D);
        for (uint64_t x = 0; x < nt; x++) {
            N[x] = f2(orc_fp_sub(0, in->multiplicities[x]), 0);
            D[x] = x < in->rows
                       ? sub2(alpha, compress_pair(beta, in->keys[x], in->values ? &in->values[x] : NULL))
                       : pad;
        }
        build_tree(&trees[2 * k + 1], ht, N, D);
        free(N);
        free(D);
        F2 r4[4] = {trees[2 * k].num, trees[2 * k].den, trees[2 * k + 1].num, trees[2 * k + 1].den};
        memcpy(roots + 8 * k, r4, sizeof r4);
    }
    orc_transcript_absorb_fp2s(tr, "logup/roots", roots, 4 * count);
    /* schedule: entry s = tree id (2k = lookup k, 2k+1 = table k) */
    uint32_t ntrees = 2 * count, tallest = 0;
    for (uint32_t s = 0; s < ntrees; s++)
        if (trees[schedule[s]].h > tallest) tallest = trees[schedule[s]].h;
    for (uint32_t s = 0; s < ntrees; s++) trees[schedule[s]].join = tallest - trees[schedule[s]].h;
    *log_len = 0;
    tree_t **act = malloc(sizeof(tree_t *) * (ntrees ? ntrees : 1));
    uint32_t *lvl = malloc(sizeof(uint32_t) * (ntrees ? ntrees : 1));
    for (uint32_t level = 0; level < tallest; level++) {
        uint32_t na = 0;
        for (uint32_t s = 0; s < ntrees; s++) {
            tree_t *t = &trees[schedule[s]];
            if (t->join <= level) {
                act[na] = t;
                lvl[na] = level - t->join;
                na++;
            }
        }
        prove_level(tr, level, act, na, lvl, log, cap, log_len);
    }
    /* values: valũe(rho_a) of the padded lookup value column (prover.rs:301-318) */
    uint64_t nv = 0;
    for (uint32_t k = 0; k < count; k++) {
        const orc_logup_instance *in = &inst[k];
        tree_t *la = &trees[2 * k], *ta = &trees[2 * k + 1];
        if (in->values) {
            uint64_t n = 1ULL << la->h;
            uint64_t *col = malloc(8 * n);
            for (uint64_t x = 0; x < n; x++) col[x] = x < in->lookups ? in->lookup_values[x] : in->values[0];
            uint64_t v[2];
            orc_mle_eval(col, 0, la->h, (const uint64_t *)la->point, v);
            values[2 * nv] = v[0];
            values[2 * nv + 1] = v[1];
            nv++;
            free(col);
        }
        /* leaf claims: lookup (num, den), table (num, den) */
        F2 cl[4] = {la->num, la->den, ta->num, ta->den};
        memcpy(claims + 8 * k, cl, sizeof cl);
    }
    orc_transcript_absorb_fp2s(tr, "logup/values", values, nv);
    *n_values = nv;
    for (uint32_t s = 0; s < ntrees; s++) free_tree(&trees[s]);
    free(trees);
    free(act);
    free(lvl);
    return 0;
}
