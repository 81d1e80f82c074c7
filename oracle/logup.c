/*
 * Test infrastructure (oracle): LogUp-GKR fraction trees and their level-by-level sumchecks.
 * Restates crates/tessera-logup/src/prover.rs:201-469 (compress, tree, lookup_tree, table_tree,
 * value_at, prove_logup), crates/tessera-logup/src/instance.rs:225-253 (multiplicities) and
 * crates/tessera-sumcheck/src/trees.rs:199-345,761-930 (FractionTree, prove_trees_in) with the
 * level prover of crates/tessera-sumcheck/src/level.rs:535-590, whose messages equal the batch
 * prover's over [eq(r_i, .) public, N0, N1, D0, D1] with the relation
 * eq (N0 D1 + N1 D0 + lambda D0 D1), degree 3, padding rounds from the claims.
 * Product forests (crates/tessera-sumcheck/src/forest.rs:30-340) join the same level loop with
 * the relation eq V0 V1 and points from the levels' root challenges.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "internal.h"

typedef orc_fp2 F2;
static F2 f2(uint64_t a, uint64_t b) { F2 r = {a, b}; return r; }
static F2 add2(F2 a, F2 b) { return fp2_add(a, b); }
static F2 sub2(F2 a, F2 b) { return fp2_sub(a, b); }
static F2 mul2(F2 a, F2 b) { return fp2_mul(a, b); }

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

/* A tree of the level loop (trees.rs:761-930): a fraction tree (levels[t] holds 2^t nodes
 * (N, D)) or a product forest of 2^s stacked trees of height h (V[t] holds 2^(s+t) values, V[h]
 * the leaves, V[t][z] = V[t+1][2z] V[t+1][2z+1]; forest.rs:306-340). */
typedef struct {
    int forest;
    uint32_t h, s;
    F2 **N, **D, **V;
    /* the current claim: point, and (num, den) or the product value in num */
    uint32_t plen;
    F2 *point;
    F2 num, den;
    int has_claim;
    uint32_t join;
} tree_t;

static void build_tree(tree_t *t, uint32_t h, const F2 *leafN, const F2 *leafD) {
    memset(t, 0, sizeof *t);
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
    t->has_claim = 1;
}

static void build_forest(tree_t *t, const orc_forest *f) {
    memset(t, 0, sizeof *t);
    uint32_t vars = 0;
    while ((1ULL << vars) < f->len) vars++;
    t->forest = 1;
    t->h = f->height;
    t->s = vars - f->height;
    t->V = calloc(t->h + 1, sizeof(F2 *));
    uint64_t n = f->len;
    t->V[t->h] = malloc(sizeof(F2) * n);
    for (uint64_t x = 0; x < n; x++) t->V[t->h][x] = f2(f->leaves[x], 0);
    for (int l = (int)t->h - 1; l >= 0; l--) {
        uint64_t m = 1ULL << (t->s + l);
        t->V[l] = malloc(sizeof(F2) * m);
        for (uint64_t z = 0; z < m; z++) t->V[l][z] = mul2(t->V[l + 1][2 * z], t->V[l + 1][2 * z + 1]);
    }
    t->has_claim = 0;
}

static void free_tree(tree_t *t) {
    for (uint32_t l = 0; l <= t->h; l++) {
        if (t->N) free(t->N[l]);
        if (t->D) free(t->D[l]);
        if (t->V) free(t->V[l]);
    }
    free(t->N);
    free(t->D);
    free(t->V);
    free(t->point);
}

/* compress: key + beta value, or the key alone (prover.rs:201-207). */
static F2 compress_pair(F2 beta, uint64_t key, const uint64_t *value) {
    if (!value) return f2(key, 0);
    return add2(fp2_mul_base(beta, *value), f2(key, 0));
}

static void log_put(uint64_t *log, uint64_t cap, uint64_t *len, uint64_t v) {
    if (*len < cap) log[*len] = v;
    (*len)++;
}
static void log_f2(uint64_t *log, uint64_t cap, uint64_t *len, F2 v) {
    log_put(log, cap, len, v.c0);
    log_put(log, cap, len, v.c1);
}

/* One level of prove_trees_in (trees.rs:835-905) over the scheduled trees `sched` (n of them),
 * with the level prover's messages (level.rs:535-590), which equal the batch prover's over
 * [eq(r_i, .) public, tables] with the relations eq (N0 D1 + N1 D0 + lambda D0 D1) for fraction
 * trees and eq V0 V1 for forests, degree 3, padding rounds from the claims. The level record in
 * the log: s_max, the root challenges, has_lambda, n_max, gamma, lambda, the rounds (n_max x 3),
 * the challenges, the count of final evaluations and the evaluations, rho, the state, counter. */
static void prove_level(orc_transcript *tr, uint32_t level, tree_t **sched, uint32_t n,
                        uint64_t *log, uint64_t cap, uint64_t *len) {
    char label[128];
    /* forests joining here take their points from the level's root challenges */
    uint32_t s_max = 0;
    for (uint32_t i = 0; i < n; i++)
        if (sched[i]->join == level && sched[i]->s > s_max) s_max = sched[i]->s;
    F2 *rootc = malloc(sizeof(F2) * (s_max ? s_max : 1));
    for (uint32_t c = 0; c < s_max; c++) {
        snprintf(label, sizeof label, "tree/%u/root/%u", level, c);
        rootc[c] = orc_transcript_challenge_fp2(tr, label, NULL);
    }
    for (uint32_t i = 0; i < n; i++) {
        tree_t *t = sched[i];
        if (t->join != level) continue;
        free(t->point);
        t->plen = t->s;
        t->point = malloc(sizeof(F2) * (t->s ? t->s : 1));
        for (uint32_t c = 0; c < t->s; c++) t->point[c] = rootc[s_max - t->s + c];
    }
    tree_t **act = malloc(sizeof(tree_t *) * (n ? n : 1));
    uint32_t count = 0, fractions = 0;
    for (uint32_t i = 0; i < n; i++)
        if (sched[i]->join <= level) {
            act[count++] = sched[i];
            fractions |= !sched[i]->forest;
        }
    snprintf(label, sizeof label, "tree/%u/gamma", level);
    F2 gamma = orc_transcript_challenge_fp2(tr, label, NULL);
    F2 lambda = f2(0, 0);
    if (fractions) {
        snprintf(label, sizeof label, "tree/%u/lambda", level);
        lambda = orc_transcript_challenge_fp2(tr, label, NULL);
    }
    uint32_t n_max = 0;
    for (uint32_t i = 0; i < count; i++)
        if (act[i]->plen > n_max) n_max = act[i]->plen;
    /* working tables per instance: eq, then N0, N1, D0, D1 (fraction) or V0, V1 (forest) */
    F2 *(*tab)[5] = calloc(count, sizeof *tab);
    F2 *w = malloc(sizeof(F2) * (count ? count : 1)), *c = malloc(sizeof(F2) * (count ? count : 1));
    F2 g = f2(1, 0);
    for (uint32_t i = 0; i < count; i++) {
        tree_t *t = act[i];
        uint32_t nv = t->plen, below = level - t->join + 1;
        uint64_t m = 1ULL << nv;
        int nt = t->forest ? 3 : 5;
        for (int k = 0; k < nt; k++) tab[i][k] = malloc(sizeof(F2) * m);
        orc_eq_table((const uint64_t *)t->point, nv, (uint64_t *)tab[i][0]);
        for (uint64_t x = 0; x < m; x++) {
            if (t->forest) {
                tab[i][1][x] = t->V[below][2 * x];
                tab[i][2][x] = t->V[below][2 * x + 1];
            } else {
                tab[i][1][x] = t->N[below][2 * x];
                tab[i][2][x] = t->N[below][2 * x + 1];
                tab[i][3][x] = t->D[below][2 * x];
                tab[i][4][x] = t->D[below][2 * x + 1];
            }
        }
        w[i] = g;
        if (t->forest) {
            if (t->has_claim) {
                c[i] = t->num;
            } else { /* the forest's root claim: Sum_z eq(point, z) V0[z] V1[z] (trees.rs:638-640) */
                F2 acc = f2(0, 0);
                for (uint64_t x = 0; x < m; x++) acc = add2(acc, mul2(tab[i][0][x], mul2(tab[i][1][x], tab[i][2][x])));
                c[i] = acc;
            }
        } else {
            c[i] = add2(t->num, mul2(lambda, t->den));
        }
        g = mul2(g, gamma);
    }
    log_put(log, cap, len, s_max);
    for (uint32_t k = 0; k < s_max; k++) log_f2(log, cap, len, rootc[k]);
    log_put(log, cap, len, fractions ? 1 : 0);
    log_put(log, cap, len, n_max);
    log_f2(log, cap, len, gamma);
    log_f2(log, cap, len, lambda);
    F2 *chal = malloc(sizeof(F2) * (n_max ? n_max : 1));
    for (uint32_t j = 0; j < n_max; j++) {
        F2 msg[3] = {f2(0, 0), f2(0, 0), f2(0, 0)};
        for (uint32_t i = 0; i < count; i++) {
            uint32_t pad = n_max - act[i]->plen;
            if (j < pad) {
                F2 v = fp2_mul_base(mul2(w[i], c[i]), 1ULL << (pad - j - 1));
                for (int z = 0; z < 3; z++) msg[z] = add2(msg[z], v);
                continue;
            }
            int nt = act[i]->forest ? 3 : 5;
            uint64_t pairs = 1ULL << (act[i]->plen - (j - pad) - 1);
            F2 s[3] = {f2(0, 0), f2(0, 0), f2(0, 0)};
            for (uint64_t q = 0; q < pairs; q++) {
                F2 at[5][3]; /* tables at z = 0, 2, 3 (slots 0..2) */
                for (int k = 0; k < nt; k++) {
                    F2 a = tab[i][k][2 * q], b = tab[i][k][2 * q + 1], d = sub2(b, a);
                    at[k][0] = a;
                    at[k][1] = add2(b, d);
                    at[k][2] = add2(at[k][1], d);
                }
                for (int z = 0; z < 3; z++) {
                    F2 inner = act[i]->forest
                                   ? mul2(at[1][z], at[2][z])
                                   : add2(add2(mul2(at[1][z], at[4][z]), mul2(at[2][z], at[3][z])),
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
            int nt = act[i]->forest ? 3 : 5;
            uint64_t half = 1ULL << (act[i]->plen - (j - pad) - 1);
            for (int k = 0; k < nt; k++)
                for (uint64_t x = 0; x < half; x++) {
                    F2 a = tab[i][k][2 * x], b = tab[i][k][2 * x + 1];
                    tab[i][k][x] = add2(a, mul2(r, sub2(b, a)));
                }
        }
    }
    for (uint32_t j = 0; j < n_max; j++) log_f2(log, cap, len, chal[j]);
    /* final evaluations: N0, N1, D0, D1 per fraction tree, V0, V1 per forest */
    F2 *fin = malloc(sizeof(F2) * 4 * (count ? count : 1));
    uint32_t nf = 0;
    for (uint32_t i = 0; i < count; i++) {
        int ne = act[i]->forest ? 2 : 4;
        for (int k = 0; k < ne; k++) fin[nf++] = tab[i][k + 1][0];
    }
    snprintf(label, sizeof label, "tree/%u/evals", level);
    orc_transcript_absorb_fp2s(tr, label, (const uint64_t *)fin, nf);
    log_put(log, cap, len, nf);
    for (uint32_t i = 0; i < nf; i++) log_f2(log, cap, len, fin[i]);
    snprintf(label, sizeof label, "tree/%u/line", level);
    F2 rho = orc_transcript_challenge_fp2(tr, label, NULL);
    log_f2(log, cap, len, rho);
    for (int b = 0; b < 4; b++) log_put(log, cap, len, orc_get_u64(tr->state + 8 * b));
    log_put(log, cap, len, tr->counter);
    /* next claims: point rho || own point (the last plen challenges), values on the line */
    uint32_t at = 0;
    for (uint32_t i = 0; i < count; i++) {
        tree_t *t = act[i];
        uint32_t nv = t->plen;
        F2 *np = malloc(sizeof(F2) * (nv + 1));
        np[0] = rho;
        for (uint32_t k = 0; k < nv; k++) np[1 + k] = chal[n_max - nv + k];
        free(t->point);
        t->point = np;
        t->plen = nv + 1;
        if (t->forest) {
            F2 v0 = fin[at], v1 = fin[at + 1];
            t->num = add2(v0, mul2(rho, sub2(v1, v0)));
            at += 2;
        } else {
            F2 n0 = fin[at], n1 = fin[at + 1], d0 = fin[at + 2], d1 = fin[at + 3];
            t->num = add2(n0, mul2(rho, sub2(n1, n0)));
            t->den = add2(d0, mul2(rho, sub2(d1, d0)));
            at += 4;
        }
        t->has_claim = 1;
        for (int k = 0; k < (t->forest ? 3 : 5); k++) free(tab[i][k]);
    }
    free(tab);
    free(w);
    free(c);
    free(chal);
    free(fin);
    free(rootc);
    free(act);
}

/* prove_trees_in's level loop over scheduled trees: every tree joins at tallest - height. */
static void prove_scheduled(orc_transcript *tr, tree_t **sched, uint32_t n, uint64_t *log,
                            uint64_t cap, uint64_t *log_len) {
    uint32_t tallest = 0;
    for (uint32_t i = 0; i < n; i++)
        if (sched[i]->h > tallest) tallest = sched[i]->h;
    for (uint32_t i = 0; i < n; i++) sched[i]->join = tallest - sched[i]->h;
    *log_len = 0;
    for (uint32_t level = 0; level < tallest; level++) prove_level(tr, level, sched, n, log, cap, log_len);
}

/* prove_trees (trees.rs:761-776) over product forests alone, in order, under tree/{l}. Shapes
 * (forest.rs:269-286): a leaf count that is a power of two, 1 <= height <= log2 of it; -1 else.
 * forest_claims: the forests' leaf claims (value per forest, Fp2); their points are the
 * challenges of the log. */
int orc_prove_trees(orc_transcript *tr, const orc_forest *forests, uint32_t nforests,
                    uint64_t *forest_claims, uint64_t *log, uint64_t cap, uint64_t *log_len) {
    for (uint32_t f = 0; f < nforests; f++) {
        uint64_t len = forests[f].len;
        if (len == 0 || (len & (len - 1)) || forests[f].height == 0 || (1ULL << forests[f].height) > len) return -1;
    }
    tree_t *trees = calloc(nforests ? nforests : 1, sizeof(tree_t));
    tree_t **sched = malloc(sizeof(tree_t *) * (nforests ? nforests : 1));
    for (uint32_t f = 0; f < nforests; f++) {
        build_forest(&trees[f], &forests[f]);
        sched[f] = &trees[f];
    }
    prove_scheduled(tr, sched, nforests, log, cap, log_len);
    for (uint32_t f = 0; f < nforests; f++) {
        forest_claims[2 * f] = trees[f].num.c0;
        forest_claims[2 * f + 1] = trees[f].num.c1;
        free_tree(&trees[f]);
    }
    free(trees);
    free(sched);
    return 0;
}

/* prove_logup (prover.rs:361-469) with product forests beside the instances: schedule entries
 * are tree ids, 2k = lookup tree k, 2k+1 = table tree k, 2 count + f = forest f. forest_claims
 * receives each forest's leaf claim value (may be NULL without forests). */
int orc_prove_logup(orc_transcript *tr, const orc_logup_instance *inst, uint32_t count,
                    const orc_forest *forests, uint32_t nforests,
                    const uint32_t *schedule, uint64_t *alpha_out, uint64_t *beta_out,
                    uint64_t *roots, uint64_t *values, uint64_t *n_values, uint64_t *claims,
                    uint64_t *forest_claims, uint64_t *log, uint64_t cap, uint64_t *log_len) {
    F2 alpha = orc_transcript_challenge_fp2(tr, "logup/alpha", NULL);
    F2 beta = orc_transcript_challenge_fp2(tr, "logup/beta", NULL);
    alpha_out[0] = alpha.c0;
    alpha_out[1] = alpha.c1;
    beta_out[0] = beta.c0;
    beta_out[1] = beta.c1;
    tree_t *trees = calloc(2 * count + nforests + 1, sizeof(tree_t));
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
            N[x] = f2(fp_sub(0, in->multiplicities[x]), 0);
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
    for (uint32_t f = 0; f < nforests; f++) build_forest(&trees[2 * count + f], &forests[f]);
    uint32_t ntrees = 2 * count + nforests;
    tree_t **sched = malloc(sizeof(tree_t *) * (ntrees ? ntrees : 1));
    for (uint32_t s = 0; s < ntrees; s++) sched[s] = &trees[schedule[s]];
    prove_scheduled(tr, sched, ntrees, log, cap, log_len);
    free(sched);
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
    for (uint32_t f = 0; f < nforests; f++) {
        forest_claims[2 * f] = trees[2 * count + f].num.c0;
        forest_claims[2 * f + 1] = trees[2 * count + f].num.c1;
    }
    for (uint32_t s = 0; s < ntrees; s++) free_tree(&trees[s]);
    free(trees);
    return 0;
}
