/*
 * Test infrastructure (oracle): multilinear tables and the batched product sumcheck.
 * Restates crates/tessera-poly/src/{eq.rs,multilinear.rs,round.rs} and
 * crates/tessera-sumcheck/src/prover.rs (prove_batch -> prove_checked, State::{new, add_round,
 * final_evals}, Bindings::bind, padding_scale) with the kernels of kernel.rs: round j's message
 * is p_j(0), p_j(2), ..., p_j(d) of sum_i gamma^i P_i (P_i front-padded), absorbed as TCE
 * Vec<Fp2>; the challenge binds the lowest variable, t[k] = t[2k] + r (t[2k+1] - t[2k]).
 * A sum-of-products instance (prover.rs:243-312, kernel.rs:502-633 term_sums) has the relation
 * sum_t w_t prod_{f in t} T_f over its table list; the reference groups terms by weight and
 * factors out common tables, which changes no value in the field.
 */
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#include "oracle.h"
#include "internal.h"

typedef orc_fp2 F2;

static F2 f2(uint64_t a, uint64_t b) { F2 r = {a, b}; return r; }
static F2 add2(F2 a, F2 b) { return fp2_add(a, b); }
static F2 sub2(F2 a, F2 b) { return fp2_sub(a, b); }
static F2 mul2(F2 a, F2 b) { return fp2_mul(a, b); }

/* eq_table (tessera-poly eq.rs:23-45): doubling per coordinate, x_1 the lowest bit. */
void orc_eq_table(const uint64_t *point, uint32_t n, uint64_t *out) {
    F2 *t = (F2 *)out;
    t[0] = f2(1, 0);
    uint64_t half = 1;
    for (uint32_t i = 0; i < n; i++) {
        F2 r = f2(point[2 * i], point[2 * i + 1]);
        for (uint64_t k = 0; k < half; k++) {
            F2 one = mul2(t[k], r);
            t[half + k] = one;
            t[k] = sub2(t[k], one);
        }
        half *= 2;
    }
}

/* evaluate / evaluate_base by successive binding of the lowest variable
 * (multilinear.rs:70-121,136-180; equal in the field to the reference's blocked form). */
void orc_mle_eval(const uint64_t *table, int ext, uint32_t n, const uint64_t *point,
                  uint64_t out[2]) {
    uint64_t len = 1ULL << n;
    F2 *t = malloc(sizeof(F2) * len);
    for (uint64_t i = 0; i < len; i++) t[i] = ext ? f2(table[2 * i], table[2 * i + 1]) : f2(table[i], 0);
    for (uint32_t v = 0; v < n; v++) {
        F2 r = f2(point[2 * v], point[2 * v + 1]);
        len /= 2;
        for (uint64_t k = 0; k < len; k++) t[k] = add2(t[2 * k], mul2(r, sub2(t[2 * k + 1], t[2 * k])));
    }
    out[0] = t[0].c0;
    out[1] = t[0].c1;
    free(t);
}

/* RoundPoly::evaluate (round.rs:91-151): Lagrange over the nodes 0..d with p(1) = claim - p(0). */
void orc_round_eval(const uint64_t *evals, uint32_t d, const uint64_t claim[2], const uint64_t r[2],
                    uint64_t out[2]) {
    F2 at[16];
    at[0] = f2(evals[0], evals[1]);
    at[1] = sub2(f2(claim[0], claim[1]), at[0]);
    for (uint32_t k = 2; k <= d; k++) at[k] = f2(evals[2 * (k - 1)], evals[2 * (k - 1) + 1]);
    F2 rr = f2(r[0], r[1]), value = f2(0, 0);
    for (uint32_t k = 0; k <= d; k++) {
        /* basis_k(r) = prod_{m != k} (r - m) / (k - m) */
        F2 num = f2(1, 0);
        uint64_t den = 1;
        for (uint32_t m = 0; m <= d; m++) {
            if (m == k) continue;
            num = mul2(num, sub2(rr, f2(m, 0)));
            den = fp_mul(den, fp_sub(k, m));
        }
        value = add2(value, mul2(at[k], fp2_mul_base(num, orc_fp_inv(den))));
    }
    out[0] = value.c0;
    out[1] = value.c1;
}

/* ---- the batched product sumcheck ---- */

typedef struct {
    uint32_t vars, pad, factors;
    const orc_table *tables;
    const orc_term *terms; /* NULL: product */
    uint32_t nterms;
    F2 **cur;      /* current (bound) values per table, NULL until the first own bind */
    F2 weight;     /* gamma^i */
    F2 wsum;       /* gamma^i * S_i for padding rounds */
    F2 first[17];  /* node sums of the first own round, from the pass that gave S_i */
    int have_first;
} inst_t;

static F2 entry(const inst_t *in, uint32_t k, uint64_t x) {
    if (in->cur[k]) return in->cur[k][x];
    const orc_table *t = &in->tables[k];
    return t->ext ? f2(t->data[2 * x], t->data[2 * x + 1]) : f2(t->data[x], 0);
}

typedef struct {
    const inst_t *in;
    uint32_t nodes; /* degree d: values at z in {0, 2, ..., d} plus z = 1 in slot d */
    uint64_t pairs;
    F2 *partial;    /* per thread: d+1 sums */
    int threads;
    int need_one;   /* also sum z = 1 (only S_i needs it; a message sends p(0), p(2..d)) */
} sums_t;

/* Every table still unbound and in the base field: the first round of Table::base inputs, whose
 * line values and node products stay in Fp (kernel.rs single/zip over base slices). */
static int all_base(const inst_t *in) {
    if (in->terms) return 0;
    for (uint32_t k = 0; k < in->factors; k++)
        if (in->cur[k] || in->tables[k].ext) return 0;
    return 1;
}

static void sums_range_base(sums_t *s, F2 *acc, uint64_t lo, uint64_t hi) {
    const inst_t *in = s->in;
    uint32_t d = s->nodes, F = in->factors;
    uint64_t sum[17] = {0};
    for (uint64_t x = lo; x < hi; x++) {
        uint64_t prod[17];
        for (uint32_t k = 0; k < F; k++) {
            const uint64_t *t = in->tables[k].data;
            uint64_t a = t[2 * x], b = t[2 * x + 1], diff = fp_sub(b, a);
            if (k == 0) {
                uint64_t line = fp_add(b, diff);
                prod[0] = a;
                prod[d] = b;
                for (uint32_t z = 2; z <= d; z++) {
                    prod[z - 1] = line;
                    line = fp_add(line, diff);
                }
                continue;
            }
            prod[0] = fp_mul(prod[0], a);
            if (s->need_one) prod[d] = fp_mul(prod[d], b);
            uint64_t line = fp_add(b, diff);
            for (uint32_t z = 2; z <= d; z++) {
                prod[z - 1] = fp_mul(prod[z - 1], line);
                line = fp_add(line, diff);
            }
        }
        for (uint32_t z = 0; z <= d; z++) sum[z] = fp_add(sum[z], prod[z]);
    }
    for (uint32_t z = 0; z <= d; z++) acc[z] = f2(sum[z], 0);
}

/* kernel::product_sums (kernel.rs:292): sum over pairs of prod_k line_k(z). Slot layout:
 * [p(0), p(2), ..., p(d), p(1)]. */
static void sums_range(void *arg, int slot, uint64_t lo, uint64_t hi) {
    sums_t *s = arg;
    const inst_t *in = s->in;
    uint32_t d = s->nodes;
    F2 *acc = s->partial + (uint64_t)slot * (d + 1);
    for (uint32_t z = 0; z <= d; z++) acc[z] = f2(0, 0);
    if (all_base(in)) {
        sums_range_base(s, acc, lo, hi);
        return;
    }
    for (uint64_t x = lo; x < hi; x++) {
        F2 prod[16];
        if (in->terms) {
            /* every table's line at every node, then sum_t w_t prod_{f in t} line_f(z) */
            F2 line[16][16];
            for (uint32_t k = 0; k < in->factors; k++) {
                F2 a = entry(in, k, 2 * x), b = entry(in, k, 2 * x + 1), diff = sub2(b, a);
                line[k][0] = a;
                line[k][d] = b;
                F2 v = add2(b, diff);
                for (uint32_t z = 2; z <= d; z++) {
                    line[k][z - 1] = v;
                    v = add2(v, diff);
                }
            }
            for (uint32_t z = 0; z <= d; z++) {
                F2 sum = f2(0, 0);
                for (uint32_t t = 0; t < in->nterms; t++) {
                    const orc_term *tm = &in->terms[t];
                    F2 p = f2(tm->weight[0], tm->weight[1]);
                    for (uint32_t f = 0; f < tm->nfactors; f++) p = mul2(p, line[tm->factors[f]][z]);
                    sum = add2(sum, p);
                }
                acc[z] = add2(acc[z], sum);
            }
            continue;
        }
        for (uint32_t k = 0; k < in->factors; k++) {
            F2 a = entry(in, k, 2 * x), b = entry(in, k, 2 * x + 1), diff = sub2(b, a);
            /* z = 0 -> a; z = 1 -> b; z >= 2 -> a + z (b - a); the first table seeds the
             * products */
            F2 line = add2(b, diff); /* z = 2 */
            if (k == 0) {
                prod[0] = a;
                prod[d] = b;
                for (uint32_t z = 2; z <= d; z++) {
                    prod[z - 1] = line;
                    line = add2(line, diff);
                }
                continue;
            }
            prod[0] = mul2(prod[0], a);
            if (s->need_one) prod[d] = mul2(prod[d], b);
            for (uint32_t z = 2; z <= d; z++) {
                prod[z - 1] = mul2(prod[z - 1], line);
                line = add2(line, diff);
            }
        }
        for (uint32_t z = 0; z <= d; z++) acc[z] = add2(acc[z], prod[z]);
    }
}

/* All d+1 node sums of an instance's current round: out[0] = p(0), out[1..d-1] = p(2..d),
 * out[d] = p(1). */
static void node_sums(const inst_t *in, uint32_t d, uint64_t pairs, F2 *out, int threads,
                      int need_one) {
    int th = threads;
    if (pairs < (1u << 12)) th = 1; /* prover.rs:502-536: at least 2^12 pairs per thread */
    else if ((uint64_t)th > pairs >> 12) th = (int)(pairs >> 12);
    if (th < 1) th = 1;
    sums_t s = {in, d, pairs, calloc((size_t)th * (d + 1), sizeof(F2)), th, need_one};
    orc_parallel_for_t(pairs, th, sums_range, &s);
    for (uint32_t z = 0; z <= d; z++) out[z] = f2(0, 0);
    for (int t = 0; t < th; t++)
        for (uint32_t z = 0; z <= d; z++) out[z] = add2(out[z], s.partial[t * (d + 1) + z]);
    free(s.partial);
}

typedef struct {
    inst_t *in;
    uint32_t k;
    F2 r;
    F2 *dst;
} bind_t;
static void bind_range(void *arg, uint64_t lo, uint64_t hi) {
    bind_t *b = arg;
    const orc_table *t = &b->in->tables[b->k];
    if (!b->in->cur[b->k] && !t->ext) {
        /* a base table: a + r (c - a) = (a + r0 (c - a), r1 (c - a)) */
        for (uint64_t x = lo; x < hi; x++) {
            uint64_t a = t->data[2 * x], dlt = fp_sub(t->data[2 * x + 1], a);
            b->dst[x] = f2(fp_add(a, fp_mul(b->r.c0, dlt)), fp_mul(b->r.c1, dlt));
        }
        return;
    }
    for (uint64_t x = lo; x < hi; x++) {
        F2 a = entry(b->in, b->k, 2 * x), c = entry(b->in, b->k, 2 * x + 1);
        b->dst[x] = add2(a, mul2(b->r, sub2(c, a)));
    }
}

static void bind_instance(inst_t *in, F2 r, uint64_t half, int threads) {
    for (uint32_t k = 0; k < in->factors; k++) {
        F2 *dst = malloc(sizeof(F2) * (half ? half : 1));
        bind_t b = {in, k, r, dst};
        orc_parallel_for(half, half >= (1u << 14) ? threads : 1, bind_range, &b);
        free(in->cur[k]);
        in->cur[k] = dst;
    }
}

static uint32_t log2_u64(uint64_t x) { return (uint32_t)__builtin_ctzll(x); }

int64_t orc_prove_batch(orc_transcript *t, const char *prefix, uint32_t degree,
                        const orc_instance *instances, uint32_t count, uint64_t *gamma_out,
                        uint64_t *claim_out, uint64_t *rounds, uint64_t *challenges,
                        uint64_t *finals, int threads) {
    if (degree < 1 || degree > 10 || count == 0) return -1;
    char label[256];
    uint32_t n_max = 0;
    inst_t *in = calloc(count, sizeof(inst_t));
    for (uint32_t i = 0; i < count; i++) {
        (void)0;
        in[i].factors = instances[i].factors;
        in[i].tables = instances[i].tables;
        in[i].vars = instances[i].vars;
        in[i].terms = instances[i].terms;
        in[i].nterms = instances[i].nterms;
        in[i].cur = calloc(instances[i].factors ? instances[i].factors : 1, sizeof(F2 *));
        /* check_instance (prover.rs:739-771): NoTables, TooManyTables or NoTerms, TermFactors,
         * TermFactor */
        if (in[i].factors == 0 || in[i].factors > 16) return -1;
        if (!in[i].terms && in[i].factors > degree) return -1;
        if (in[i].terms) {
            if (in[i].nterms == 0) return -1;
            for (uint32_t t = 0; t < in[i].nterms; t++) {
                const orc_term *tm = &in[i].terms[t];
                if (tm->nfactors == 0 || tm->nfactors > degree) return -1;
                for (uint32_t f = 0; f < tm->nfactors; f++)
                    if (tm->factors[f] >= in[i].factors) return -1;
            }
        }
        if (in[i].vars > n_max) n_max = in[i].vars;
    }
    snprintf(label, sizeof label, "%s/gamma", prefix);
    F2 gamma = orc_transcript_challenge_fp2(t, label, NULL);
    F2 w = f2(1, 0), claim = f2(0, 0);
    for (uint32_t i = 0; i < count; i++) {
        in[i].weight = w;
        in[i].pad = n_max - in[i].vars;
        /* S_i, the sum over the own hypercube (State::new, prover.rs:924-980) */
        F2 S;
        if (in[i].vars == 0 && in[i].terms) {
            S = f2(0, 0);
            for (uint32_t t = 0; t < in[i].nterms; t++) {
                const orc_term *tm = &in[i].terms[t];
                F2 p = f2(tm->weight[0], tm->weight[1]);
                for (uint32_t f = 0; f < tm->nfactors; f++) p = mul2(p, entry(&in[i], tm->factors[f], 0));
                S = add2(S, p);
            }
        } else if (in[i].vars == 0) {
            S = f2(1, 0);
            for (uint32_t k = 0; k < in[i].factors; k++) S = mul2(S, entry(&in[i], k, 0));
        } else {
            node_sums(&in[i], degree, 1ULL << (in[i].vars - 1), in[i].first, threads, 1);
            in[i].have_first = 1;
            S = add2(in[i].first[0], in[i].first[degree]);
        }
        in[i].wsum = mul2(w, S);
        claim = add2(claim, fp2_mul_base(in[i].wsum, fp_pow(2, in[i].pad)));
        w = mul2(w, gamma);
    }
    for (uint32_t j = 0; j < n_max; j++) {
        F2 msg[16];
        for (uint32_t z = 0; z < degree; z++) msg[z] = f2(0, 0);
        for (uint32_t i = 0; i < count; i++) {
            inst_t *x = &in[i];
            if (j < x->pad) {
                /* padding round: 2^{pad - j - 1} S_i at every node (prover.rs:1020-1031) */
                F2 v = fp2_mul_base(x->wsum, 1ULL << (x->pad - j - 1));
                for (uint32_t z = 0; z < degree; z++) msg[z] = add2(msg[z], v);
                continue;
            }
            uint64_t pairs = 1ULL << (x->vars - (j - x->pad) - 1);
            F2 sums[17];
            if (x->have_first) {
                memcpy(sums, x->first, sizeof sums); /* no table has been bound since */
                x->have_first = 0;
            } else {
                node_sums(x, degree, pairs, sums, threads, 0);
            }
            msg[0] = add2(msg[0], mul2(x->weight, sums[0]));
            for (uint32_t z = 1; z < degree; z++) msg[z] = add2(msg[z], mul2(x->weight, sums[z]));
        }
        snprintf(label, sizeof label, "%s/round/%u", prefix, j);
        orc_transcript_absorb_fp2s(t, label, (const uint64_t *)msg, degree);
        snprintf(label, sizeof label, "%s/r/%u", prefix, j);
        F2 r = orc_transcript_challenge_fp2(t, label, NULL);
        memcpy(rounds + 2 * (uint64_t)degree * j, msg, sizeof(F2) * degree);
        challenges[2 * j] = r.c0;
        challenges[2 * j + 1] = r.c1;
        for (uint32_t i = 0; i < count; i++) {
            inst_t *x = &in[i];
            if (j < x->pad) continue;
            uint64_t half = 1ULL << (x->vars - (j - x->pad) - 1);
            bind_instance(x, r, half, threads);
        }
    }
    int64_t nf = 0;
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t k = 0; k < in[i].factors; k++) {
            if (in[i].tables[k].is_public) continue;
            F2 v = entry(&in[i], k, 0);
            finals[2 * nf] = v.c0;
            finals[2 * nf + 1] = v.c1;
            nf++;
        }
    }
    snprintf(label, sizeof label, "%s/evals", prefix);
    orc_transcript_absorb_fp2s(t, label, finals, (size_t)nf);
    gamma_out[0] = gamma.c0;
    gamma_out[1] = gamma.c1;
    claim_out[0] = claim.c0;
    claim_out[1] = claim.c1;
    for (uint32_t i = 0; i < count; i++) {
        for (uint32_t k = 0; k < in[i].factors; k++) free(in[i].cur[k]);
        free(in[i].cur);
    }
    free(in);
    (void)log2_u64;
    return nf;
}
