/*
 * Test infrastructure (oracle): radix-2 NTTs and the Reed-Solomon row encoder.
 * Restates crates/tessera-pcs/src/ntt.rs (twiddle tables, Cooley-Tukey / Gentleman-Sande
 * kernels, depth-first to 2^12-entry leaves, natural-order transforms) and
 * crates/tessera-pcs/src/encoding.rs (RowEncoder: CT with the inverse table, spread, masks,
 * GS with the forward table). Rows are split over pthreads for the CPU baseline.
 */
#include <pthread.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"
#include "internal.h"

#define LEAF (1u << 12) /* ntt.rs:59 */

static inline uint64_t fadd(uint64_t a, uint64_t b) { return fp_add(a, b); }
static inline uint64_t fsub(uint64_t a, uint64_t b) { return fp_sub(a, b); }
static inline uint64_t fmul(uint64_t a, uint64_t b) { return fp_mul(a, b); }

/* reverse_bits over `bits` bits (ntt.rs:79-84). */
static uint64_t rev(uint64_t i, uint32_t bits) {
    uint64_t r = 0;
    for (uint32_t b = 0; b < bits; b++) r |= ((i >> b) & 1) << (bits - 1 - b);
    return r;
}

/* bit_reversed_powers: out[t] = scale * base^{rev(t)}, t < 2^bits (ntt.rs:86-108). */
static void bit_reversed_powers(uint64_t base, uint32_t bits, uint64_t scale, uint64_t *out) {
    uint64_t squares[33];
    uint64_t sq = base;
    for (uint32_t i = 0; i < bits; i++) {
        squares[i] = sq;
        sq = fmul(sq, sq);
    }
    out[0] = scale;
    uint64_t len = 1;
    for (int e = (int)bits - 1; e >= 0; e--) {
        for (uint64_t t = 0; t < len; t++) out[len + t] = fmul(out[t], squares[e]);
        len *= 2;
    }
}

/* twiddles (ntt.rs:110-125): Z[t] = w^{rev(t)}, w = w_n or w_n^{-1}, t < n/2. */
void orc_twiddles(uint32_t log_n, int inverse, uint64_t *out) {
    if (log_n == 0) return;
    uint64_t root = orc_fp_root(log_n);
    uint64_t base = inverse ? fp_pow(root, (1ULL << log_n) - 1) : root;
    bit_reversed_powers(base, log_n - 1, 1, out);
}

/* inverse_of_power_of_two (ntt.rs:127-131). */
static uint64_t inv_pow2(uint32_t log_n) { return fp_pow(ORC_P / 2 + 1, log_n); }

/* The shift of a kernel: NULL for the subgroup, else s^{2^e} for e < 32 (ntt.rs:133-185). */
static inline uint64_t twiddle(const uint64_t *shift_pows, uint64_t z, uint64_t half) {
    if (!shift_pows) return z;
    return fmul(z, shift_pows[__builtin_ctzll(half)]);
}

/* The butterfly loops take two pairs per iteration: measured on x86-64 (gcc 11), the one-pair
 * GS loop ran at a third of the two-pair loop's rate although its iterations are independent. */
static void ct_bfly(uint64_t *restrict lo, uint64_t *restrict hi, uint64_t n, uint64_t c, int unit) {
    uint64_t i = 0;
    if (unit) {
        for (; i < n; i++) {
            uint64_t a = lo[i], t = hi[i];
            hi[i] = fsub(a, t);
            lo[i] = fadd(a, t);
        }
        return;
    }
    for (; i + 2 <= n; i += 2) {
        uint64_t a0 = lo[i], a1 = lo[i + 1];
        uint64_t t0 = fmul(hi[i], c), t1 = fmul(hi[i + 1], c);
        hi[i] = fsub(a0, t0);
        hi[i + 1] = fsub(a1, t1);
        lo[i] = fadd(a0, t0);
        lo[i + 1] = fadd(a1, t1);
    }
    for (; i < n; i++) {
        uint64_t a = lo[i], t = fmul(hi[i], c);
        hi[i] = fsub(a, t);
        lo[i] = fadd(a, t);
    }
}
static void gs_bfly(uint64_t *restrict lo, uint64_t *restrict hi, uint64_t n, uint64_t c, int unit) {
    uint64_t i = 0;
    if (unit) {
        for (; i < n; i++) {
            uint64_t a = lo[i], b = hi[i];
            lo[i] = fadd(a, b);
            hi[i] = fsub(a, b);
        }
        return;
    }
    for (; i + 2 <= n; i += 2) {
        uint64_t a0 = lo[i], b0 = hi[i], a1 = lo[i + 1], b1 = hi[i + 1];
        uint64_t d0 = fsub(a0, b0), d1 = fsub(a1, b1);
        lo[i] = fadd(a0, b0);
        lo[i + 1] = fadd(a1, b1);
        hi[i] = fmul(d0, c);
        hi[i + 1] = fmul(d1, c);
    }
    for (; i < n; i++) {
        uint64_t a = lo[i], b = hi[i];
        lo[i] = fadd(a, b);
        hi[i] = fmul(fsub(a, b), c);
    }
}

/* ct_leaf (ntt.rs:257-281). */
static void ct_leaf(uint64_t *v, uint64_t len, const uint64_t *tw, const uint64_t *sp,
                    uint64_t group) {
    uint64_t half = len / 2, first = group, count = 1;
    while (half > 0) {
        for (uint64_t g = 0; g < count; g++) {
            uint64_t *chunk = v + g * 2 * half;
            int unit = (!sp && first + g == 0);
            ct_bfly(chunk, chunk + half, half, twiddle(sp, tw[first + g], half), unit);
        }
        half /= 2;
        first *= 2;
        count *= 2;
    }
}

/* ct_group (ntt.rs:232-255): depth-first until a group fits the 2^12-entry leaf. */
static void ct_group(uint64_t *v, uint64_t len, const uint64_t *tw, const uint64_t *sp,
                     uint64_t group) {
    if (len <= LEAF) {
        ct_leaf(v, len, tw, sp, group);
        return;
    }
    uint64_t half = len / 2;
    ct_bfly(v, v + half, half, twiddle(sp, tw[group], half), !sp && group == 0);
    ct_group(v, half, tw, sp, 2 * group);
    ct_group(v + half, half, tw, sp, 2 * group + 1);
}

/* gs_leaf (ntt.rs:318-340). */
static void gs_leaf(uint64_t *v, uint64_t len, const uint64_t *tw, const uint64_t *sp,
                    uint64_t group) {
    uint64_t half = 1, count = len / 2, first = group * count;
    while (count > 0) {
        for (uint64_t g = 0; g < count; g++) {
            uint64_t *chunk = v + g * 2 * half;
            int unit = (!sp && first + g == 0);
            gs_bfly(chunk, chunk + half, half, twiddle(sp, tw[first + g], half), unit);
        }
        half *= 2;
        first /= 2;
        count /= 2;
    }
}

/* gs_group (ntt.rs:296-316). */
static void gs_group(uint64_t *v, uint64_t len, const uint64_t *tw, const uint64_t *sp,
                     uint64_t group) {
    if (len <= LEAF) {
        gs_leaf(v, len, tw, sp, group);
        return;
    }
    uint64_t half = len / 2;
    gs_group(v, half, tw, sp, 2 * group);
    gs_group(v + half, half, tw, sp, 2 * group + 1);
    gs_bfly(v, v + half, half, twiddle(sp, tw[group], half), !sp && group == 0);
}

/* bit_reverse_scaled (ntt.rs:343-393): position rev(i) receives scale * v[i].
 * Small sizes swap pairs directly. Larger ones split an index into a high part h and a low part
 * l of RB bits around a middle m: rev(h, m, l) = (rev(l), rev(m), rev(h)), so the RT x RT block
 * of middle m (rows h, columns l) lands transposed, both parts reversed, in the block of rev(m).
 * Blocks m and rev(m) are read into two buffers and written back crosswise, so every access is
 * a run of RT consecutive words instead of a scattered single word. */
#define RB 5u
#define RT (1u << RB)
/* rev over `bits` bits by swapping halves, quarters, ... of a 64-bit word. */
static uint64_t rev_fast(uint64_t x, uint32_t bits) {
    x = ((x >> 1) & 0x5555555555555555ULL) | ((x & 0x5555555555555555ULL) << 1);
    x = ((x >> 2) & 0x3333333333333333ULL) | ((x & 0x3333333333333333ULL) << 2);
    x = ((x >> 4) & 0x0F0F0F0F0F0F0F0FULL) | ((x & 0x0F0F0F0F0F0F0F0FULL) << 4);
    x = __builtin_bswap64(x);
    return bits ? x >> (64 - bits) : 0;
}
static void bit_reverse_scaled(uint64_t *v, uint32_t bits, const uint64_t *scale) {
    uint64_t n = 1ULL << bits;
    if (bits < 2 * RB) {
        for (uint64_t i = 0; i < n; i++) {
            uint64_t j = rev(i, bits);
            if (i < j) {
                uint64_t x = v[i], y = v[j];
                v[i] = scale ? fmul(y, *scale) : y;
                v[j] = scale ? fmul(x, *scale) : x;
            } else if (i == j && scale) {
                v[i] = fmul(v[i], *scale);
            }
        }
        return;
    }
    uint32_t mid = bits - 2 * RB;
    uint64_t row = (uint64_t)RT << mid; /* index distance between consecutive h */
    uint32_t rt[RT];
    for (uint32_t t = 0; t < RT; t++) rt[t] = (uint32_t)rev(t, RB);
    uint64_t A[RT * RT], B[RT * RT];
    for (uint64_t m = 0; m < (1ULL << mid); m++) {
        uint64_t mr = rev_fast(m, mid);
        if (mr < m) continue; /* the pair (mr, m) was done at mr */
        for (uint32_t h = 0; h < RT; h++) {
            memcpy(A + h * RT, v + h * row + m * RT, 8 * RT);
            memcpy(B + h * RT, v + h * row + mr * RT, 8 * RT);
        }
        /* destination row r = rev(l), column c = rev(h) of block mr gets A[h][l] */
        for (uint32_t r = 0; r < RT; r++) {
            uint64_t *dst = v + r * row + mr * RT;
            uint32_t l = rt[r];
            for (uint32_t c = 0; c < RT; c++) {
                uint64_t x = A[rt[c] * RT + l];
                dst[c] = scale ? fmul(x, *scale) : x;
            }
        }
        if (mr != m) {
            for (uint32_t r = 0; r < RT; r++) {
                uint64_t *dst = v + r * row + m * RT;
                uint32_t l = rt[r];
                for (uint32_t c = 0; c < RT; c++) {
                    uint64_t x = B[rt[c] * RT + l];
                    dst[c] = scale ? fmul(x, *scale) : x;
                }
            }
        }
    }
}

static void shift_powers(uint64_t s, uint64_t out[32]) {
    out[0] = s;
    for (int e = 1; e < 32; e++) out[e] = fmul(out[e - 1], out[e - 1]);
}

/* NttPlan::ntt / intt / coset_ntt / coset_intt (ntt.rs:461-516). */
void orc_ntt(uint64_t *v, uint32_t log_n) {
    uint64_t *tw = malloc(sizeof(uint64_t) * ((1ULL << log_n) / 2 + 1));
    orc_twiddles(log_n, 0, tw);
    ct_group(v, 1ULL << log_n, tw, NULL, 0);
    bit_reverse_scaled(v, log_n, NULL);
    free(tw);
}
void orc_intt(uint64_t *v, uint32_t log_n) {
    uint64_t *tw = malloc(sizeof(uint64_t) * ((1ULL << log_n) / 2 + 1));
    uint64_t s = inv_pow2(log_n);
    orc_twiddles(log_n, 1, tw);
    bit_reverse_scaled(v, log_n, &s);
    gs_group(v, 1ULL << log_n, tw, NULL, 0);
    free(tw);
}
void orc_coset_ntt(uint64_t *v, uint32_t log_n, uint64_t shift) {
    uint64_t *tw = malloc(sizeof(uint64_t) * ((1ULL << log_n) / 2 + 1));
    uint64_t sp[32];
    orc_twiddles(log_n, 0, tw);
    shift_powers(shift, sp);
    ct_group(v, 1ULL << log_n, tw, sp, 0);
    bit_reverse_scaled(v, log_n, NULL);
    free(tw);
}
void orc_coset_intt(uint64_t *v, uint32_t log_n, uint64_t shift) {
    uint64_t *tw = malloc(sizeof(uint64_t) * ((1ULL << log_n) / 2 + 1));
    uint64_t sp[32], s = inv_pow2(log_n);
    orc_twiddles(log_n, 1, tw);
    shift_powers(orc_fp_inv(shift), sp);
    bit_reverse_scaled(v, log_n, &s);
    gs_group(v, 1ULL << log_n, tw, sp, 0);
    free(tw);
}

/* ---- RowEncoder (encoding.rs:126-340) ---- */
typedef struct {
    uint32_t log_k, log_r;
    uint64_t P;
    uint64_t *inverse; /* w_k^{-rev(t)}, t < k/2 */
    uint64_t *forward; /* w_n^{rev(t)},  t < n/2 */
    uint64_t *spread;  /* 7^{rev_k(t)}/k, t < k */
    uint64_t shift_pow_k;
    const uint64_t *rows, *masks;
    uint64_t *codewords;
} encoder_t;

/* spread (encoding.rs:323-340): entry t times scale[t] to position s*t, zeros after it. */
static void spread(uint64_t *cw, uint64_t n, const uint64_t *scale, uint64_t k) {
    uint64_t stride = n / k;
    for (uint64_t m = k; m >= 2; m /= 2) {
        for (uint64_t t = m / 2; t < m; t++) {
            uint64_t y = cw[t];
            cw[stride * t] = fmul(y, scale[t]);
            for (uint64_t z = 1; z < stride; z++) cw[stride * t + z] = 0;
        }
    }
    cw[0] = fmul(cw[0], scale[0]);
    for (uint64_t z = 1; z < stride; z++) cw[z] = 0;
}

/* encode (encoding.rs:292-320). */
static void encode_one(const encoder_t *e, const uint64_t *row, const uint64_t *mask,
                       uint64_t *cw) {
    uint64_t k = 1ULL << e->log_k, n = k << e->log_r, stride = 1ULL << e->log_r;
    memcpy(cw, row, k * 8);
    ct_group(cw, k, e->inverse, NULL, 0);
    spread(cw, n, e->spread, k);
    uint64_t power = 1;
    for (uint64_t i = 0; i < e->P; i++) {
        uint64_t scaled = fmul(mask[i], power);
        uint64_t low = stride * rev(i, e->log_k);
        cw[low] = fsub(cw[low], scaled);
        cw[low + stride / 2] = fmul(scaled, e->shift_pow_k);
        power = fmul(power, 7);
    }
    gs_group(cw, n, e->forward, NULL, 0);
}

static void encode_range(void *arg, uint64_t lo, uint64_t hi) {
    const encoder_t *e = arg;
    uint64_t k = 1ULL << e->log_k, n = k << e->log_r;
    for (uint64_t r = lo; r < hi; r++)
        encode_one(e, e->rows + r * k, e->masks ? e->masks + r * e->P : NULL, e->codewords + r * n);
}

int orc_encode_rows(uint32_t log_k, uint32_t log_r, uint64_t P, const uint64_t *rows,
                    const uint64_t *masks, uint64_t count, uint64_t *codewords, int threads) {
    if (log_r == 0 || log_k + log_r > 32 || P > (1ULL << log_k)) return -1;
    encoder_t e;
    uint64_t k = 1ULL << log_k, n = k << log_r;
    e.log_k = log_k;
    e.log_r = log_r;
    e.P = P;
    e.inverse = malloc(8 * (k / 2 + 1));
    e.forward = malloc(8 * (n / 2 + 1));
    e.spread = malloc(8 * k);
    orc_twiddles(log_k, 1, e.inverse);
    orc_twiddles(log_k + log_r, 0, e.forward);
    bit_reversed_powers(7, log_k, inv_pow2(log_k), e.spread);
    e.shift_pow_k = fp_pow(7, k);
    e.rows = rows;
    e.masks = masks;
    e.codewords = codewords;
    orc_parallel_for(count, threads, encode_range, &e);
    free(e.inverse);
    free(e.forward);
    free(e.spread);
    return 0;
}

/* ---- a small static-split thread pool ---- */
typedef struct {
    void (*fn)(void *, int, uint64_t, uint64_t);
    void *arg;
    int t;
    uint64_t lo, hi;
} job_t;
static void *job_run(void *p) {
    job_t *j = p;
    j->fn(j->arg, j->t, j->lo, j->hi);
    return NULL;
}
void orc_parallel_for_t(uint64_t count, int threads,
                        void (*fn)(void *, int, uint64_t, uint64_t), void *arg) {
    if (threads <= 1 || count <= 1) {
        fn(arg, 0, 0, count);
        return;
    }
    if ((uint64_t)threads > count) threads = (int)count;
    if (threads > 256) threads = 256;
    pthread_t tid[256];
    int spawned[256] = {0};
    job_t jobs[256];
    for (int t = 0; t < threads; t++) {
        jobs[t].fn = fn;
        jobs[t].arg = arg;
        jobs[t].t = t;
        jobs[t].lo = count * t / threads;
        jobs[t].hi = count * (t + 1) / threads;
    }
    for (int t = 1; t < threads; t++) {
        if (pthread_create(&tid[t], NULL, job_run, &jobs[t]) == 0) spawned[t] = 1;
        else job_run(&jobs[t]); /* spawn failure: run it on this thread */
    }
    job_run(&jobs[0]);
    for (int t = 1; t < threads; t++)
        if (spawned[t]) pthread_join(tid[t], NULL);
}

typedef struct {
    void (*fn)(void *, uint64_t, uint64_t);
    void *arg;
} plain_t;
static void plain_run(void *p, int t, uint64_t lo, uint64_t hi) {
    (void)t;
    plain_t *x = p;
    x->fn(x->arg, lo, hi);
}
void orc_parallel_for(uint64_t count, int threads, void (*fn)(void *, uint64_t, uint64_t),
                      void *arg) {
    plain_t x = {fn, arg};
    orc_parallel_for_t(count, threads, plain_run, &x);
}
