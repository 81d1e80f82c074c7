/*
 * Test infrastructure (oracle): Goldilocks Fp and Fp2 = Fp[X]/(X^2 - 7).
 * Restates crates/tessera-field/src/goldilocks.rs and ext2.rs with unsigned __int128.
 */
#include "oracle.h"

typedef unsigned __int128 u128;

static const uint64_t EPS = 0xFFFFFFFFULL; /* 2^64 - p */

/* canonicalize: x < 2^64 < 2p, one conditional subtraction (goldilocks.rs:34-39). */
static inline uint64_t canon(uint64_t x) { return x >= ORC_P ? x - ORC_P : x; }

/* add (goldilocks.rs:41-48): a + b mod p for canonical inputs. */
uint64_t orc_fp_add(uint64_t a, uint64_t b) {
    uint64_t s = a + b;
    if (s < a) s += EPS; /* wrapped: 2^64 = EPS mod p */
    return canon(s);
}

/* sub (goldilocks.rs:50-56). */
uint64_t orc_fp_sub(uint64_t a, uint64_t b) {
    uint64_t d = a - b;
    if (a < b) d -= EPS; /* borrow: -2^64 = -EPS mod p */
    return d;
}

/* reduce128 (goldilocks.rs:58-75): x = lo + 2^64 (hi_lo + 2^32 hi_hi)
 *   = lo - hi_hi + hi_lo (2^32 - 1)  (mod p). */
uint64_t orc_fp_reduce128(uint64_t lo, uint64_t hi) {
    uint64_t hi_hi = hi >> 32, hi_lo = hi & EPS;
    uint64_t t0 = lo - hi_hi;
    if (lo < hi_hi) t0 -= EPS;
    uint64_t t1 = hi_lo * EPS;
    uint64_t t2 = t0 + t1;
    if (t2 < t0) t2 += EPS;
    return canon(t2);
}

/* mul (goldilocks.rs:77-80). */
uint64_t orc_fp_mul(uint64_t a, uint64_t b) {
    u128 x = (u128)a * b;
    return orc_fp_reduce128((uint64_t)x, (uint64_t)(x >> 64));
}

/* pow by square-and-multiply (goldilocks.rs:82-95). */
uint64_t orc_fp_pow(uint64_t a, uint64_t e) {
    uint64_t acc = 1, sq = a;
    while (e) {
        if (e & 1) acc = orc_fp_mul(acc, sq);
        sq = orc_fp_mul(sq, sq);
        e >>= 1;
    }
    return acc;
}

/* inverse = x^(p-2), zero maps to zero (goldilocks.rs:338-356). */
uint64_t orc_fp_inv(uint64_t a) { return orc_fp_pow(a, ORC_P - 2); }

/* two_adic_generator: 7^((p-1)/2^32) squared 32-log_n times (goldilocks.rs:168-180, 397). */
uint64_t orc_fp_root(uint32_t log_n) {
    uint64_t r = orc_fp_pow(7, (ORC_P - 1) >> 32);
    for (uint32_t i = log_n; i < 32; i++) r = orc_fp_mul(r, r);
    return r;
}

orc_fp2 orc_fp2_add(orc_fp2 a, orc_fp2 b) {
    orc_fp2 r = {orc_fp_add(a.c0, b.c0), orc_fp_add(a.c1, b.c1)};
    return r;
}
orc_fp2 orc_fp2_sub(orc_fp2 a, orc_fp2 b) {
    orc_fp2 r = {orc_fp_sub(a.c0, b.c0), orc_fp_sub(a.c1, b.c1)};
    return r;
}
/* Karatsuba (ext2.rs:233-247): c0 = a0 b0 + 7 a1 b1, c1 = (a0+a1)(b0+b1) - a0 b0 - a1 b1. */
orc_fp2 orc_fp2_mul(orc_fp2 a, orc_fp2 b) {
    uint64_t v0 = orc_fp_mul(a.c0, b.c0), v1 = orc_fp_mul(a.c1, b.c1);
    uint64_t cross = orc_fp_mul(orc_fp_add(a.c0, a.c1), orc_fp_add(b.c0, b.c1));
    orc_fp2 r = {orc_fp_add(v0, orc_fp_mul(v1, 7)), orc_fp_sub(orc_fp_sub(cross, v0), v1)};
    return r;
}
orc_fp2 orc_fp2_mul_base(orc_fp2 a, uint64_t b) {
    orc_fp2 r = {orc_fp_mul(a.c0, b), orc_fp_mul(a.c1, b)};
    return r;
}
/* inverse via the norm c0^2 - 7 c1^2 (ext2.rs norm/conjugate). */
orc_fp2 orc_fp2_inv(orc_fp2 a) {
    uint64_t norm = orc_fp_sub(orc_fp_mul(a.c0, a.c0), orc_fp_mul(7, orc_fp_mul(a.c1, a.c1)));
    uint64_t ni = orc_fp_inv(norm);
    orc_fp2 r = {orc_fp_mul(a.c0, ni), orc_fp_sub(0, orc_fp_mul(a.c1, ni))};
    return r;
}
void orc_fp_mul_vec(const uint64_t *a, const uint64_t *b, uint64_t *out, uint64_t n) {
    for (uint64_t i = 0; i < n; i++) out[i] = orc_fp_mul(a[i], b[i]);
}
void orc_fp2_mul_v(const uint64_t a[2], const uint64_t b[2], uint64_t out[2]) {
    orc_fp2 x = {a[0], a[1]}, y = {b[0], b[1]};
    orc_fp2 r = orc_fp2_mul(x, y);
    out[0] = r.c0;
    out[1] = r.c1;
}
void orc_fp2_inv_v(const uint64_t a[2], uint64_t out[2]) {
    orc_fp2 x = {a[0], a[1]};
    orc_fp2 r = orc_fp2_inv(x);
    out[0] = r.c0;
    out[1] = r.c1;
}
