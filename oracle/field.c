/*
 * Test infrastructure (oracle): Goldilocks Fp and Fp2 = Fp[X]/(X^2 - 7).
 * Restates crates/tessera-field/src/goldilocks.rs and ext2.rs with unsigned __int128.
 */
#include "oracle.h"

/* The arithmetic itself is in fp.h (static inline, so the other oracle files inline it); these
 * are the exported entry points the Python bindings call. */
uint64_t orc_fp_add(uint64_t a, uint64_t b) { return fp_add(a, b); }
uint64_t orc_fp_sub(uint64_t a, uint64_t b) { return fp_sub(a, b); }
uint64_t orc_fp_reduce128(uint64_t lo, uint64_t hi) { return fp_reduce128(lo, hi); }
uint64_t orc_fp_mul(uint64_t a, uint64_t b) { return fp_mul(a, b); }

/* pow by square-and-multiply (goldilocks.rs:82-95). */
uint64_t orc_fp_pow(uint64_t a, uint64_t e) { return fp_pow(a, e); }

/* inverse = x^(p-2), zero maps to zero (goldilocks.rs:338-356). */
uint64_t orc_fp_inv(uint64_t a) { return orc_fp_pow(a, ORC_P - 2); }

/* two_adic_generator: 7^((p-1)/2^32) squared 32-log_n times (goldilocks.rs:168-180, 397). */
uint64_t orc_fp_root(uint32_t log_n) {
    uint64_t r = orc_fp_pow(7, (ORC_P - 1) >> 32);
    for (uint32_t i = log_n; i < 32; i++) r = orc_fp_mul(r, r);
    return r;
}

orc_fp2 orc_fp2_add(orc_fp2 a, orc_fp2 b) { return fp2_add(a, b); }
orc_fp2 orc_fp2_sub(orc_fp2 a, orc_fp2 b) { return fp2_sub(a, b); }
orc_fp2 orc_fp2_mul(orc_fp2 a, orc_fp2 b) { return fp2_mul(a, b); }
orc_fp2 orc_fp2_mul_base(orc_fp2 a, uint64_t b) { return fp2_mul_base(a, b); }
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
