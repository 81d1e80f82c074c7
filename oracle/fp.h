/*
 * Test infrastructure (oracle): Goldilocks Fp and Fp2 = Fp[X]/(X^2 - 7) as static inline
 * functions, so every oracle file's inner loops inline the arithmetic instead of calling the
 * library's exported symbols (which a -fPIC shared object reaches through the PLT).
 * Restates crates/tessera-field/src/goldilocks.rs:34-82 and ext2.rs:233-247 with
 * unsigned __int128; field.c exports the same operations under the orc_fp_* names.
 */
#ifndef TESSERA_ORACLE_FP_H
#define TESSERA_ORACLE_FP_H
#include <stdint.h>

#ifndef ORC_P
#define ORC_P 0xFFFFFFFF00000001ULL
#endif
#define ORC_EPS 0xFFFFFFFFULL /* 2^64 - p */

typedef unsigned __int128 orc_u128;

/* The carries below are masks, not branches: on uniform inputs each carry is a coin flip, and
 * a mispredicted branch costs more than the whole operation (the reference selects with
 * select_unpredictable for the same reason, goldilocks.rs:34-75). */
static inline uint64_t fp_mask(int c) { return (uint64_t)0 - (uint64_t)c; }

/* canonicalize: x < 2^64 < 2p, one conditional subtraction (goldilocks.rs:34-39). */
static inline uint64_t fp_canon(uint64_t x) { return x - (ORC_P & fp_mask(x >= ORC_P)); }

/* add (goldilocks.rs:41-48): a + b mod p for canonical inputs; 2^64 = EPS mod p. */
static inline uint64_t fp_add(uint64_t a, uint64_t b) {
    uint64_t s;
    int c = __builtin_add_overflow(a, b, &s);
    return fp_canon(s + (ORC_EPS & fp_mask(c)));
}

/* sub (goldilocks.rs:50-56); -2^64 = -EPS mod p. */
static inline uint64_t fp_sub(uint64_t a, uint64_t b) {
    uint64_t d;
    int c = __builtin_sub_overflow(a, b, &d);
    return d - (ORC_EPS & fp_mask(c));
}

/* reduce128 (goldilocks.rs:58-75): x = lo + 2^64 (hi_lo + 2^32 hi_hi)
 *   = lo - hi_hi + hi_lo (2^32 - 1)  (mod p). */
static inline uint64_t fp_reduce128(uint64_t lo, uint64_t hi) {
    uint64_t hi_hi = hi >> 32, hi_lo = hi & ORC_EPS;
    uint64_t t0, t2;
    int b = __builtin_sub_overflow(lo, hi_hi, &t0);
    t0 -= ORC_EPS & fp_mask(b);
    uint64_t t1 = hi_lo * ORC_EPS;
    int c = __builtin_add_overflow(t0, t1, &t2);
    return fp_canon(t2 + (ORC_EPS & fp_mask(c)));
}

/* mul (goldilocks.rs:77-80). */
static inline uint64_t fp_mul(uint64_t a, uint64_t b) {
    orc_u128 x = (orc_u128)a * b;
    return fp_reduce128((uint64_t)x, (uint64_t)(x >> 64));
}

static inline uint64_t fp_pow(uint64_t a, uint64_t e) {
    uint64_t acc = 1, sq = a;
    while (e) {
        if (e & 1) acc = fp_mul(acc, sq);
        sq = fp_mul(sq, sq);
        e >>= 1;
    }
    return acc;
}

typedef struct { uint64_t c0, c1; } orc_fp2;

static inline orc_fp2 fp2_add(orc_fp2 a, orc_fp2 b) {
    orc_fp2 r = {fp_add(a.c0, b.c0), fp_add(a.c1, b.c1)};
    return r;
}
static inline orc_fp2 fp2_sub(orc_fp2 a, orc_fp2 b) {
    orc_fp2 r = {fp_sub(a.c0, b.c0), fp_sub(a.c1, b.c1)};
    return r;
}
/* Karatsuba (ext2.rs:233-247): c0 = a0 b0 + 7 a1 b1, c1 = (a0+a1)(b0+b1) - a0 b0 - a1 b1. */
static inline orc_fp2 fp2_mul(orc_fp2 a, orc_fp2 b) {
    uint64_t v0 = fp_mul(a.c0, b.c0), v1 = fp_mul(a.c1, b.c1);
    uint64_t cross = fp_mul(fp_add(a.c0, a.c1), fp_add(b.c0, b.c1));
    orc_fp2 r = {fp_add(v0, fp_mul(v1, 7)), fp_sub(fp_sub(cross, v0), v1)};
    return r;
}
static inline orc_fp2 fp2_mul_base(orc_fp2 a, uint64_t b) {
    orc_fp2 r = {fp_mul(a.c0, b), fp_mul(a.c1, b)};
    return r;
}
#endif
