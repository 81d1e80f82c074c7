// Goldilocks field arithmetic for gfx950 kernels and the host side of the library.
//
// p = 2^64 - 2^32 + 1, so 2^64 = 2^32 - 1 (mod p) and 2^96 = -1 (mod p). Elements are u64;
// every function here returns the canonical representative in [0, p). Corrections are
// branch-free selects (v_cndmask on the device), following the reference's branch-free field
// (crates/tessera-field/src/goldilocks.rs:34-82). Fp2 = Fp[X]/(X^2 - 7) (ext2.rs:104-247).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define TSR_HD __host__ __device__ __forceinline__

namespace gl {

constexpr uint64_t P = 0xFFFFFFFF00000001ull;
constexpr uint64_t EPS = 0xFFFFFFFFull;  // 2^64 mod p

TSR_HD uint64_t canon(uint64_t x) { return x >= P ? x - P : x; }

// a, b < p, so a + b < 2p. With s = a + b mod 2^64: a carry out means the sum is 2^64 + s
// = s + EPS (mod p), and then s + EPS < p; without one, s + EPS carries exactly when s >= p
// (2^64 - EPS = p), and then s + EPS mod 2^64 = s - p. Either way the result is s + EPS.
TSR_HD uint64_t add(uint64_t a, uint64_t b) {
    const uint64_t s = a + b;
    const uint64_t t = s + EPS;
    return (s < a || t < s) ? t : s;
}

// A borrow means the difference is a - b + 2^64; adding p (mod 2^64) gives a - b + p.
TSR_HD uint64_t sub(uint64_t a, uint64_t b) {
    const uint64_t d = a - b;
    return a < b ? d + P : d;
}

TSR_HD uint64_t neg(uint64_t a) { return sub(0, a); }

TSR_HD uint64_t mulhi(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    return __umul64hi(a, b);
#else
    return (uint64_t)(((unsigned __int128)a * b) >> 64);
#endif
}

// x = lo + 2^64 hi_lo + 2^96 hi_hi = lo - hi_hi + hi_lo (2^32 - 1) (mod p). The borrow of
// lo - hi_hi and the carry of the sum are taken from the subtraction and the addition themselves
// (a borrow leaves t0 >= 2^64 - 2^32, so t0 - EPS does not wrap; a carry leaves t2 < t1 <= 2^64 -
// 2^33 + 1, so t2 + EPS does not either): 15 VALU instructions on gfx950 against 18 with compares.
TSR_HD uint64_t reduce128(uint64_t lo, uint64_t hi) {
    const uint32_t hi_hi = (uint32_t)(hi >> 32);
    const uint32_t hi_lo = (uint32_t)hi;
    uint64_t t0;
    const bool borrow = __builtin_sub_overflow(lo, (uint64_t)hi_hi, &t0);
    t0 -= borrow ? EPS : 0;
    // hi_lo (2^32 - 1) = (hi_lo << 32) - hi_lo
    const uint64_t t1 = ((uint64_t)hi_lo << 32) - hi_lo;
    uint64_t t2;
    const bool carry = __builtin_add_overflow(t0, t1, &t2);
    t2 += carry ? EPS : 0;
    return canon(t2);
}

TSR_HD uint64_t mul(uint64_t a, uint64_t b) {
#ifdef __HIP_DEVICE_COMPILE__
    // The 128-bit product from four 32 x 32 + 64 multiply-adds (v_mad_u64_u32). Written as a * b
    // and __umul64hi, the compiler computes the low half twice (two more v_mul_lo_u32): the
    // butterfly rate is 18% lower that way.
    const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t p01 = a0 * b1 + (p00 >> 32);
    const uint64_t p10 = a1 * b0 + (uint32_t)p01;
    const uint64_t hi = a1 * b1 + (p01 >> 32) + (p10 >> 32);
    return reduce128((p10 << 32) | (uint32_t)p00, hi);
#else
    return reduce128(a * b, mulhi(a, b));
#endif
}

// x * 2^E for E in {24, 48, 72}: the 8th and 4th roots of unity are +-2^24, +-2^72 and +-2^48
// (2^96 = -1), so the radix-8 rounds' constant twiddles are shifts. With x = x0 + 2^32 x1:
//   2^24: the 128-bit product (x << 24, x >> 40) reduced;
//   2^48: x0 2^48 = rotl(x0, 16) 2^32 - (x0 >> 16) and x1 2^80 = x1 (2^48 - 2^16), so
//         x 2^48 = (rotl(x0, 16) + rotl(x1, 16)) 2^32 - (x0 >> 16) - (x1 >> 16) - x1 2^16;
//   2^72: 2^72 = 2^40 - 2^8 and 2^104 = -2^8, so x 2^72 = rotl(x0, 8) 2^32 - ((x0 + x1) 2^8 + (x0 >> 24)).
// The last two are A - B with A <= (2^32 - 1) 2^32 < p and 0 <= B < 2^49: A - B, plus p when
// negative, is already canonical (12 and 17 VALU instructions, against 28 and 22 through a
// 128-bit reduction).
TSR_HD uint32_t rotl32(uint32_t x, int r) { return (x << r) | (x >> (32 - r)); }

template <int E>
TSR_HD uint64_t mul_pow2(uint64_t x) {
    static_assert(E == 24 || E == 48 || E == 72, "a root-of-unity shift");
    const uint32_t x0 = (uint32_t)x, x1 = (uint32_t)(x >> 32);
    if constexpr (E == 24) {
        return reduce128(x << 24, x >> 40);
    } else if constexpr (E == 48) {
        const uint64_t R = (uint64_t)rotl32(x0, 16) + rotl32(x1, 16);
        const uint32_t c = (uint32_t)(R >> 32);  // R 2^32 = (R mod 2^32 + c) 2^32 - c
        const uint64_t A = (uint64_t)((uint32_t)R + c) << 32;
        const uint64_t B = ((uint64_t)x1 << 16) + (uint64_t)((x0 >> 16) + (x1 >> 16) + c);
        const uint64_t r = A - B;
        return A < B ? r - EPS : r;  // r + p (mod 2^64)
    } else {
        const uint64_t A = (uint64_t)rotl32(x0, 8) << 32;
        const uint64_t B = (((uint64_t)x0 + x1) << 8) | (x0 >> 24);
        const uint64_t r = A - B;
        return A < B ? r - EPS : r;
    }
}

TSR_HD uint64_t sqr(uint64_t a) { return mul(a, a); }

// Sums of products reduced once (the delayed reduction of crates/tessera-field/src/
// goldilocks.rs:216-250 and ext2.rs:126-178: field-equal to reducing every product). `wide` holds
// a 160-bit integer in 32-bit limbs; mac adds a 128-bit product of two canonical values with one
// carry chain (14 VALU instructions on gfx950 against 31 for a reduced product and a field
// addition). The top limb counts carries past 2^128, one at most per product.
struct wide {
    uint32_t w[5];
};

TSR_HD void mul128(uint64_t a, uint64_t b, uint64_t &lo, uint64_t &hi) {
#ifdef __HIP_DEVICE_COMPILE__
    const uint64_t a0 = (uint32_t)a, a1 = a >> 32, b0 = (uint32_t)b, b1 = b >> 32;
    const uint64_t p00 = a0 * b0;
    const uint64_t p01 = a0 * b1 + (p00 >> 32);
    const uint64_t p10 = a1 * b0 + (uint32_t)p01;
    hi = a1 * b1 + (p01 >> 32) + (p10 >> 32);
    lo = (p10 << 32) | (uint32_t)p00;
#else
    lo = a * b;
    hi = mulhi(a, b);
#endif
}

TSR_HD void mac(wide &x, uint64_t a, uint64_t b) {
    uint64_t lo, hi;
    mul128(a, b, lo, hi);
    unsigned c;
    x.w[0] = __builtin_addc(x.w[0], (uint32_t)lo, 0u, &c);
    x.w[1] = __builtin_addc(x.w[1], (uint32_t)(lo >> 32), c, &c);
    x.w[2] = __builtin_addc(x.w[2], (uint32_t)hi, c, &c);
    x.w[3] = __builtin_addc(x.w[3], (uint32_t)(hi >> 32), c, &c);
    x.w[4] += c;
}

// x = L + 2^64 H + 2^128 T with 2^128 = (2^32 - 1)^2 = -2^32 (mod p), and T 2^32 < p.
TSR_HD uint64_t reduce(const wide &x) {
    const uint64_t L = ((uint64_t)x.w[1] << 32) | x.w[0], H = ((uint64_t)x.w[3] << 32) | x.w[2];
    return sub(reduce128(L, H), (uint64_t)x.w[4] << 32);
}

inline __host__ __device__ uint64_t pow(uint64_t a, uint64_t e) {
    uint64_t acc = 1;
    while (e) {
        if (e & 1) acc = mul(acc, a);
        a = mul(a, a);
        e >>= 1;
    }
    return acc;
}

inline __host__ __device__ uint64_t inv(uint64_t a) { return pow(a, P - 2); }

// A primitive 2^log_n-th root of unity: 7^((p-1)/2^32) squared 32 - log_n times.
inline __host__ __device__ uint64_t root(uint32_t log_n) {
    uint64_t r = pow(7, (P - 1) >> 32);
    for (uint32_t i = log_n; i < 32; i++) r = mul(r, r);
    return r;
}

struct fp2 {
    uint64_t c0, c1;
};

TSR_HD fp2 f2(uint64_t c0, uint64_t c1) { return fp2{c0, c1}; }
TSR_HD fp2 add(fp2 a, fp2 b) { return fp2{add(a.c0, b.c0), add(a.c1, b.c1)}; }
TSR_HD fp2 sub(fp2 a, fp2 b) { return fp2{sub(a.c0, b.c0), sub(a.c1, b.c1)}; }
TSR_HD fp2 neg(fp2 a) { return fp2{neg(a.c0), neg(a.c1)}; }
TSR_HD fp2 mul_base(fp2 a, uint64_t b) { return fp2{mul(a.c0, b), mul(a.c1, b)}; }
TSR_HD uint64_t mul7(uint64_t a) {
    // a*7 < 2^67: lo + hi*(2^64 mod p) with hi < 8.
    uint64_t lo = a * 7, hi = mulhi(a, 7);
    uint64_t t = lo + hi * EPS;
    t += (t < lo) ? EPS : 0;
    return canon(t);
}
TSR_HD fp2 mul(fp2 a, fp2 b) {
    uint64_t v0 = mul(a.c0, b.c0), v1 = mul(a.c1, b.c1);
    uint64_t cross = mul(add(a.c0, a.c1), add(b.c0, b.c1));
    return fp2{add(v0, mul7(v1)), sub(sub(cross, v0), v1)};
}
TSR_HD bool eq(fp2 a, fp2 b) { return a.c0 == b.c0 && a.c1 == b.c1; }

// An Fp2 sum of products, reduced once: c0 += a0 b0 + 7 a1 b1, c1 += a0 b1 + a1 b0.
struct wide2 {
    wide c0, c1;
};
TSR_HD void mac(wide2 &x, fp2 a, fp2 b) {
    mac(x.c0, a.c0, b.c0);
    mac(x.c0, a.c1, mul7(b.c1));
    mac(x.c1, a.c0, b.c1);
    mac(x.c1, a.c1, b.c0);
}
// x += a b for an Fp2 a and a base b.
TSR_HD void mac_base(wide2 &x, fp2 a, uint64_t b) {
    mac(x.c0, a.c0, b);
    mac(x.c1, a.c1, b);
}
TSR_HD fp2 reduce(const wide2 &x) { return fp2{reduce(x.c0), reduce(x.c1)}; }
inline __host__ __device__ fp2 inv(fp2 a) {
    uint64_t norm = sub(mul(a.c0, a.c0), mul7(mul(a.c1, a.c1)));
    uint64_t ni = inv(norm);
    return fp2{mul(a.c0, ni), neg(mul(a.c1, ni))};
}

}  // namespace gl
