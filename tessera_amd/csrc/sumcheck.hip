// Batched sumchecks on gfx950, with the round loop and its transcript on the device.
//
// Semantics: crates/tessera-sumcheck/src/prover.rs:1122-1410 (prove_batch / prove_checked,
// State::{new, add_round, final_evals}, padding_scale) and level.rs:535-590 (prove_level):
// instance i on n_i variables, weight gamma^i, front-padded to n_max; round j's message is
// p_j(0), p_j(2), ..., p_j(d) of the batched polynomial, absorbed as TCE Vec<Fp2> under
// <prefix>/round/{j}; r_j is squeezed under <prefix>/r/{j}; a padding round adds
// 2^{pad - j - 1} gamma^i S_i at every node (S_i from the tables for prove_batch, the claim for a
// tree level: equal for honest tables); binding is t[k] = t[2k] + r (t[2k+1] - t[2k])
// (tessera-poly multilinear.rs:70-121).
//
// Three relations: a product of the instance's tables (prove_batch's Instance::new), a
// fraction-tree level eq (N0 D1 + N1 D0 + lambda D0 D1) over [eq, N0, N1, D0, D1], and a sum of
// weighted products over the instance's tables (Instance::sum_of_products, prover.rs:243-312;
// kernel.rs:502-633 term_sums, whose grouping by weight and common factors changes no value).
//
// Execution: the caller squeezes gamma (and lambda) on the host transcript, then run_batch
// enqueues the whole loop on the context's stream: the first-round sums of every instance
// (k_sc_round, no bind), then for each round j a one-block k_sc_transcript (reduces the
// per-block partial sums, forms the message, absorbs it, squeezes r_j with the device
// transcript) followed by k_sc_round for every instance in its own rounds (binds all its tables
// with r_j and, fused in the same pass, sums the next round's node products). One
// synchronization at the end; the host then absorbs the final evaluations.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <type_traits>
#include <vector>
#include "common.h"
#include "devtranscript.h"
#include "field.h"
#include "sumcheck.h"
#include "transcript.h"

namespace tsr {

using gl::fp2;

constexpr int SC_MAXF = 10;
constexpr int SC_MAXD = 10;
constexpr int SC_THREADS = 256;
#ifndef TSR_SC_MAX_BLOCKS
#define TSR_SC_MAX_BLOCKS 1024
#endif
constexpr int SC_MAX_BLOCKS = TSR_SC_MAX_BLOCKS;  // workgroups of a round launch at most
// A level round over 2^m pairs splits eq into eq_lo (2^a entries) and eq_hi (2^(m - a)); a = m
// up to this many variables (eq_lo at most 128 KiB), else max(this, m / 2).
constexpr uint32_t SC_EQ_LO_BITS = 13;

constexpr int SC_MAXT = 16;

// The terms of a REL_TERMS instance: term t is weight[t] * prod_k T_k^mult[t][k].
struct ScTerms {
    uint32_t nterms;
    uint32_t unit;  // bit t: weight[t] is 1
    fp2 weight[SC_MAXT];
    uint8_t mult[SC_MAXT][SC_MAXF];
};

struct ScTables {
    const void *src[SC_MAXF];  // nullptr: the constant 1 (a fraction tree's unit numerators)
    fp2 *dst[SC_MAXF];
    uint8_t shift[SC_MAXF], off[SC_MAXF];  // first round: entry x is src[(x << shift) + off]
};

__device__ __forceinline__ fp2 ld_entry(const void *p, bool base, uint64_t x) {
    if (!p) return fp2{1, 0};
    if (base) return fp2{((const uint64_t *)p)[x], 0};
    return ((const fp2 *)p)[x];
}

// Node values of one pair, slots [p(0), p(2), ..., p(D), p(1)] (D+1 of them).
// FT > 0: exactly FT tables, known at compile time (else F at run time, at most SC_MAXF).
template <int D, int REL, int FT = 0>
__device__ __forceinline__ void pair_values(const fp2 *a, const fp2 *b, uint32_t F,
                                            const fp2 &lambda, const ScTerms &terms,
                                            fp2 out[D + 1], uint32_t basem = 0) {
    if (REL == REL_TERMS) {
        // line values of every table at z = 0, 2, ..., D, then 1; the tables are indexed at
        // compile time (unrolled), the terms and multiplicities are uniform across the wave
        fp2 line[SC_MAXF];
#pragma unroll
        for (int s = 0; s <= D; s++) {
            fp2 v[SC_MAXF];
#pragma unroll
            for (int k = 0; k < SC_MAXF; k++) {
                if (k >= (int)F) break;
                if (s == 0) {
                    v[k] = a[k];
                } else if (s == D) {
                    v[k] = b[k];
                } else {
                    line[k] = s == 1 ? gl::add(b[k], gl::sub(b[k], a[k])) : gl::add(line[k], gl::sub(b[k], a[k]));
                    v[k] = line[k];
                }
            }
            fp2 sum{0, 0};
            for (uint32_t t = 0; t < terms.nterms; t++) {
                fp2 p = terms.weight[t];
                bool first = (terms.unit >> t) & 1;
#pragma unroll
                for (int k = 0; k < SC_MAXF; k++) {
                    if (k >= (int)F) break;
                    for (uint32_t m = 0; m < terms.mult[t][k]; m++) {
                        p = first ? v[k] : gl::mul(p, v[k]);
                        first = false;
                    }
                }
                sum = gl::add(sum, p);
            }
            out[s] = sum;
        }
    } else if (REL == REL_PRODUCT) {
        // basem bit k: table k's values are base (c1 = 0), so its products are base products
#pragma unroll
        for (int k = 0; k < (FT > 0 ? FT : SC_MAXF); k++) {
            if (FT == 0 && k >= (int)F) break;
            const bool bk = (basem >> k) & 1;
            fp2 diff = gl::sub(b[k], a[k]);
            fp2 line = gl::add(b[k], diff);  // z = 2
            auto times = [&](const fp2 &x, const fp2 &v) { return bk ? gl::mul_base(x, v.c0) : gl::mul(x, v); };
            if (k == 0) {
                out[0] = a[k];
                out[D] = b[k];
#pragma unroll
                for (int z = 2; z <= D; z++) {
                    out[z - 1] = line;
                    line = gl::add(line, diff);
                }
            } else {
                out[0] = times(out[0], a[k]);
                out[D] = times(out[D], b[k]);
#pragma unroll
                for (int z = 2; z <= D; z++) {
                    out[z - 1] = times(out[z - 1], line);
                    line = gl::add(line, diff);
                }
            }
        }
    }
}

__device__ __forceinline__ uint64_t shfl_xor_u64(uint64_t v, int mask) {
    uint32_t lo = (uint32_t)v, hi = (uint32_t)(v >> 32);
    lo = __shfl_xor(lo, mask, 64);
    hi = __shfl_xor(hi, mask, 64);
    return ((uint64_t)hi << 32) | lo;
}

// One item q of an own round in k_sc_round's modes: binds the tables at q and/or adds the
// pair's node values to acc.
template <int D, int REL, int FT = 0>
__device__ __forceinline__ void sc_item(const ScTables &t, uint32_t F, uint32_t base_mask, bool first,
                                        bool bind, uint64_t pairs, const fp2 &r, const fp2 &lambda,
                                        const ScTerms &terms, uint64_t q, fp2 (&acc)[D + 1]) {
    // the tables are indexed at compile time (unrolled up to SC_MAXF, F uniform), so a and b
    // stay in registers; a base table of the first round is bound with base products
    constexpr int KMAX = FT > 0 ? FT : SC_MAXF;
    fp2 a[KMAX], b[KMAX];
    // e0 + r (e1 - e0), with two base products when e0, e1 are base
    auto bindv = [&](const fp2 &e0, const fp2 &e1, bool base) {
        if (base) {
            fp2 v = gl::mul_base(r, gl::sub(e1.c0, e0.c0));
            v.c0 = gl::add(v.c0, e0.c0);
            return v;
        }
        return gl::add(e0, gl::mul(r, gl::sub(e1, e0)));
    };
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        if (FT == 0 && k >= (int)F) break;
        const bool base = first && ((base_mask >> k) & 1);
        const uint32_t sh = first ? t.shift[k] : 0, of = first ? t.off[k] : 0;
        auto at = [&](uint64_t x) { return ld_entry(t.src[k], base, (x << sh) + of); };
        if (!bind) {
            a[k] = at(2 * q);
            b[k] = at(2 * q + 1);
        } else if (pairs > 0) {
            fp2 e0 = at(4 * q), e1 = at(4 * q + 1), e2 = at(4 * q + 2), e3 = at(4 * q + 3);
            a[k] = bindv(e0, e1, base);
            b[k] = bindv(e2, e3, base);
            if (t.dst[k]) {
                t.dst[k][2 * q] = a[k];
                t.dst[k][2 * q + 1] = b[k];
            }
        } else {
            fp2 e0 = at(2 * q), e1 = at(2 * q + 1);
            if (t.dst[k]) t.dst[k][q] = bindv(e0, e1, base);
        }
    }
    if (pairs > 0) {
        fp2 val[D + 1];
        // first round without binding: base tables multiply as base values
        pair_values<D, REL, FT>(a, b, F, lambda, terms, val, bind ? 0u : (first ? base_mask : 0u));
#pragma unroll
        for (int z = 0; z <= D; z++) acc[z] = gl::add(acc[z], val[z]);
    }
}

// One own round of one instance: bind every table with r (when out_len > 0) and/or sum the node
// values over `pairs` pairs (when pairs > 0), writing one partial per block.
//   sums only:   pairs of the unbound tables (entries 2q, 2q+1)
//   bind + sums: dst[2q], dst[2q+1] from src[4q..4q+3]; the pair (dst[2q], dst[2q+1]) summed
//   bind only:   dst[q] from src[2q], src[2q+1] for q < out_len
// `first`: the sources are the instance's original tables (base entries, strides, offsets).
template <int D, int REL, int FT = 0>
__global__ void __launch_bounds__(SC_THREADS)
k_sc_round(ScTables t, uint32_t F, uint32_t base_mask, bool first, uint64_t out_len,
           uint64_t pairs, const fp2 *__restrict__ r_ptr, fp2 lambda, fp2 *__restrict__ partials,
           ScTerms terms) {
    const bool bind = out_len > 0;
    const fp2 r = bind ? *r_ptr : fp2{0, 0};
    const uint64_t items = pairs > 0 ? pairs : out_len;
    const uint64_t stride = (uint64_t)gridDim.x * SC_THREADS;
    fp2 acc[D + 1];
#pragma unroll
    for (int z = 0; z <= D; z++) acc[z] = fp2{0, 0};

    for (uint64_t q = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x; q < items; q += stride)
        sc_item<D, REL, FT>(t, F, base_mask, first, bind, pairs, r, lambda, terms, q, acc);
    if (pairs == 0) return;

    // wave reduction, then across the block's waves through LDS
    __shared__ fp2 red[SC_THREADS / 64][D + 1];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int z = 0; z <= D; z++) {
        fp2 v = acc[z];
        for (int m = 32; m >= 1; m >>= 1) {
            fp2 o{shfl_xor_u64(v.c0, m), shfl_xor_u64(v.c1, m)};
            v = gl::add(v, o);
        }
        if (lane == 0) red[wave][z] = v;
    }
    __syncthreads();
    if (threadIdx.x <= D) {
        fp2 s = red[0][threadIdx.x];
        for (int w = 1; w < SC_THREADS / 64; w++) s = gl::add(s, red[w][threadIdx.x]);
        partials[(uint64_t)blockIdx.x * (D + 1) + threadIdx.x] = s;
    }
}

// Node sums: x y added into an Fp2 sum, or into a wide sum reduced later (gl::wide2).
__device__ __forceinline__ void accumulate(fp2 &acc, const fp2 &x, const fp2 &y) { acc = gl::add(acc, gl::mul(x, y)); }
__device__ __forceinline__ void accumulate(gl::wide2 &acc, const fp2 &x, const fp2 &y) { gl::mac(acc, x, y); }
__device__ __forceinline__ void accumulate_base(fp2 &acc, const fp2 &x, uint64_t y) {
    acc = gl::add(acc, gl::mul_base(x, y));
}
__device__ __forceinline__ void accumulate_base(gl::wide2 &acc, const fp2 &x, uint64_t y) { gl::mac_base(acc, x, y); }
__device__ __forceinline__ fp2 reduced(const fp2 &x) { return x; }
__device__ __forceinline__ fp2 reduced(const gl::wide2 &x) { return gl::reduce(x); }

template <int REL, bool UNIT, int NA, typename ACC = fp2>
__device__ __forceinline__ void sc_eq_nodes(const fp2 *a, const fp2 *b, const fp2 &e,
                                            const fp2 &lambda, ACC (&acc)[NA]);

// One item q of a level round in k_sc_round_eq's modes: binds the tables at q and/or adds E F
// at z = 0, 1, 2 to acc. LO: the pair's weight is eq_lo[q mod 2^eq_a] alone (the caller applies
// eq_hi[q >> eq_a] to the sums of a run of pairs that share it, or eq_hi is the single entry 1).
// NA = 6 (unit numerators only): acc[0..2] take E (D0 + D1) and acc[3..5] E D0 D1, so that
// lambda multiplies the sums instead of every pair.
template <int REL, bool UNIT, bool BASE, bool LO = false, int NA = 3, typename ACC = fp2>
__device__ __forceinline__ void sc_item_eq(const ScTables &t, uint32_t base_mask, bool first,
                                           bool bind, uint64_t pairs, const fp2 &r,
                                           const fp2 &lambda, const fp2 *__restrict__ eq_lo,
                                           const fp2 *__restrict__ eq_hi, uint32_t eq_a,
                                           uint64_t q, ACC (&acc)[NA]) {
    constexpr int F = REL == REL_PRODUCT ? 2 : 4;
    constexpr int K0 = (REL == REL_FRACTION && UNIT) ? 2 : 0;
    auto weight = [&](uint64_t x) {
        const fp2 lo = eq_lo[x & ((1ull << eq_a) - 1)];
        return LO ? lo : gl::mul(lo, eq_hi[x >> eq_a]);
    };
    fp2 a[F], b[F];
    if (BASE) {
        const uint64_t *s0 = (const uint64_t *)t.src[0], *s1 = (const uint64_t *)t.src[1];
        auto at0 = [&](uint64_t x) { return s0[(x << t.shift[0]) + t.off[0]]; };
        auto at1 = [&](uint64_t x) { return s1[(x << t.shift[1]) + t.off[1]]; };
        // e0 + r (e1 - e0) for base e0, e1
        auto bindb = [&](uint64_t e0, uint64_t e1) {
            fp2 v = gl::mul_base(r, gl::sub(e1, e0));
            v.c0 = gl::add(v.c0, e0);
            return v;
        };
        if (!bind) {
            const uint64_t a0 = at0(2 * q), b0 = at0(2 * q + 1), a1 = at1(2 * q), b1 = at1(2 * q + 1);
            const fp2 e = weight(q);
            const uint64_t f0 = gl::mul(a0, a1), f1 = gl::mul(b0, b1);
            const uint64_t f2 = gl::mul(gl::sub(gl::add(b0, b0), a0), gl::sub(gl::add(b1, b1), a1));
            accumulate_base(acc[0], e, f0);
            accumulate_base(acc[1], e, f1);
            accumulate_base(acc[2], e, f2);
            return;
        }
        if (pairs == 0) {
            t.dst[0][q] = bindb(at0(2 * q), at0(2 * q + 1));
            t.dst[1][q] = bindb(at1(2 * q), at1(2 * q + 1));
            return;
        }
        a[0] = bindb(at0(4 * q), at0(4 * q + 1));
        b[0] = bindb(at0(4 * q + 2), at0(4 * q + 3));
        a[1] = bindb(at1(4 * q), at1(4 * q + 1));
        b[1] = bindb(at1(4 * q + 2), at1(4 * q + 3));
#pragma unroll
        for (int k = 0; k < 2; k++) {
            t.dst[k][2 * q] = a[k];
            t.dst[k][2 * q + 1] = b[k];
        }
    }
    if (!BASE) {
        // every table's entries are loaded before any bound value is stored (the tables are
        // separate arrays, but the compiler cannot know it: a store between two tables' loads
        // would hold the second table's loads back)
        fp2 ent[F][4];
        const fp2 e = pairs > 0 ? weight(q) : fp2{0, 0};
#pragma unroll
        for (int k = K0; k < F; k++) {
            const bool base = first && ((base_mask >> k) & 1);
            const uint32_t sh = first ? t.shift[k] : 0, of = first ? t.off[k] : 0;
            auto at = [&](uint64_t x) { return ld_entry(t.src[k], base, (x << sh) + of); };
            if (bind && pairs > 0) {
#pragma unroll
                for (int m = 0; m < 4; m++) ent[k][m] = at(4 * q + m);
            } else {
                ent[k][0] = at(2 * q);
                ent[k][1] = at(2 * q + 1);
            }
        }
#pragma unroll
        for (int k = K0; k < F; k++) {
            if (!bind) {
                a[k] = ent[k][0];
                b[k] = ent[k][1];
            } else if (pairs > 0) {
                a[k] = gl::add(ent[k][0], gl::mul(r, gl::sub(ent[k][1], ent[k][0])));
                b[k] = gl::add(ent[k][2], gl::mul(r, gl::sub(ent[k][3], ent[k][2])));
            } else {
                a[k] = gl::add(ent[k][0], gl::mul(r, gl::sub(ent[k][1], ent[k][0])));
            }
        }
#pragma unroll
        for (int k = K0; k < F; k++) {
            if (!bind) break;
            if (pairs > 0) {
                t.dst[k][2 * q] = a[k];
                t.dst[k][2 * q + 1] = b[k];
            } else {
                t.dst[k][q] = a[k];
            }
        }
        if (pairs == 0) return;
        sc_eq_nodes<REL, UNIT, NA, ACC>(a, b, e, lambda, acc);
        return;
    }
    sc_eq_nodes<REL, UNIT, NA, ACC>(a, b, weight(q), lambda, acc);
}

// E F(z) at z = 0, 1, 2 of one pair added to acc (see k_sc_round_eq).
template <int REL, bool UNIT, int NA, typename ACC>
__device__ __forceinline__ void sc_eq_nodes(const fp2 *a, const fp2 *b, const fp2 &e,
                                            const fp2 &lambda, ACC (&acc)[NA]) {
    auto line2 = [](const fp2 &x, const fp2 &y) { return gl::sub(gl::add(y, y), x); };
    if constexpr (REL == REL_PRODUCT) {
        // E folded into T0's line; the three node products go straight into the sums
        const fp2 x0 = gl::mul(e, a[0]), x1 = gl::mul(e, b[0]);
        accumulate(acc[0], x0, a[1]);
        accumulate(acc[1], x1, b[1]);
        accumulate(acc[2], line2(x0, x1), line2(a[1], b[1]));
        return;
    } else if constexpr (NA == 6) {
        static_assert(REL == REL_FRACTION && UNIT, "split sums are for unit numerators");
        // E (D0 + D1) is linear in z: its sums at z = 0, 1 give the sum at z = 2 (the caller
        // forms 2 acc[1] - acc[0]; acc[2] stays unused); (E D0) D1 at z = 0, 1, 2
        const fp2 s0 = gl::add(a[2], a[3]), s1 = gl::add(b[2], b[3]);
        const fp2 x0 = gl::mul(e, a[2]), x1 = gl::mul(e, b[2]);
        accumulate(acc[0], e, s0);
        accumulate(acc[1], e, s1);
        accumulate(acc[3], x0, a[3]);
        accumulate(acc[4], x1, b[3]);
        accumulate(acc[5], line2(x0, x1), line2(a[3], b[3]));
        return;
    } else if constexpr (UNIT) {
        // E (D0 + D1) + (lambda E D0) D1
        const fp2 u = gl::mul(lambda, e);
        const fp2 s0 = gl::add(a[2], a[3]), s1 = gl::add(b[2], b[3]);
        const fp2 p0 = gl::mul(u, a[2]), p1 = gl::mul(u, b[2]);
        accumulate(acc[0], e, s0);
        accumulate(acc[0], p0, a[3]);
        accumulate(acc[1], e, s1);
        accumulate(acc[1], p1, b[3]);
        accumulate(acc[2], e, line2(s0, s1));
        accumulate(acc[2], line2(p0, p1), line2(a[3], b[3]));
    } else {
        // (E N0) D1 + (E D0)(N1 + lambda D1)
        const fp2 n0 = gl::mul(e, a[0]), n1 = gl::mul(e, b[0]);
        const fp2 d0 = gl::mul(e, a[2]), d1 = gl::mul(e, b[2]);
        const fp2 w0 = gl::add(a[1], gl::mul(lambda, a[3])), w1 = gl::add(b[1], gl::mul(lambda, b[3]));
        accumulate(acc[0], n0, a[3]);
        accumulate(acc[0], d0, w0);
        accumulate(acc[1], n1, b[3]);
        accumulate(acc[1], d1, w1);
        accumulate(acc[2], line2(n0, n1), line2(a[3], b[3]));
        accumulate(acc[2], line2(d0, d1), line2(w0, w1));
    }
}

// The first round's sums of a fraction level with stored numerators, from the level it proves:
// F(z) at z = 0, 1 is n + lambda d of the parent nodes 2q, 2q + 1 (the level's values are
// N0 D1 + N1 D0 and D0 D1 of these very tables), and only z = 2 multiplies the tables' line
// values: 8 Fp2 products per pair instead of 12.
template <bool LO, typename ACC = fp2>
__device__ __forceinline__ void sc_item_eq_parents(const ScTables &t, uint32_t base_mask,
                                                   const fp2 &lambda, const fp2 *__restrict__ eq_lo,
                                                   const fp2 *__restrict__ eq_hi, uint32_t eq_a,
                                                   const fp2 *__restrict__ pn,
                                                   const fp2 *__restrict__ pd, uint64_t q,
                                                   ACC (&acc)[3]) {
    auto line2 = [](const fp2 &x, const fp2 &y) { return gl::sub(gl::add(y, y), x); };
    fp2 x2[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const bool base = (base_mask >> k) & 1;
        const fp2 a = ld_entry(t.src[k], base, ((2 * q) << t.shift[k]) + t.off[k]);
        const fp2 b = ld_entry(t.src[k], base, ((2 * q + 1) << t.shift[k]) + t.off[k]);
        x2[k] = line2(a, b);
    }
    const fp2 n0 = pn[2 * q], n1 = pn[2 * q + 1], d0 = pd[2 * q], d1 = pd[2 * q + 1];
    const fp2 lo = eq_lo[q & ((1ull << eq_a) - 1)];
    const fp2 e = LO ? lo : gl::mul(lo, eq_hi[q >> eq_a]);
    const fp2 f2 = gl::add(gl::mul(x2[0], x2[3]), gl::mul(x2[2], gl::add(x2[1], gl::mul(lambda, x2[3]))));
    accumulate(acc[0], e, gl::add(n0, gl::mul(lambda, d0)));
    accumulate(acc[1], e, gl::add(n1, gl::mul(lambda, d1)));
    accumulate(acc[2], e, f2);
}

// One own round of a level instance whose relation is eq(r, x) F(x) with eq a factor, not a
// table (degree 3). Modes as k_sc_round's. Pair q of the round is weighted by
// E[q] = eq_lo[q mod 2^eq_a] eq_hi[q >> eq_a], the eq of the point on the variables above the one
// the round binds; the transcript kernel multiplies in the eq of the bound variables and the
// line of the current one. Without that line, E F(z) is quadratic in z: each pair gives its
// values at z = 0, 1, 2, and each block's node 3 follows as f(3) = f(0) + 3 (f(2) - f(1)).
//   REL_PRODUCT:  F = T0 T1 (tables T0, T1), E folded into T0's line
//   REL_FRACTION: F = N0 D1 + N1 D0 + lambda D0 D1 (tables N0, N1, D0, D1), computed as
//                 (E N0) D1 + (E D0)(N1 + lambda D1); UNIT (N0 = N1 = 1, tables 0 and 1 absent):
//                 E (D0 + D1) + (lambda E D0) D1
// Line values: X(0) = a, X(1) = b, X(2) = 2b - a.
// BASE (REL_PRODUCT, first round of a level, both tables base, as a forest's level values): the
// sums multiply in the base field and scale by E once per node (3 base products + 3 Fp2 x Fp
// per pair instead of 6 Fp2 products), the binding takes 2 base products per entry.
// LO (rounds with sums; eq_hi a single entry 1, or 2^eq_a >= SC_THREADS): block b takes the
// pairs [b c, (b + 1) c), c a multiple of SC_THREADS, so that the pairs of one pass of the block
// share eq_hi[q >> eq_a]. A pair is weighted by eq_lo alone and each run of pairs with one eq_hi
// entry scales its sums by it once: one Fp2 product less per pair. Unit numerators also keep
// E (D0 + D1) and E D0 D1 apart and multiply by lambda once per thread.
// BIND: out_len > 0 (a round without binding holds half the loaded entries: fewer registers).
// PAR (a first round of a fraction level with stored numerators, no binding): the sums at
// z = 0, 1 from the parent level's values pn, pd (sc_item_eq_parents).
template <int REL, bool UNIT, bool BASE = false, bool LO = false, bool BIND = true, bool PAR = false>
__global__ void __launch_bounds__(SC_THREADS)
k_sc_round_eq(ScTables t, uint32_t base_mask, bool first, uint64_t out_len, uint64_t pairs,
              const fp2 *__restrict__ r_ptr, fp2 lambda, fp2 *__restrict__ partials,
              const fp2 *__restrict__ eq_lo, const fp2 *__restrict__ eq_hi, uint32_t eq_a,
              const fp2 *__restrict__ pn, const fp2 *__restrict__ pd) {
    static_assert(!PAR || (REL == REL_FRACTION && !UNIT && !BASE && !BIND), "parents: fraction sums");
    constexpr bool bind = BIND;
    const fp2 r = bind ? *r_ptr : fp2{0, 0};
    const uint64_t items = pairs > 0 ? pairs : out_len;
    fp2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};

    if constexpr (!LO) {
        const uint64_t stride = (uint64_t)gridDim.x * SC_THREADS;
        for (uint64_t q = (uint64_t)blockIdx.x * SC_THREADS + threadIdx.x; q < items; q += stride) {
            if constexpr (PAR)
                sc_item_eq_parents<false>(t, base_mask, lambda, eq_lo, eq_hi, eq_a, pn, pd, q, acc);
            else
                sc_item_eq<REL, UNIT, BASE>(t, base_mask, first, bind, pairs, r, lambda, eq_lo, eq_hi, eq_a, q, acc);
        }
        if (pairs == 0) return;
    } else {
        constexpr int NA = (REL == REL_FRACTION && UNIT) ? 6 : 3;
        // the run sums are wide: one reduction per run instead of one per product
        using RunT = gl::wide2;
        const uint64_t per = (pairs + gridDim.x - 1) / gridDim.x;
        const uint64_t chunk = (per + SC_THREADS - 1) / SC_THREADS * SC_THREADS;
        const uint64_t q0 = (uint64_t)blockIdx.x * chunk;
        const uint64_t q1 = q0 + chunk < pairs ? q0 + chunk : pairs;
        const bool hi_one = (pairs >> eq_a) <= 1;
        // a run's sums, folded into acc (times lambda for the E D0 D1 half, times its eq_hi)
        RunT run[NA];
#pragma unroll
        for (int z = 0; z < NA; z++) run[z] = RunT{};
        auto fold = [&](const fp2 *w) {
            fp2 lin[3];  // NA = 6: the linear part's node 2 from its nodes 0 and 1 (sc_eq_nodes)
            if constexpr (NA == 6) {
                lin[0] = reduced(run[0]);
                lin[1] = reduced(run[1]);
                lin[2] = gl::sub(gl::add(lin[1], lin[1]), lin[0]);
            }
#pragma unroll
            for (int z = 0; z < 3; z++) {
                fp2 v;
                if constexpr (NA == 6)
                    v = gl::add(lin[z], gl::mul(lambda, reduced(run[3 + z])));
                else
                    v = reduced(run[z]);
                acc[z] = gl::add(acc[z], w ? gl::mul(*w, v) : v);
            }
#pragma unroll
            for (int z = 0; z < NA; z++) run[z] = RunT{};
        };
        uint64_t h = q0 >> eq_a;
#pragma unroll 1
        for (uint64_t qb = q0; qb < q1; qb += SC_THREADS) {
            if (!hi_one && (qb >> eq_a) != h) {
                const fp2 w = eq_hi[h];
                fold(&w);
                h = qb >> eq_a;
            }
            const uint64_t q = qb + threadIdx.x;
            if (q < q1) {
                if constexpr (PAR)
                    sc_item_eq_parents<true, RunT>(t, base_mask, lambda, eq_lo, eq_hi, eq_a, pn, pd, q, run);
                else
                    sc_item_eq<REL, UNIT, BASE, true, NA, RunT>(t, base_mask, first, bind, pairs, r, lambda, eq_lo, eq_hi, eq_a, q, run);
            }
        }
        if (q0 < q1) {
            if (hi_one) {
                fold(nullptr);
            } else {
                const fp2 w = eq_hi[h];
                fold(&w);
            }
        }
    }

    __shared__ fp2 red[SC_THREADS / 64][3];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int z = 0; z < 3; z++) {
        fp2 v = acc[z];
        for (int m = 32; m >= 1; m >>= 1) {
            fp2 o{shfl_xor_u64(v.c0, m), shfl_xor_u64(v.c1, m)};
            v = gl::add(v, o);
        }
        if (lane == 0) red[wave][z] = v;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        fp2 f[3];
        for (int z = 0; z < 3; z++) {
            f[z] = red[0][z];
            for (int w = 1; w < SC_THREADS / 64; w++) f[z] = gl::add(f[z], red[w][z]);
        }
        const fp2 d = gl::sub(f[2], f[1]);
        fp2 *out = partials + (uint64_t)blockIdx.x * 4;  // slots z = 0, 2, 3, 1
        out[0] = f[0];
        out[1] = f[2];
        out[2] = gl::add(f[0], gl::add(gl::add(d, d), d));
        out[3] = f[1];
    }
}

// The facts of one instance in one round (a table of rounds x instances, built on the host).
struct ScRoundInst {
    uint32_t nblk;        // blocks of partials this round's message reads
    uint32_t pad;         // padding rounds in front of the instance's own rounds
    uint32_t eqf;         // 1: eq(point, x) is a factor of the relation (see k_sc_round_eq)
    uint32_t wsum_given;  // 1: wsum set by the host (claims, or instances on no variable)
    fp2 weight;           // gamma^i
    fp2 rt;               // eqf and an own round t: coordinate t of the eq point
};

// The per-instance state the rounds update.
static_assert(sizeof(ScRoundInst) % 16 == 0, "round tables are staged in LDS as 16-byte words");

struct ScInstState {
    fp2 wsum;  // gamma^i S_i
    fp2 eq_k;  // eqf: gamma^i times the product over the bound own rounds t of eq(point_t, c_t)
};

// Batches of up to SC_STAGED instances have their facts staged in LDS by the transcript kernel.
constexpr uint32_t SC_STAGED = 64;
// The transcript kernel's LDS: its static arrays and count x (D+1) sums (one workgroup).
constexpr size_t SC_TRANSCRIPT_LDS = 64 * 1024;
constexpr size_t SC_TRANSCRIPT_STATIC = 8 * 1024;

// W L(x) for the eq line of round t: W ((1 - r_t)(1 - x) + r_t x) at the node x (additions only).
__device__ __forceinline__ fp2 eq_line_at(const fp2 &w, const fp2 &rt, uint32_t node) {
    const fp2 l0 = gl::sub(fp2{1, 0}, rt), step = gl::sub(rt, l0);
    fp2 l = l0;
    for (uint32_t n = 0; n < node; n++) l = gl::add(l, step);
    return gl::mul(w, l);
}

// eq(r_t, c) = r_t c + (1 - r_t)(1 - c) = 1 - r_t - c + 2 r_t c (one product).
__device__ __forceinline__ fp2 eq_at(const fp2 &rt, const fp2 &c) {
    const fp2 m = gl::mul(rt, c);
    return gl::sub(gl::sub(gl::add(fp2{1, 0}, gl::add(m, m)), rt), c);
}

// gamma^i S_i of a padded instance without a claim, from the sums of its first own round
// (slot 0: node 0, slot D: node 1): p(0) + p(1), with an eq factor's line (1 - r_0, r_0) at the
// two nodes (the host puts coordinate 0 of its point in rt for round 0).
__device__ __forceinline__ fp2 first_wsum(const ScRoundInst &in, const fp2 *s, uint32_t D) {
    if (!in.eqf) return gl::mul(in.weight, gl::add(s[0], s[D]));
    return gl::add(gl::mul(eq_line_at(in.weight, in.rt, 0), s[0]),
                   gl::mul(eq_line_at(in.weight, in.rt, 1), s[D]));
}

struct ScCtl {
    DevTranscript t;
    fp2 r;  // the current round's challenge
};

// Round j: reduce the partial sums, form the message, absorb, squeeze r_j.
// Instance i's partials are partials[(i * SC_MAX_BLOCKS + b) * (D + 1) + slot] for
// b < ri[i].nblk <= nblk_max. Every global load the kernel makes has an address known at launch
// (the round's instance table, the instance states, the partials up to nblk_max, the labels, the
// transcript state), so all of them are in flight at once; the transcript runs on wave 0 from LDS
// and registers.
template <int D>
__global__ void __launch_bounds__(SC_THREADS)
k_sc_transcript(uint32_t j, const ScRoundInst *__restrict__ ri, ScInstState *__restrict__ ist,
                uint32_t count, uint32_t nblk_max, const fp2 *__restrict__ partials,
                ScCtl *__restrict__ ctl, DevKeys keys, const uint8_t *__restrict__ labels,
                const uint32_t *__restrict__ label_lens, fp2 *__restrict__ rounds_out,
                fp2 *__restrict__ chals_out) {
    extern __shared__ fp2 sums[];  // count x (D+1)
    __shared__ ScRoundInst sri[SC_STAGED];
    __shared__ ScInstState sis[SC_STAGED];
    __shared__ uint32_t hbuf[DEV_BUF_WORDS];
    __shared__ uint32_t lab[2 * LABEL_SLOT / 4];  // the labels of round j (absorb, then squeeze)
    __shared__ uint32_t st[8], outw[8];
    __shared__ fp2 msg[D];
    __shared__ fp2 chal;
    __shared__ WaveKeys wk;
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // Loads into registers first, unconditionally (count >= 1; addresses clamped), LDS stores
    // after the partials' loads are issued: no wave waits on memory before the reduction.
    const uint32_t ci = threadIdx.x < count ? threadIdx.x : count - 1;
    const ScRoundInst my = ri[ci];
    const ScInstState mys = ist[ci];
    // vector loads (lane-indexed) only: a scalar load from global memory would make the first
    // wait on the kernel arguments (scalar loads complete out of order) wait for it too
    const uint32_t t_word = ((const uint32_t *)&ctl->t)[lane % 10];  // state[0..8), counter
    const uint32_t len_word = label_lens[2 * j + (lane & 1)];
    const uint32_t key_word = (&keys.absorb[0])[lane < 24 ? lane : 23];
    const uint32_t lab_word = ((const uint32_t *)(labels + (size_t)2 * j * LABEL_SLOT))[threadIdx.x & 31];
    // each wave reduces slots x = wave, wave + 4, ...: lane l loads blocks l, l + 64, ... below
    // nblk_max (all at once), keeps those below the instance's own count, then a shuffle tree
    // adds the 64 lane sums (field addition: any order gives the same sum)
    constexpr int PER_LANE = SC_MAX_BLOCKS / 64;
    for (uint32_t x = wave; x < count * (D + 1); x += SC_THREADS / 64) {
        const uint32_t i = x / (D + 1), z = x % (D + 1);
        const fp2 *p = partials + (uint64_t)i * SC_MAX_BLOCKS * (D + 1) + z;
        const uint32_t nb = ri[i].nblk;
        fp2 v[PER_LANE];
#pragma unroll
        for (int k = 0; k < PER_LANE; k++) {
            const uint32_t b = lane + 64 * k;
            v[k] = b < nblk_max ? p[(uint64_t)b * (D + 1)] : fp2{0, 0};
        }
#pragma unroll
        for (int k = 0; k < PER_LANE; k++)
            if (lane + 64 * k >= nb) v[k] = fp2{0, 0};
        if (nblk_max > 64) {  // uniform: more than one partial per lane
#pragma unroll
            for (int w = PER_LANE / 2; w >= 1; w /= 2)
#pragma unroll
                for (int k = 0; k < w; k++) v[k] = gl::add(v[k], v[k + w]);
        }
        fp2 s = v[0];
        // lanes at or above nblk_max hold zero: only the steps that reach a lower lane
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            if ((uint32_t)off >= nblk_max) continue;
            fp2 o;
            o.c0 = __shfl_xor(s.c0, off, 64);
            o.c1 = __shfl_xor(s.c1, off, 64);
            s = gl::add(s, o);
        }
        if (lane == 0) sums[x] = s;
    }
    if (wave == 0 && lane < 8) st[lane] = t_word;
    if (wave == 0 && lane < 24) (&wk.absorb[0])[lane] = key_word;
    if (threadIdx.x < 2 * LABEL_SLOT / 4) lab[threadIdx.x] = lab_word;
    if (count <= SC_STAGED && threadIdx.x < count) {
        sri[threadIdx.x] = my;
        sis[threadIdx.x] = mys;
    }
    __syncthreads();
    // S_i of the padded instances that have no claim: p(0) + p(1) of their first sums
    auto first_sum = [&](uint32_t i, const ScRoundInst &in, ScInstState &is) {
        if (j != 0 || in.pad == 0 || in.wsum_given) return;
        is.wsum = first_wsum(in, sums + (uint64_t)i * (D + 1), D);
        ist[i].wsum = is.wsum;
    };
    // each instance's share of the message, in place of its sums at the nodes 0, 2, ..., D (the
    // message slots): weight (times K L(node) for an eq factor) times the sum, or its padding
    // share 2^(pad-j-1) gamma^i S_i
    auto share = [&](uint32_t i, uint32_t z, const ScRoundInst &in, const ScInstState &is) {
        fp2 *s = sums + (uint64_t)i * (D + 1) + z;
        if (j < in.pad) {
            *s = gl::mul_base(is.wsum, 1ull << (in.pad - j - 1));
            return;
        }
        // weight, or for an eq factor gamma^i K times the eq line of this round at the node
        const fp2 f = in.eqf ? eq_line_at(is.eq_k, in.rt, z == 0 ? 0 : z + 1) : in.weight;
        *s = gl::mul(*s, f);
    };
    if (count <= SC_STAGED) {
        // the facts staged in LDS; one thread per instance and slot (each reads its own slot)
        if (threadIdx.x < count) first_sum(threadIdx.x, sri[threadIdx.x], sis[threadIdx.x]);
        __syncthreads();
        for (uint32_t x = threadIdx.x; x < count * D; x += SC_THREADS)
            share(x / D, x % D, sri[x / D], sis[x / D]);
    } else {
        for (uint32_t i = threadIdx.x; i < count; i += SC_THREADS) {
            const ScRoundInst in = ri[i];
            ScInstState is = ist[i];
            first_sum(i, in, is);
            for (uint32_t z = 0; z < D; z++) share(i, z, in, is);
        }
    }
    __syncthreads();
    if (wave == 0) {
        if (lane < D) {
            fp2 m{0, 0};
            for (uint32_t i = 0; i < count; i++) m = gl::add(m, sums[(uint64_t)i * (D + 1) + lane]);
            msg[lane] = m;
            rounds_out[(uint64_t)j * D + lane] = m;
        }
        wave_sync();
        uint64_t counter = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)t_word, 9) << 32) |
                           (uint32_t)__builtin_amdgcn_readlane((int)t_word, 8);
        const uint32_t len_absorb = (uint32_t)__builtin_amdgcn_readlane((int)len_word, 0);
        const uint32_t len_squeeze = (uint32_t)__builtin_amdgcn_readlane((int)len_word, 1);
        const uint8_t *lb = (const uint8_t *)lab;
        wave_absorb_fp2s(st, wk, lb, len_absorb, msg, D, hbuf);
        const fp2 r = wave_challenge_fp2(st, counter, wk, lb + LABEL_SLOT, len_squeeze, hbuf, outw);
        if (lane < 8) ctl->t.state[lane] = st[lane];
        if (lane == 0) {
            ctl->t.counter = counter;
            ctl->r = r;
            chal = r;
            chals_out[j] = r;
        }
    }
    __syncthreads();
    const fp2 r = chal;
    // gamma^i K *= eq(r_t, c)
    auto fold = [&](uint32_t i, const ScRoundInst &in, const fp2 &k) {
        if (!in.eqf || j < in.pad) return;
        ist[i].eq_k = gl::mul(k, eq_at(in.rt, r));
    };
    if (threadIdx.x < count) fold(threadIdx.x, my, mys.eq_k);
    for (uint32_t i = threadIdx.x + SC_THREADS; i < count; i += SC_THREADS) fold(i, ri[i], ist[i].eq_k);
}

// ---------------------------------------------------------------------------------------------
// The tail of a batch: once every remaining round has at most SC_TAIL_PAIRS pairs per instance,
// one workgroup runs all of them (the message of the first from the partials of the grid
// launches, then per round: the instances' bind + sums in the block, the shares, the hashes on
// wave 0, the eq folds), the final binds and the gather of the final evaluations. Nothing leaves
// the workgroup between rounds: no kernel boundary and no trip through another XCD's cache per
// round.

constexpr uint32_t SC_TAIL_LOG = 8;
constexpr uint64_t SC_TAIL_PAIRS = 1ull << SC_TAIL_LOG;
constexpr int SC_TAIL_THREADS = 512;  // the instances' items run side by side in thread groups
constexpr uint32_t SC_TAIL_MAX = 8;   // instances a tail takes (its steps are staged in LDS)

enum ScTailKind : uint32_t {
    TAIL_NONE = 0,        // padding round (or an instance on no variable at the final bind)
    TAIL_PRODUCT,         // k_sc_round<D, REL_PRODUCT>
    TAIL_TERMS,           // k_sc_round<D, REL_TERMS>
    TAIL_EQ_PRODUCT,      // k_sc_round_eq<REL_PRODUCT, false, false>
    TAIL_EQ_PRODUCT_BASE, // k_sc_round_eq<REL_PRODUCT, false, true>
    TAIL_EQ_UNIT,         // k_sc_round_eq<REL_FRACTION, true, false>
    TAIL_EQ_FRACTION,     // k_sc_round_eq<REL_FRACTION, false, false>
};

// One instance's work in one tail round (or the final bind): as the grid launch would do it.
struct ScTailStep {
    ScTables t;
    const fp2 *eq_lo, *eq_hi;
    uint32_t eq_a, kind, F, base_mask;
    uint32_t first, pad_;
    uint64_t out_len, pairs;
};

static_assert(sizeof(ScTailStep) % 16 == 0, "tail steps are staged in LDS as 16-byte words");

// The final evaluations: out[e] = entry off[e] of the buffer at ptrs[e] (base or Fp2), e < n.
struct ScFinals {
    const void *const *ptrs;
    const uint32_t *base, *off;
    fp2 *out;
    uint32_t n;
};

// A level round's pairs in the tail, eight lanes per pair: lane 2k + h of the pair binds (or
// loads) half h of table K0 + k and leaves it in wbuf; lanes z = 0, 1, 2 then take the line
// values at z, E and the relation, so one pair's products are spread over four lanes and no
// single lane runs the whole chain (a wave executes an instruction in four cycles whatever the
// number of its active lanes). acc: slots 0, 2, 3, 1 (node 3 = f(0) - 3 f(1) + 3 f(2)).
template <int REL, bool UNIT>
__device__ __forceinline__ void tail_eq_wide(const ScTailStep &sp, const fp2 &r, const fp2 &lambda,
                                             uint32_t gt, uint32_t G, fp2 *wbuf, fp2 (&acc)[4]) {
    constexpr int F = REL == REL_PRODUCT ? 2 : 4;
    constexpr int K0 = (REL == REL_FRACTION && UNIT) ? 2 : 0;
    constexpr uint32_t NV = 2 * (F - K0);
    const bool bind = sp.out_len > 0, first = sp.first != 0;
    const uint32_t role = gt & 7, slots = G / 8;
    fp2 *my = wbuf + (threadIdx.x & ~7u);
    for (uint64_t q = gt >> 3; q < sp.pairs; q += slots) {
        // the pair's eq weight, loaded with the tables' entries (the wave barrier below would
        // hold a load placed after it back by a whole memory latency)
        const fp2 e = role < 3 ? sp.eq_lo[q] : fp2{0, 0};  // eq_hi is [1] in the tail's rounds
        if (role < NV) {
            const uint32_t k = K0 + (role >> 1), h = role & 1;
            const bool base = first && ((sp.base_mask >> k) & 1);
            const uint32_t sh = first ? sp.t.shift[k] : 0, of = first ? sp.t.off[k] : 0;
            const void *src = sp.t.src[k];
            fp2 v;
            if (!bind) {
                v = ld_entry(src, base, ((2 * q + h) << sh) + of);
            } else {
                const fp2 e0 = ld_entry(src, base, ((4 * q + 2 * h) << sh) + of);
                const fp2 e1 = ld_entry(src, base, ((4 * q + 2 * h + 1) << sh) + of);
                v = gl::add(e0, gl::mul(r, gl::sub(e1, e0)));
                sp.t.dst[k][2 * q + h] = v;
            }
            my[role] = v;
        }
        wave_sync();
        if (role < 3) {
            fp2 x[F];
#pragma unroll
            for (int k = K0; k < F; k++) {
                const fp2 a = my[2 * (k - K0)], b = my[2 * (k - K0) + 1];
                x[k] = role == 0 ? a : role == 1 ? b : gl::sub(gl::add(b, b), a);
            }
            fp2 f;
            if (REL == REL_PRODUCT)
                f = gl::mul(gl::mul(e, x[0]), x[1]);
            else if (UNIT)
                f = gl::mul(e, gl::add(gl::add(x[2], x[3]), gl::mul(lambda, gl::mul(x[2], x[3]))));
            else
                f = gl::mul(e, gl::add(gl::add(gl::mul(x[0], x[3]), gl::mul(x[1], x[2])),
                                       gl::mul(lambda, gl::mul(x[2], x[3]))));
            const fp2 f3 = gl::add(gl::add(f, f), f);
            if (role == 0) {
                acc[0] = gl::add(acc[0], f);
                acc[2] = gl::add(acc[2], f);
            } else if (role == 1) {
                acc[3] = gl::add(acc[3], f);
                acc[2] = gl::sub(acc[2], f3);
            } else {
                acc[1] = gl::add(acc[1], f);
                acc[2] = gl::add(acc[2], f3);
            }
        }
        wave_sync();  // the pair's values are read before the next pair's overwrite them
    }
}

// This thread's items of one step, summed into acc (slots 0, 2..D, 1).
template <int D>
__device__ __forceinline__ void tail_items(const ScTailStep &sp, const ScTerms &terms, const fp2 &r,
                                           const fp2 &lambda, uint32_t q0, uint32_t stride,
                                           fp2 *wbuf, fp2 (&acc)[D + 1]) {
    const bool bind = sp.out_len > 0, first = sp.first != 0;
    const uint64_t items = sp.pairs > 0 ? sp.pairs : sp.out_len;
    if constexpr (D == 3) {
        // level rounds with sums whose pairs fit the group eight lanes each (more pairs: one
        // lane per pair, the group is busy either way)
        if (sp.pairs > 0 && sp.pairs * 8 <= stride) {
            switch (sp.kind) {
            case TAIL_EQ_PRODUCT:
            case TAIL_EQ_PRODUCT_BASE:
                tail_eq_wide<REL_PRODUCT, false>(sp, r, lambda, q0, stride, wbuf, acc);
                return;
            case TAIL_EQ_UNIT:
                tail_eq_wide<REL_FRACTION, true>(sp, r, lambda, q0, stride, wbuf, acc);
                return;
            case TAIL_EQ_FRACTION:
                tail_eq_wide<REL_FRACTION, false>(sp, r, lambda, q0, stride, wbuf, acc);
                return;
            default:
                break;
            }
        }
    }
    switch (sp.kind) {
    case TAIL_PRODUCT:
        if (sp.F == D) {  // a product of D tables: the tables known at compile time
            for (uint64_t q = q0; q < items; q += stride)
                sc_item<D, REL_PRODUCT, D>(sp.t, sp.F, sp.base_mask, first, bind, sp.pairs, r, lambda, terms, q, acc);
            return;
        }
        for (uint64_t q = q0; q < items; q += stride)
            sc_item<D, REL_PRODUCT>(sp.t, sp.F, sp.base_mask, first, bind, sp.pairs, r, lambda, terms, q, acc);
        return;
    case TAIL_TERMS:
        for (uint64_t q = q0; q < items; q += stride)
            sc_item<D, REL_TERMS>(sp.t, sp.F, sp.base_mask, first, bind, sp.pairs, r, lambda, terms, q, acc);
        return;
    default:
        break;
    }
    if constexpr (D == 3) {
        fp2 a3[3] = {{0, 0}, {0, 0}, {0, 0}};
        switch (sp.kind) {
        case TAIL_EQ_PRODUCT:
            for (uint64_t q = q0; q < items; q += stride)
                sc_item_eq<REL_PRODUCT, false, false, true>(sp.t, sp.base_mask, first, bind, sp.pairs, r, lambda, sp.eq_lo, sp.eq_hi, sp.eq_a, q, a3);
            break;
        case TAIL_EQ_PRODUCT_BASE:
            for (uint64_t q = q0; q < items; q += stride)
                sc_item_eq<REL_PRODUCT, false, true, true>(sp.t, sp.base_mask, first, bind, sp.pairs, r, lambda, sp.eq_lo, sp.eq_hi, sp.eq_a, q, a3);
            break;
        case TAIL_EQ_UNIT:
            for (uint64_t q = q0; q < items; q += stride)
                sc_item_eq<REL_FRACTION, true, false, true>(sp.t, sp.base_mask, first, bind, sp.pairs, r, lambda, sp.eq_lo, sp.eq_hi, sp.eq_a, q, a3);
            break;
        case TAIL_EQ_FRACTION:
            for (uint64_t q = q0; q < items; q += stride)
                sc_item_eq<REL_FRACTION, false, false, true>(sp.t, sp.base_mask, first, bind, sp.pairs, r, lambda, sp.eq_lo, sp.eq_hi, sp.eq_a, q, a3);
            break;
        default:
            return;
        }
        // E F is quadratic in the bound variable: node 3 = f(0) + 3 (f(2) - f(1)) (linear, so
        // per thread as well as per block)
        const fp2 d = gl::sub(a3[2], a3[1]);
        acc[0] = gl::add(acc[0], a3[0]);
        acc[1] = gl::add(acc[1], a3[2]);
        acc[2] = gl::add(acc[2], gl::add(a3[0], gl::add(gl::add(d, d), d)));
        acc[3] = gl::add(acc[3], a3[1]);
    }
}

// One workgroup of SC_TAIL_THREADS: instance i's items run on the thread group
// [i G, (i + 1) G) (G = SC_TAIL_THREADS / the power of two >= count), so the instances' chains of
// products overlap; the steps, the rounds' instance tables and their labels are staged in LDS at
// the start.
template <int D>
__global__ void __launch_bounds__(SC_TAIL_THREADS)
k_sc_tail(uint32_t j0, uint32_t n_max, const ScRoundInst *__restrict__ rtab,
          ScInstState *__restrict__ ist, uint32_t count, uint32_t nblk_max0,
          const fp2 *__restrict__ partials, const ScTailStep *__restrict__ steps_g,
          const ScTerms *__restrict__ terms, fp2 lambda, ScCtl *__restrict__ ctl, DevKeys keys,
          const uint8_t *__restrict__ labels, const uint32_t *__restrict__ label_lens,
          fp2 *__restrict__ rounds_out, fp2 *__restrict__ chals_out, ScFinals fin) {
    extern __shared__ uint4 dyn[];  // the steps, then count x (D+1) sums
    constexpr int WAVES = SC_TAIL_THREADS / 64;
    __shared__ ScRoundInst sri[SC_TAIL_MAX];
    __shared__ ScInstState sis[SC_TAIL_MAX];
    __shared__ uint32_t hbuf[DEV_BUF_WORDS];
    __shared__ uint32_t lab[2 * LABEL_SLOT / 4];
    __shared__ uint32_t st[8], outw[8];
    __shared__ fp2 msg[D];
    __shared__ fp2 chal;
    __shared__ WaveKeys wk;
    __shared__ fp2 red[WAVES][D + 1];
    __shared__ fp2 wbuf[SC_TAIL_THREADS];  // eight lanes per pair: the pair's bound values
    const uint32_t n_steps = (n_max - j0) * count, nr_t = n_max - j0;
    ScTailStep *steps = (ScTailStep *)dyn;
    fp2 *sums = (fp2 *)(steps + n_steps);
    // the rounds' instance tables, labels and label lengths, staged with the steps, so a round
    // starts without a load from global memory
    ScRoundInst *rst = (ScRoundInst *)(sums + count * (D + 1));
    uint32_t *lst = (uint32_t *)(rst + (size_t)nr_t * count);
    uint32_t *nst = lst + (size_t)nr_t * (2 * LABEL_SLOT / 4);
    for (uint32_t w = threadIdx.x; w < n_steps * (sizeof(ScTailStep) / 16); w += SC_TAIL_THREADS)
        dyn[w] = ((const uint4 *)steps_g)[w];
    for (uint32_t w = threadIdx.x; w < nr_t * count * (sizeof(ScRoundInst) / 16); w += SC_TAIL_THREADS)
        ((uint4 *)rst)[w] = ((const uint4 *)(rtab + (size_t)j0 * count))[w];
    for (uint32_t w = threadIdx.x; w < nr_t * (2 * LABEL_SLOT / 4); w += SC_TAIL_THREADS)
        lst[w] = ((const uint32_t *)(labels + (size_t)2 * j0 * LABEL_SLOT))[w];
    for (uint32_t w = threadIdx.x; w < 2 * nr_t; w += SC_TAIL_THREADS) nst[w] = label_lens[2 * j0 + w];
    const uint32_t lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const uint32_t t_word = ((const uint32_t *)&ctl->t)[lane % 10];
    if (wave == 0 && lane < 8) st[lane] = t_word;
    if (wave == 0 && lane < 24) (&wk.absorb[0])[lane] = (&keys.absorb[0])[lane];
    if (threadIdx.x < count) sis[threadIdx.x] = ist[threadIdx.x];
    uint64_t counter = ((uint64_t)(uint32_t)__builtin_amdgcn_readlane((int)t_word, 9) << 32) |
                       (uint32_t)__builtin_amdgcn_readlane((int)t_word, 8);
    // the thread groups: G threads (whole waves) per instance
    uint32_t groups = 1;
    while (groups < count) groups <<= 1;
    const uint32_t G = SC_TAIL_THREADS / groups, gi = threadIdx.x / G, gt = threadIdx.x % G;
    const uint32_t gw = G / 64;  // waves per group
    fp2 r{0, 0};
    __syncthreads();  // the staged steps and round tables
    for (uint32_t j = j0; j < n_max; j++) {
        const ScRoundInst *ri = rst + (size_t)(j - j0) * count;
        if (threadIdx.x < count) sri[threadIdx.x] = ri[threadIdx.x];
        if (threadIdx.x < 2 * LABEL_SLOT / 4) lab[threadIdx.x] = lst[(j - j0) * (2 * LABEL_SLOT / 4) + threadIdx.x];
        const uint32_t len_word = nst[2 * (j - j0) + (lane & 1)];
        if (j == j0) {
            // the grid launches' partials, as k_sc_transcript reduces them
            for (uint32_t x = wave; x < count * (D + 1); x += WAVES) {
                const uint32_t i = x / (D + 1), z = x % (D + 1);
                const fp2 *p = partials + (uint64_t)i * SC_MAX_BLOCKS * (D + 1) + z;
                const uint32_t nb = ri[i].nblk;
                fp2 s{0, 0};
                for (uint32_t b = lane; b < nblk_max0; b += 64)
                    if (b < nb) s = gl::add(s, p[(uint64_t)b * (D + 1)]);
#pragma unroll
                for (int off = 32; off > 0; off >>= 1) {
                    fp2 o;
                    o.c0 = __shfl_xor(s.c0, off, 64);
                    o.c1 = __shfl_xor(s.c1, off, 64);
                    s = gl::add(s, o);
                }
                if (lane == 0) sums[x] = s;
            }
        } else {
            // bind at r_{j-1} and the sums of round j: instance gi on this thread's group
            fp2 acc[D + 1];
#pragma unroll
            for (int z = 0; z <= D; z++) acc[z] = fp2{0, 0};
            if (gi < count)
                tail_items<D>(steps[(size_t)(j - j0 - 1) * count + gi], terms[gi], r, lambda, gt, G, wbuf, acc);
#pragma unroll
            for (int z = 0; z <= D; z++) {
                fp2 v = acc[z];
                for (int m = 32; m >= 1; m >>= 1) {
                    fp2 o{shfl_xor_u64(v.c0, m), shfl_xor_u64(v.c1, m)};
                    v = gl::add(v, o);
                }
                if (lane == 0) red[wave][z] = v;
            }
            __syncthreads();
            for (uint32_t x = threadIdx.x; x < count * (D + 1); x += SC_TAIL_THREADS) {
                const uint32_t i = x / (D + 1), z = x % (D + 1);
                fp2 v{0, 0};
                for (uint32_t w = i * gw; w < (i + 1) * gw; w++) v = gl::add(v, red[w][z]);
                sums[x] = v;
            }
        }
        __syncthreads();
        // S_i of padded instances without a claim (round 0 only), then the shares in place
        if (j == 0) {
            if (threadIdx.x < count) {
                const uint32_t i = threadIdx.x;
                const ScRoundInst &in = sri[i];
                if (in.pad > 0 && !in.wsum_given) {
                    sis[i].wsum = first_wsum(in, sums + (uint64_t)i * (D + 1), D);
                }
            }
            __syncthreads();
        }
        for (uint32_t x = threadIdx.x; x < count * D; x += SC_TAIL_THREADS) {
            const uint32_t i = x / D, z = x % D;
            const ScRoundInst &in = sri[i];
            fp2 *s = sums + (uint64_t)i * (D + 1) + z;
            if (j < in.pad) {
                *s = gl::mul_base(sis[i].wsum, 1ull << (in.pad - j - 1));
            } else {
                const fp2 f = in.eqf ? eq_line_at(sis[i].eq_k, in.rt, z == 0 ? 0 : z + 1) : in.weight;
                *s = gl::mul(*s, f);
            }
        }
        __syncthreads();
        if (wave == 0) {
            if (lane < D) {
                fp2 m{0, 0};
                for (uint32_t i = 0; i < count; i++) m = gl::add(m, sums[(uint64_t)i * (D + 1) + lane]);
                msg[lane] = m;
                rounds_out[(uint64_t)j * D + lane] = m;
            }
            wave_sync();
            const uint32_t len_absorb = (uint32_t)__builtin_amdgcn_readlane((int)len_word, 0);
            const uint32_t len_squeeze = (uint32_t)__builtin_amdgcn_readlane((int)len_word, 1);
            const uint8_t *lb = (const uint8_t *)lab;
            wave_absorb_fp2s(st, wk, lb, len_absorb, msg, D, hbuf);
            const fp2 c = wave_challenge_fp2(st, counter, wk, lb + LABEL_SLOT, len_squeeze, hbuf, outw);
            if (lane == 0) {
                chal = c;
                chals_out[j] = c;
            }
        }
        __syncthreads();
        r = chal;
        if (threadIdx.x < count) {
            const ScRoundInst &in = sri[threadIdx.x];
            if (in.eqf && j >= in.pad) sis[threadIdx.x].eq_k = gl::mul(sis[threadIdx.x].eq_k, eq_at(in.rt, r));
        }
        __syncthreads();
    }
    // the last binds: every instance's last own round at r_{n_max-1}
    if (gi < count) {
        fp2 acc[D + 1];
#pragma unroll
        for (int z = 0; z <= D; z++) acc[z] = fp2{0, 0};
        tail_items<D>(steps[(size_t)(n_max - 1 - j0) * count + gi], terms[gi], r, lambda, gt, G, wbuf, acc);
    }
    if (threadIdx.x < count) ist[threadIdx.x] = sis[threadIdx.x];
    if (wave == 0) {
        if (lane < 8) ctl->t.state[lane] = st[lane];
        if (lane == 0) {
            ctl->t.counter = counter;
            ctl->r = r;
        }
    }
    // the final evaluations, from the last binds above (as k_sc_gather would, one launch less)
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < fin.n; e += SC_TAIL_THREADS)
        fin.out[e] = ld_entry(fin.ptrs[e], fin.base[e] != 0, fin.off[e]);
}

// Gathers each final evaluation: entry off[e] of the buffer at ptrs[e] (base or Fp2).
__global__ void k_sc_gather(const void *const *__restrict__ ptrs, const uint32_t *__restrict__ base,
                            const uint32_t *__restrict__ off, uint32_t n, fp2 *__restrict__ out) {
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e < n) out[e] = ld_entry(ptrs[e], base[e] != 0, off[e]);
}

// eq tables of segments of one point: table g holds eq(point[start_g .. start_g + n_g), x) for
// x < 2^n_g at out + off_g (blockIdx.y = g).
struct EqSeg {
    uint64_t off;
    uint32_t start, n;
};

__global__ void __launch_bounds__(256)
k_eq_segments(const fp2 *__restrict__ point, const EqSeg *__restrict__ segs, fp2 *__restrict__ out) {
    const EqSeg g = segs[blockIdx.y];
    const uint64_t len = 1ull << g.n;
    for (uint64_t x = (uint64_t)blockIdx.x * 256 + threadIdx.x; x < len; x += (uint64_t)gridDim.x * 256) {
        fp2 e{1, 0};
        for (uint32_t i = 0; i < g.n; i++) {
            const fp2 r = point[g.start + i];
            e = gl::mul(e, ((x >> i) & 1) ? r : gl::sub(fp2{1, 0}, r));
        }
        out[g.off + x] = e;
    }
}

// ---------------------------------------------------------------------------------------------

// The eq factor of one launch: the pair weights' half tables (lo == nullptr: none).
struct EqSplit {
    const fp2 *lo = nullptr, *hi = nullptr;
    uint32_t a = 0;
};

template <int D, int REL, int FT = 0>
static void launch_round(dim3 grid, hipStream_t s, const ScTables &t, uint32_t F,
                         uint32_t base_mask, bool first, uint64_t out_len, uint64_t pairs,
                         const fp2 *r, fp2 lambda, fp2 *partials, const ScTerms &terms) {
    hipLaunchKernelGGL((k_sc_round<D, REL, FT>), grid, dim3(SC_THREADS), 0, s, t, F, base_mask,
                       first, out_len, pairs, r, lambda, partials, terms);
}

template <int REL, bool UNIT, bool BASE = false>
static void launch_round_eq(dim3 grid, hipStream_t s, const ScTables &t, uint32_t base_mask,
                            bool first, uint64_t out_len, uint64_t pairs, const fp2 *r, fp2 lambda,
                            fp2 *partials, const EqSplit &eq, const fp2 *pn = nullptr,
                            const fp2 *pd = nullptr) {
    // LO when eq_hi is the single entry 1, or when its runs span whole passes of a block and
    // every thread has at least 8 pairs (one run's scaling, 3 products, then costs less than the
    // product per pair it saves)
    const bool lo = pairs > 0 && ((pairs >> eq.a) <= 1 ||
                                  ((1ull << eq.a) >= SC_THREADS && pairs >= 8ull * grid.x * SC_THREADS));
#define TSR_ROUND_EQ(LO_, BIND_, PAR_)                                                              \
    hipLaunchKernelGGL((k_sc_round_eq<REL, UNIT, BASE, LO_, BIND_, PAR_>), grid, dim3(SC_THREADS), 0, s, \
                       t, base_mask, first, out_len, pairs, r, lambda, partials, eq.lo, eq.hi, eq.a, \
                       pn, pd)
    if constexpr (REL == REL_FRACTION && !UNIT && !BASE) {
        if (pn && pd && first && out_len == 0 && pairs > 0) {
            if (lo)
                TSR_ROUND_EQ(true, false, true);
            else
                TSR_ROUND_EQ(false, false, true);
            return;
        }
    }
    if (lo && out_len == 0)
        TSR_ROUND_EQ(true, false, false);
    else if (lo)
        TSR_ROUND_EQ(true, true, false);
    else if (out_len == 0)
        TSR_ROUND_EQ(false, false, false);
    else
        TSR_ROUND_EQ(false, true, false);
#undef TSR_ROUND_EQ
}

static void launch_round_d(uint32_t d, int rel, dim3 grid, hipStream_t s, const ScTables &t,
                           uint32_t F, uint32_t base_mask, bool first, uint64_t out_len,
                           uint64_t pairs, const fp2 *r, fp2 lambda, fp2 *partials,
                           const ScTerms &terms, const EqSplit &eq = EqSplit(),
                           const fp2 *pn = nullptr, const fp2 *pd = nullptr) {
    if (eq.lo) {  // tree levels: degree 3, tables T0 T1 or N0 N1 D0 D1
        if (rel == REL_PRODUCT && first && (base_mask & 3) == 3)
            launch_round_eq<REL_PRODUCT, false, true>(grid, s, t, base_mask, first, out_len, pairs,
                                                      r, lambda, partials, eq);
        else if (rel == REL_PRODUCT)
            launch_round_eq<REL_PRODUCT, false>(grid, s, t, base_mask, first, out_len, pairs, r,
                                                lambda, partials, eq);
        else if (!t.src[0])
            launch_round_eq<REL_FRACTION, true>(grid, s, t, base_mask, first, out_len, pairs, r,
                                                lambda, partials, eq);
        else
            launch_round_eq<REL_FRACTION, false>(grid, s, t, base_mask, first, out_len, pairs, r,
                                                 lambda, partials, eq, pn, pd);
        return;
    }
    switch (d) {
#define CASE(D)                                                                              \
    case D:                                                                                  \
        if (rel == REL_TERMS)                                                                \
            launch_round<D, REL_TERMS>(grid, s, t, F, base_mask, first, out_len, pairs, r,   \
                                       lambda, partials, terms);                             \
        else if (F == D) /* a product of D tables: the tables known at compile time */      \
            launch_round<D, REL_PRODUCT, D>(grid, s, t, F, base_mask, first, out_len, pairs, \
                                            r, lambda, partials, terms);                     \
        else                                                                                 \
            launch_round<D, REL_PRODUCT>(grid, s, t, F, base_mask, first, out_len, pairs, r, \
                                         lambda, partials, terms);                           \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
    }
}

static void launch_transcript_d(uint32_t d, hipStream_t s, uint32_t j, const ScRoundInst *ri,
                                ScInstState *ist, uint32_t count, uint32_t nblk_max,
                                const fp2 *partials, ScCtl *ctl, const DevKeys &keys,
                                const uint8_t *labels, const uint32_t *lens, fp2 *rounds, fp2 *chals) {
    size_t shm = sizeof(fp2) * count * (d + 1);
    switch (d) {
#define CASE(D)                                                                               \
    case D:                                                                                   \
        hipLaunchKernelGGL(k_sc_transcript<D>, dim3(1), dim3(SC_THREADS), shm, s, j, ri, ist,  \
                           count, nblk_max, partials, ctl, keys, labels, lens, rounds, chals); \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
    }
}

static void launch_tail_d(uint32_t d, hipStream_t s, uint32_t j0, uint32_t n_max,
                          const ScRoundInst *rtab, ScInstState *ist, uint32_t count,
                          uint32_t nblk_max0, const fp2 *partials, const ScTailStep *steps,
                          const ScTerms *terms, fp2 lambda, ScCtl *ctl, const DevKeys &keys,
                          const uint8_t *labels, const uint32_t *lens, fp2 *rounds, fp2 *chals,
                          const ScFinals &fin) {
    size_t shm = sizeof(ScTailStep) * (size_t)(n_max - j0) * count + sizeof(fp2) * count * (d + 1) +
                 (size_t)(n_max - j0) * (sizeof(ScRoundInst) * count + 2 * LABEL_SLOT + 8);
    switch (d) {
#define CASE(D)                                                                               \
    case D:                                                                                   \
        hipLaunchKernelGGL(k_sc_tail<D>, dim3(1), dim3(SC_TAIL_THREADS), shm, s, j0, n_max,    \
                           rtab, ist, count, nblk_max0, partials, steps, terms, lambda, ctl,   \
                           keys,                                                               \
                           labels, lens, rounds, chals, fin);                                  \
        break;
        CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8) CASE(9) CASE(10)
#undef CASE
    }
}

static uint32_t blocks_for_pairs(const Ctx *ctx, uint64_t items) {
    const uint64_t cap = ctx->round_blocks ? ctx->round_blocks : SC_MAX_BLOCKS;
    uint64_t b = (items + SC_THREADS - 1) / SC_THREADS;
    return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(b, cap));
}

DevKeys dev_keys(const Ctx *ctx) {
    DevKeys k;
    memcpy(k.absorb, ctx->keys[DOM_TRANSCRIPT_ABSORB], 32);
    memcpy(k.squeeze, ctx->keys[DOM_TRANSCRIPT_SQUEEZE], 32);
    memcpy(k.ratchet, ctx->keys[DOM_TRANSCRIPT_RATCHET], 32);
    return k;
}

void load_state_words(const uint8_t st[32], uint32_t w[8]) {
    for (int i = 0; i < 8; i++)
        w[i] = (uint32_t)st[4 * i] | ((uint32_t)st[4 * i + 1] << 8) |
               ((uint32_t)st[4 * i + 2] << 16) | ((uint32_t)st[4 * i + 3] << 24);
}

void store_state_words(const uint32_t w[8], uint8_t st[32]) {
    for (int i = 0; i < 8; i++)
        for (int j = 0; j < 4; j++) st[4 * i + j] = (uint8_t)(w[i] >> (8 * j));
}

Status run_batch(Ctx *ctx, Transcript &tr, const std::string &p, uint32_t D,
                 std::vector<GenInst> &insts, fp2 gamma, fp2 lambda, SumcheckOut &out) {
    hipStream_t s = ctx->stream;
    const uint32_t count = (uint32_t)insts.size();
    uint32_t n_max = 0;
    for (auto &in : insts) n_max = std::max(n_max, in.vars);
    out.gamma = gamma;
    if (count == 0) return Status::fail(TSR_BAD_SHAPE, "NoInstances");
    if (SC_TRANSCRIPT_STATIC + sizeof(fp2) * count * (D + 1) > SC_TRANSCRIPT_LDS)
        return Status::fail(TSR_DECLINED, "a batch of " + std::to_string(count) +
                                              " instances at degree " + std::to_string(D) +
                                              " exceeds the transcript kernel's LDS");

    std::vector<uint32_t> pad(count), wgiven(count);
    std::vector<fp2> weight(count);
    std::vector<ScInstState> hstate(count);
    fp2 w{1, 0};
    for (uint32_t i = 0; i < count; i++) {
        pad[i] = n_max - insts[i].vars;
        weight[i] = w;
        wgiven[i] = insts[i].claim_given ? 1 : 0;
        hstate[i].wsum = insts[i].claim_given ? gl::mul(w, insts[i].claim) : fp2{0, 0};
        hstate[i].eq_k = w;  // gamma^i: the eq factor's K starts at 1
        w = gl::mul(w, gamma);
    }
    Status st;

    // eq factors: the points on the device, and for each own round t of each such instance the
    // half tables of E_t = eq(point[t+1 ..], .) over the pairs of round t (lo: the low half of
    // the pair index, hi: the rest)
    std::vector<fp2> eq_pts;
    std::vector<EqSeg> segs;
    std::vector<std::vector<EqSplit>> splits(count);
    std::vector<uint32_t> pt_off(count, 0);
    uint64_t seg_len = 0;
    for (uint32_t i = 0; i < count; i++) {
        if (insts[i].rel == REL_FRACTION && !insts[i].eq_factor)
            return Status::fail(TSR_DECLINED, "a fraction relation carries its eq factor");
        if (!insts[i].eq_factor) continue;
        const GenInst &g = insts[i];
        const size_t nt = g.rel == REL_PRODUCT ? 2 : 4;
        // a fraction level always brings its claim; a product (prove_batch's Sum eq a b) may
        // leave S_i to its first sums
        bool shape = D == 3 && (g.rel == REL_PRODUCT || g.rel == REL_FRACTION) &&
                     (g.claim_given || g.rel == REL_PRODUCT) && g.eq_point.size() == g.vars &&
                     g.tabs.size() == nt;
        // tables present, except both numerators of a fraction with unit numerators
        for (size_t k = 0; shape && k < nt; k++)
            shape = g.tabs[k].src || (g.rel == REL_FRACTION && k < 2 && !g.tabs[0].src && !g.tabs[1].src);
        if (g.rel == REL_FRACTION && shape) shape = !g.tabs[0].src == !g.tabs[1].src;
        if (!shape)
            return Status::fail(TSR_DECLINED, "an eq factor needs degree 3, a claim, a point of vars "
                                              "coordinates and tables T0 T1 or N0 N1 D0 D1");
        pt_off[i] = (uint32_t)eq_pts.size();
        eq_pts.insert(eq_pts.end(), g.eq_point.begin(), g.eq_point.end());
        for (uint32_t t = 0; t < g.vars; t++) {
            // round t weights its 2^m pairs by eq_lo[q mod 2^a] eq_hi[q >> a]; a = m (eq_hi = [1])
            // up to SC_EQ_LO_BITS variables, so that small rounds and the tail use eq_lo alone
            const uint32_t m = g.vars - t - 1;
            const uint32_t a = m <= SC_EQ_LO_BITS ? m : std::max<uint32_t>(SC_EQ_LO_BITS, m / 2);
            segs.push_back(EqSeg{seg_len, pt_off[i] + t + 1, a});
            seg_len += 1ull << a;
            segs.push_back(EqSeg{seg_len, pt_off[i] + t + 1 + a, m - a});
            seg_len += 1ull << (m - a);
        }
    }
    // working buffers: per table a (2^{n-1}) and b (2^{n-2}) Fp2
    std::vector<std::vector<DevBuf<fp2>>> bufa(count), bufb(count);
    // instance i's partials: SC_MAX_BLOCKS x (D+1) from partials + i SC_MAX_BLOCKS (D+1)
    DevBuf<fp2> partials;
    const uint64_t pstride = (uint64_t)SC_MAX_BLOCKS * (D + 1);
    if ((st = partials.alloc(ctx, pstride * std::max(count, 1u)))) return st;
    for (uint32_t i = 0; i < count; i++) {
        const uint32_t v = insts[i].vars;
        const uint64_t len = 1ull << v;
        bufa[i] = std::vector<DevBuf<fp2>>(insts[i].tabs.size());
        bufb[i] = std::vector<DevBuf<fp2>>(insts[i].tabs.size());
        for (size_t k = 0; k < insts[i].tabs.size(); k++) {
            if (!insts[i].tabs[k].src || v == 0) continue;
            if ((st = bufa[i][k].alloc(ctx, len / 2))) return st;
            if (v >= 2 && (st = bufb[i][k].alloc(ctx, len / 4))) return st;
        }
    }

    // the terms of REL_TERMS instances, as the round kernel reads them
    std::vector<ScTerms> sterms(count);
    for (uint32_t i = 0; i < count; i++) {
        ScTerms &st_ = sterms[i];
        memset(&st_, 0, sizeof st_);
        if (insts[i].rel != REL_TERMS) continue;
        if (insts[i].terms.size() > SC_MAXT || insts[i].tabs.size() > SC_MAXF)
            return Status::fail(TSR_DECLINED, "more than 16 terms or 10 tables in an instance");
        st_.nterms = (uint32_t)insts[i].terms.size();
        for (uint32_t t = 0; t < st_.nterms; t++) {
            const GenTerm &g = insts[i].terms[t];
            st_.weight[t] = g.weight;
            if (g.weight.c0 == 1 && g.weight.c1 == 0) st_.unit |= 1u << t;
            for (size_t k = 0; k < g.mult.size() && k < SC_MAXF; k++) st_.mult[t][k] = g.mult[k];
        }
    }

    // S_i of instances on no variable (no claim given): the relation at the single entries
    for (uint32_t i = 0; i < count; i++) {
        if (insts[i].vars != 0 || pad[i] == 0 || insts[i].claim_given) continue;
        std::vector<fp2> single;
        for (auto &tab : insts[i].tabs) {
            fp2 v{1, 0};
            if (tab.src) {
                v = fp2{0, 0};
                TSR_HIP(hipMemcpyAsync(&v, (const uint8_t *)tab.src + (tab.base ? 8 : 16) * tab.off,
                                       tab.base ? 8 : 16, hipMemcpyDeviceToHost, s));
                TSR_HIP(tsr::stream_sync(s));
            }
            single.push_back(v);
        }
        fp2 rel{1, 0};
        if (insts[i].rel == REL_TERMS) {
            rel = fp2{0, 0};
            for (const GenTerm &g : insts[i].terms) {
                fp2 p = g.weight;
                for (size_t k = 0; k < g.mult.size(); k++)
                    for (uint32_t m = 0; m < g.mult[k]; m++) p = gl::mul(p, single[k]);
                rel = gl::add(rel, p);
            }
        } else {
            for (auto &v : single) rel = gl::mul(rel, v);
        }
        hstate[i].wsum = gl::mul(weight[i], rel);
        wgiven[i] = 1;
    }

    // labels of every round, the control block, per-round block counts
    const uint32_t nr = std::max(n_max, 1u);
    std::vector<uint8_t> labels((size_t)2 * nr * LABEL_SLOT, 0);
    std::vector<uint32_t> lens(2 * nr, 0);
    for (uint32_t j = 0; j < n_max; j++) {
        std::string a = p + "/round/" + std::to_string(j), b = p + "/r/" + std::to_string(j);
        if (a.size() > LABEL_SLOT || b.size() > LABEL_SLOT)
            return Status::fail(TSR_DECLINED, "label longer than 64 bytes");
        memcpy(&labels[(2 * j) * LABEL_SLOT], a.data(), a.size());
        memcpy(&labels[(2 * j + 1) * LABEL_SLOT], b.data(), b.size());
        lens[2 * j] = (uint32_t)a.size();
        lens[2 * j + 1] = (uint32_t)b.size();
    }
    // the round table: per round j and instance i the blocks of partials round j's message
    // reads (round 0 also reads a padded instance's first sums for S_i) and the eq coordinate
    std::vector<ScRoundInst> rtab((size_t)nr * count);
    std::vector<uint32_t> nblk_max(nr, 0);
    for (uint32_t j = 0; j < nr; j++) {
        for (uint32_t i = 0; i < count; i++) {
            ScRoundInst &e = rtab[(size_t)j * count + i];
            const uint32_t v = insts[i].vars;
            e.nblk = 0;
            e.pad = pad[i];
            e.eqf = insts[i].eq_factor ? 1 : 0;
            e.wsum_given = wgiven[i];
            e.weight = weight[i];
            e.rt = (e.eqf && j >= pad[i] && j < n_max) ? insts[i].eq_point[j - pad[i]] : fp2{0, 0};
            // a padded eq instance without a claim: S_i in round 0 takes the line of its first
            // own round (coordinate 0 of the point)
            if (e.eqf && j == 0 && pad[i] > 0 && !wgiven[i]) e.rt = insts[i].eq_point[0];
            if (v == 0 || j >= n_max || (j < pad[i] && j != 0)) continue;
            const uint32_t t = j < pad[i] ? 0 : j - pad[i];
            e.nblk = blocks_for_pairs(ctx, 1ull << (v - t - 1));
            nblk_max[j] = std::max(nblk_max[j], e.nblk);
        }
    }
    ScCtl hctl;
    load_state_words(tr.state, hctl.t.state);
    hctl.t.counter = tr.counter;
    hctl.r = fp2{0, 0};

    // final evaluations: each non-public table's single remaining entry (1 for a constant table)
    std::vector<const void *> fptr;
    std::vector<uint32_t> fbase, foff;
    for (uint32_t i = 0; i < count; i++) {
        for (size_t k = 0; k < insts[i].tabs.size(); k++) {
            const GenTable &tab = insts[i].tabs[k];
            if (tab.is_public) continue;
            if (!tab.src) {
                fptr.push_back(nullptr);
                fbase.push_back(0);
                foff.push_back(0);
            } else if (insts[i].vars == 0) {
                fptr.push_back(tab.src);
                fbase.push_back(tab.base ? 1 : 0);
                foff.push_back(tab.off);
            } else {
                // the last own round t = vars-1 wrote into a (t even) or b (t odd)
                fptr.push_back(((insts[i].vars - 1) % 2 == 0) ? (const void *)bufa[i][k].ptr
                                                               : (const void *)bufb[i][k].ptr);
                fbase.push_back(0);
                foff.push_back(0);
            }
        }
    }
    const size_t nf = fptr.size();

    // The tail: from round j0 on, every round has at most SC_TAIL_PAIRS pairs per instance
    // (round j has 2^(n_max - j - 1) in every instance it binds), and one workgroup runs it all.
    const bool tail = count <= SC_TAIL_MAX && n_max >= 1;
    const uint32_t j0 = !tail ? n_max : (n_max >= SC_TAIL_LOG + 2 ? n_max - SC_TAIL_LOG - 2 : 0);
    const size_t n_steps = tail ? (size_t)(n_max - j0) * count : 0;

    // Everything the rounds read from the host goes up in one copy and everything the host reads
    // back comes down in one: one device block laid out as
    //   [eq points | eq segments | labels | label lengths | nblk | instances | finals' sources]
    //   [control block | rounds | challenges | finals]
    // (the control block is in both copies: its initial state up, its final state down).
    size_t at = 0;
    auto place = [&](size_t bytes) {
        at = (at + 255) & ~(size_t)255;
        const size_t o = at;
        at += bytes;
        return o;
    };
    const size_t o_pts = place(16 * eq_pts.size()), o_segs = place(sizeof(EqSeg) * segs.size());
    const size_t o_labels = place(labels.size()), o_lens = place(4 * lens.size());
    const size_t o_rtab = place(sizeof(ScRoundInst) * rtab.size()), o_state = place(sizeof(ScInstState) * count);
    const size_t o_fptr = place(8 * nf), o_fbase = place(4 * nf), o_foff = place(4 * nf);
    const size_t o_steps = place(sizeof(ScTailStep) * n_steps);
    const size_t o_terms = place(tail ? sizeof(ScTerms) * count : 0);
    const size_t o_ctl = place(sizeof(ScCtl));
    const size_t up_end = at;
    const size_t o_rounds = place(16 * (size_t)nr * D), o_chals = place(16 * (size_t)nr);
    const size_t o_fin = place(16 * std::max<size_t>(nf, 1));
    DevBuf<uint8_t> blob;
    if ((st = blob.alloc(ctx, at))) return st;
    uint8_t *const dev = blob.ptr;
    fp2 *const d_eqpt = (fp2 *)(dev + o_pts);
    EqSeg *const d_segs = (EqSeg *)(dev + o_segs);
    const ScRoundInst *const d_rtab = (const ScRoundInst *)(dev + o_rtab);
    ScInstState *const d_state = (ScInstState *)(dev + o_state);
    ScCtl *const d_ctl = (ScCtl *)(dev + o_ctl);

    // the eq half tables of every own round (written by k_eq_segments after the upload)
    DevBuf<fp2> d_eqtab;
    if (!segs.empty()) {
        if ((st = d_eqtab.alloc(ctx, seg_len))) return st;
        size_t g = 0;
        for (uint32_t i = 0; i < count; i++) {
            if (!insts[i].eq_factor) continue;
            for (uint32_t t = 0; t < insts[i].vars; t++, g += 2)
                splits[i].push_back(EqSplit{d_eqtab.ptr + segs[g].off, d_eqtab.ptr + segs[g + 1].off, segs[g].n});
        }
    }
    // own round t binds from (t == 0: original, odd t: a, even t: b) into (even t: a, odd: b)
    auto tables_of = [&](uint32_t i, uint32_t t, ScTables &tt, uint32_t &base_mask) {
        base_mask = 0;
        for (uint32_t k = 0; k < insts[i].tabs.size(); k++) {
            const GenTable &tab = insts[i].tabs[k];
            tt.shift[k] = tab.shift;
            tt.off[k] = tab.off;
            if (!tab.src) {
                tt.src[k] = nullptr;
                tt.dst[k] = nullptr;
                continue;
            }
            if (t == 0) {
                tt.src[k] = tab.src;
                if (tab.base) base_mask |= 1u << k;
            } else {
                tt.src[k] = (t % 2 == 1) ? (const void *)bufa[i][k].ptr : (const void *)bufb[i][k].ptr;
            }
            tt.dst[k] = (t % 2 == 0) ? bufa[i][k].ptr : bufb[i][k].ptr;
        }
    };

    // the tail's steps: for in-block round j (j0 < j < n_max) instance i binds its own round
    // t - 1 at r_{j-1} and sums own round t = j - pad_i (sums only when t == 0), as the grid
    // launch after round j - 1 would; the last step binds own round vars_i - 1
    std::vector<ScTailStep> steps(n_steps);
    auto kind_of = [&](uint32_t i, bool first, uint32_t bm, const ScTables &tt) -> uint32_t {
        const GenInst &g = insts[i];
        if (g.eq_factor) {
            if (g.rel == REL_PRODUCT) return (first && (bm & 3) == 3) ? TAIL_EQ_PRODUCT_BASE : TAIL_EQ_PRODUCT;
            return tt.src[0] ? TAIL_EQ_FRACTION : TAIL_EQ_UNIT;
        }
        return g.rel == REL_TERMS ? TAIL_TERMS : TAIL_PRODUCT;
    };
    for (uint32_t j = j0 + 1; tail && j <= n_max; j++) {
        for (uint32_t i = 0; i < count; i++) {
            ScTailStep &sp = steps[(size_t)(j - j0 - 1) * count + i];
            memset(&sp, 0, sizeof sp);
            const uint32_t v = insts[i].vars;
            if (v == 0 || j < pad[i]) continue;  // TAIL_NONE
            const uint32_t t = j - pad[i];       // own round summed (t == v: the last bind only)
            uint32_t bm = 0;
            if (t == 0) {
                tables_of(i, 0, sp.t, bm);
                sp.first = 1;
                sp.out_len = 0;
                sp.pairs = 1ull << (v - 1);
            } else {
                tables_of(i, t - 1, sp.t, bm);
                sp.first = t - 1 == 0;
                sp.out_len = 1ull << (v - t);
                sp.pairs = t < v ? sp.out_len / 2 : 0;
            }
            sp.F = (uint32_t)insts[i].tabs.size();
            sp.base_mask = bm;
            sp.kind = kind_of(i, sp.first != 0, bm, sp.t);
            if (!splits[i].empty()) {
                const EqSplit &e = splits[i][std::min<uint32_t>(t, v - 1)];
                sp.eq_lo = e.lo;
                sp.eq_hi = e.hi;
                sp.eq_a = e.a;
                // the tail weights a pair by eq_lo alone: its rounds have at most 2^9 pairs
                if (sp.pairs > (1ull << e.a))
                    return Status::fail(TSR_BAD_SHAPE, "a tail round with more than one eq_hi entry");
            }
        }
    }
    // one pinned host block for both copies: the upload image, then the readback
    uint8_t *up;
    if ((st = pinned_host(ctx, std::max(up_end, at - o_ctl), &up))) return st;
    memset(up, 0, up_end);
    auto put = [&](size_t o, const void *src, size_t bytes) {
        if (bytes) memcpy(up + o, src, bytes);
    };
    put(o_pts, eq_pts.data(), 16 * eq_pts.size());
    put(o_segs, segs.data(), sizeof(EqSeg) * segs.size());
    put(o_labels, labels.data(), labels.size());
    put(o_lens, lens.data(), 4 * lens.size());
    put(o_rtab, rtab.data(), sizeof(ScRoundInst) * rtab.size());
    put(o_state, hstate.data(), sizeof(ScInstState) * count);
    put(o_fptr, fptr.data(), 8 * nf);
    put(o_fbase, fbase.data(), 4 * nf);
    put(o_foff, foff.data(), 4 * nf);
    put(o_steps, steps.data(), sizeof(ScTailStep) * n_steps);
    if (tail) put(o_terms, sterms.data(), sizeof(ScTerms) * count);
    put(o_ctl, &hctl, sizeof(ScCtl));
    TSR_HIP(hipMemcpyAsync(dev, up, up_end, hipMemcpyHostToDevice, s));
    const DevKeys keys = dev_keys(ctx);

    if (!segs.empty()) {
        uint32_t widest = 0;
        for (auto &g : segs) widest = std::max(widest, g.n);
        const unsigned gx = (unsigned)std::min<uint64_t>(((1ull << widest) + 255) / 256, 1024);
        for (size_t g0 = 0; g0 < segs.size(); g0 += 65535) {
            const unsigned ny = (unsigned)std::min<size_t>(65535, segs.size() - g0);
            hipLaunchKernelGGL(k_eq_segments, dim3(gx, ny), dim3(256), 0, s, d_eqpt, d_segs + g0,
                               d_eqtab.ptr);
        }
        if ((st = check_launch("eq segments"))) return st;
    }


    {
        KernelTimer timer(ctx, "k_sc_round", s);
        for (uint32_t i = 0; i < count; i++) {
            if (insts[i].vars == 0) continue;
            ScTables t{};
            uint32_t bm;
            tables_of(i, 0, t, bm);
            uint64_t pairs = 1ull << (insts[i].vars - 1);
            launch_round_d(D, insts[i].rel, dim3(blocks_for_pairs(ctx, pairs)), s, t,
                           (uint32_t)insts[i].tabs.size(), bm, true, 0, pairs, nullptr, lambda,
                           partials.ptr + pstride * i, sterms[i],
                           splits[i].empty() ? EqSplit() : splits[i][0], insts[i].parent_n,
                           insts[i].parent_d);
        }
    }
    for (uint32_t j = 0; j < j0; j++) {
        {
            KernelTimer timer(ctx, "k_sc_transcript", s);
            launch_transcript_d(D, s, j, d_rtab + (size_t)j * count, d_state, count, nblk_max[j],
                                partials.ptr, d_ctl, keys, dev + o_labels,
                                (const uint32_t *)(dev + o_lens), (fp2 *)(dev + o_rounds),
                                (fp2 *)(dev + o_chals));
        }
        KernelTimer timer(ctx, "k_sc_round", s);
        for (uint32_t i = 0; i < count; i++) {
            if (j < pad[i]) continue;
            uint32_t t = j - pad[i];
            ScTables tt{};
            uint32_t bm;
            tables_of(i, t, tt, bm);
            uint64_t out_len = 1ull << (insts[i].vars - t - 1);
            uint64_t pairs = out_len / 2;  // next own round's pairs (0 after the last)
            // the pairs summed are those of own round t + 1 (none after the last round)
            const EqSplit eq = splits[i].empty() ? EqSplit() : splits[i][std::min<uint32_t>(t + 1, insts[i].vars - 1)];
            launch_round_d(D, insts[i].rel, dim3(blocks_for_pairs(ctx, pairs ? pairs : out_len)), s,
                           tt, (uint32_t)insts[i].tabs.size(), bm, t == 0, out_len, pairs,
                           &d_ctl->r, lambda, partials.ptr + pstride * i, sterms[i], eq);
        }
    }
    const ScFinals fin{(const void *const *)(dev + o_fptr), (const uint32_t *)(dev + o_fbase),
                       (const uint32_t *)(dev + o_foff), (fp2 *)(dev + o_fin), (uint32_t)nf};
    if (tail) {
        KernelTimer timer(ctx, "k_sc_tail", s);
        launch_tail_d(D, s, j0, n_max, d_rtab, d_state, count, nblk_max[j0], partials.ptr,
                      (const ScTailStep *)(dev + o_steps), (const ScTerms *)(dev + o_terms), lambda,
                      d_ctl, keys, dev + o_labels, (const uint32_t *)(dev + o_lens),
                      (fp2 *)(dev + o_rounds), (fp2 *)(dev + o_chals), fin);
    }
    if ((st = check_launch("sumcheck rounds"))) return st;
    if (nf && !tail) {
        hipLaunchKernelGGL(k_sc_gather, dim3((unsigned)((nf + 63) / 64)), dim3(64), 0, s,
                           (const void *const *)(dev + o_fptr), (const uint32_t *)(dev + o_fbase),
                           (const uint32_t *)(dev + o_foff), (uint32_t)nf, (fp2 *)(dev + o_fin));
        if ((st = check_launch("sumcheck finals"))) return st;
    }
    // the readback reuses the pinned block: the stream has finished the upload before it
    uint8_t *const down = up;
    TSR_HIP(hipMemcpyAsync(down, dev + o_ctl, at - o_ctl, hipMemcpyDeviceToHost, s));
    TSR_HIP(tsr::stream_sync(s));
    memcpy(&hctl, &down[0], sizeof(ScCtl));
    out.rounds.resize((size_t)n_max * D);
    out.challenges.resize(n_max);
    out.finals.resize(nf);
    if (n_max) {
        memcpy(out.rounds.data(), &down[o_rounds - o_ctl], 16 * (size_t)n_max * D);
        memcpy(out.challenges.data(), &down[o_chals - o_ctl], 16 * (size_t)n_max);
    }
    if (nf) memcpy(out.finals.data(), &down[o_fin - o_ctl], 16 * nf);
    store_state_words(hctl.t.state, tr.state);
    tr.counter = hctl.t.counter;
    tr.absorb_fp2s((p + "/evals").c_str(), (const uint64_t *)out.finals.data(), out.finals.size());
    return Status::ok();
}

// The point of a TSR_TABLE_EQ table (host or device memory) on the host.
static Status eq_point_of(Ctx *ctx, const tsr_table &tt, uint32_t vars, std::vector<fp2> &pt) {
    pt.assign(vars, fp2{0, 0});
    if (vars == 0) return Status::ok();
    if (!is_device_ptr(tt.data)) {
        memcpy(pt.data(), tt.data, 16ull * vars);  // host memory: no trip through the stream
    } else {
        TSR_HIP(hipMemcpyAsync(pt.data(), tt.data, 16ull * vars, hipMemcpyDeviceToHost, ctx->stream));
        TSR_HIP(tsr::stream_sync(ctx->stream));
    }
    for (auto &c : pt)
        if (c.c0 >= gl::P || c.c1 >= gl::P)
            return Status::fail(TSR_BAD_ARG, "an eq point coordinate is not canonical");
    return Status::ok();
}

Status prove_batch(Ctx *ctx, Transcript &tr, const char *prefix, uint32_t degree,
                   const tsr_instance *instances, uint32_t count, SumcheckOut &out,
                   std::vector<DevBuf<uint64_t>> &uploads) {
    hipStream_t s = ctx->stream;
    std::string p(prefix);
    std::vector<GenInst> insts(count);
    uploads.reserve(64);
    for (uint32_t i = 0; i < count; i++) {
        const tsr_instance &ti = instances[i];
        insts[i].vars = ti.vars;
        insts[i].rel = ti.terms ? REL_TERMS : REL_PRODUCT;
        for (uint32_t t = 0; ti.terms && t < ti.nterms; t++) {
            const tsr_term &tm = ti.terms[t];
            GenTerm g;
            g.weight = fp2{tm.weight[0], tm.weight[1]};
            g.mult.assign(ti.factors, 0);
            for (uint32_t f = 0; f < tm.nfactors; f++) g.mult[tm.factors[f]]++;
            insts[i].terms.push_back(g);
        }
        // Sum eq(r, x) a b with eq public: eq stays a point (k_sc_round_eq's product relation)
        uint32_t n_eq = 0, eq_at_k = 0;
        for (uint32_t k = 0; k < ti.factors; k++)
            if (ti.tables[k].ext == TSR_TABLE_EQ) n_eq++, eq_at_k = k;
        const bool eq_path = !ti.terms && ti.factors == 3 && n_eq == 1 && degree == 3 &&
                             ti.vars >= 1 && ti.tables[eq_at_k].is_public;
        if (eq_path) {
            insts[i].eq_factor = true;
            Status st = eq_point_of(ctx, ti.tables[eq_at_k], ti.vars, insts[i].eq_point);
            if (st) return st;
        }
        const uint64_t len = 1ull << ti.vars;
        for (uint32_t k = 0; k < ti.factors; k++) {
            const tsr_table &tt = ti.tables[k];
            if (eq_path && k == eq_at_k) continue;
            GenTable tab;
            tab.is_public = tt.is_public != 0;
            if (tt.ext == TSR_TABLE_EQ) {
                // any other use of an eq table: built on the device from its point
                std::vector<fp2> pt;
                Status st = eq_point_of(ctx, tt, ti.vars, pt);
                if (st) return st;
                DevBuf<uint64_t> dpt;
                if ((st = dpt.alloc(ctx, 2 * std::max<uint64_t>(ti.vars, 1)))) return st;
                if (ti.vars)
                    TSR_HIP(hipMemcpyAsync(dpt.ptr, pt.data(), 16ull * ti.vars, hipMemcpyHostToDevice, s));
                uploads.emplace_back();
                if ((st = uploads.back().alloc(ctx, 2 * len))) return st;
                if ((st = eq_table(ctx, (const fp2 *)dpt.ptr, ti.vars, (fp2 *)uploads.back().ptr, s))) return st;
                tab.base = false;
                tab.src = uploads.back().ptr;
                insts[i].tabs.push_back(tab);
                continue;
            }
            tab.base = !tt.ext;
            const uint64_t words = len * (tt.ext ? 2 : 1);
            if (is_device_ptr(tt.data)) {
                tab.src = tt.data;
            } else {
                uploads.emplace_back();
                Status st = uploads.back().alloc(ctx, words);
                if (st) return st;
                TSR_HIP(hipMemcpyAsync(uploads.back().ptr, tt.data, 8 * words,
                                       hipMemcpyHostToDevice, s));
                tab.src = uploads.back().ptr;
            }
            insts[i].tabs.push_back(tab);
        }
    }
    // gamma on the host transcript (prover.rs:1147-1150), once the tables are in place
    fp2 gamma = tr.challenge_fp2((p + "/gamma").c_str(), nullptr);
    return run_batch(ctx, tr, p, degree, insts, gamma, fp2{0, 0}, out);
}

}  // namespace tsr

using namespace tsr;

extern "C" tsr_status tsr_sumcheck_prove(tsr_ctx *ctx_, tsr_transcript *t, const char *prefix,
                                         uint32_t degree, const tsr_instance *instances,
                                         uint32_t count, uint64_t *gamma_out,
                                         uint64_t *rounds_out, uint64_t *challenges_out,
                                         uint64_t *finals_out, uint64_t finals_cap,
                                         uint64_t *n_finals) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !t || !prefix) return Status::fail(TSR_BAD_ARG, "null argument").code;
    // SumcheckError order (prover.rs:1196-1230): degree, no instances, then per instance.
    if (degree < 1 || degree > SC_MAXD)
        return Status::fail(TSR_BAD_SHAPE, "Degree: " + std::to_string(degree) +
                                               " is outside 1..=10").code;
    if (count == 0 || !instances) return Status::fail(TSR_BAD_SHAPE, "NoInstances").code;
    if (!label_ok(prefix)) return Status::fail(TSR_BAD_ARG, "a label is printable ASCII").code;
    uint64_t n_final = 0;
    for (uint32_t i = 0; i < count; i++) {
        const tsr_instance &in = instances[i];
        if (in.factors == 0 || !in.tables)
            return Status::fail(TSR_BAD_SHAPE, "NoTables: instance " + std::to_string(i)).code;
        if (!in.terms && in.factors > degree)
            return Status::fail(TSR_BAD_SHAPE, "TooManyTables: instance " + std::to_string(i) +
                                                   " has " + std::to_string(in.factors) +
                                                   " tables at degree " + std::to_string(degree)).code;
        if (in.terms && in.nterms == 0)
            return Status::fail(TSR_BAD_SHAPE, "NoTerms: instance " + std::to_string(i)).code;
        for (uint32_t t = 0; in.terms && t < in.nterms; t++) {
            const tsr_term &tm = in.terms[t];
            if (tm.nfactors == 0 || tm.nfactors > degree)
                return Status::fail(TSR_BAD_SHAPE, "TermFactors: instance " + std::to_string(i) +
                                                       " term " + std::to_string(t) + " has " +
                                                       std::to_string(tm.nfactors) + " factors at degree " +
                                                       std::to_string(degree)).code;
            if (!tm.factors) return Status::fail(TSR_BAD_ARG, "null term factors").code;
            for (uint32_t f = 0; f < tm.nfactors; f++)
                if (tm.factors[f] >= in.factors)
                    return Status::fail(TSR_BAD_SHAPE, "TermFactor: instance " + std::to_string(i) +
                                                           " term " + std::to_string(t) + " names table " +
                                                           std::to_string(tm.factors[f]) + " of " +
                                                           std::to_string(in.factors)).code;
        }
        if (in.vars > 40)
            return Status::fail(TSR_DECLINED, "tables above 2^40 entries").code;
        if (in.factors > SC_MAXF || (in.terms && in.nterms > SC_MAXT))
            return Status::fail(TSR_DECLINED, "more than 10 tables or 16 terms in an instance").code;
        for (uint32_t k = 0; k < in.factors; k++) {
            if (!in.tables[k].data) return Status::fail(TSR_BAD_ARG, "null table").code;
            if (in.tables[k].ext > TSR_TABLE_EQ)
                return Status::fail(TSR_BAD_ARG, "a table is base (0), ext (1) or an eq point (2)").code;
            if (!in.tables[k].is_public) n_final++;
        }
    }
    if (count > 256) return Status::fail(TSR_DECLINED, "more than 256 instances in one batch").code;
    if (n_finals) *n_finals = n_final;
    if (n_final > finals_cap || !finals_out)
        return Status::fail(TSR_BAD_SHAPE, "final evaluation buffer too small").code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    SumcheckOut out;
    std::vector<DevBuf<uint64_t>> uploads;
    // the proof runs on a copy of the transcript, which replaces the caller's only when every
    // step succeeded: a refusal (no memory, a decline) leaves the transcript as it was
    Transcript work = t->t;
    Status st = prove_batch(ctx, work, prefix, degree, instances, count, out, uploads);
    if (st) return st.code;
    t->t = work;
    if (gamma_out) {
        gamma_out[0] = out.gamma.c0;
        gamma_out[1] = out.gamma.c1;
    }
    if (rounds_out) memcpy(rounds_out, out.rounds.data(), 16 * out.rounds.size());
    if (challenges_out) memcpy(challenges_out, out.challenges.data(), 16 * out.challenges.size());
    memcpy(finals_out, out.finals.data(), 16 * out.finals.size());
    return TSR_OK;
}
