// Reed-Solomon row encoding on gfx950: the NTT kernels.
//
// What is computed is fixed by the reference's RowEncoder (crates/tessera-pcs/src/encoding.rs:
// 292-320): codeword symbol j = g(7 * w_n^j), g = f + (X^k - 1) m. The kernels follow the same
// algebra — Cooley-Tukey with the inverse twiddles of size k (natural -> bit-reversed, giving
// k f_{rev(t)}), a spread by 7^{rev_k(t)}/k to stride 2^r with the mask terms, then
// Gentleman-Sande with the forward twiddles of size n (bit-reversed -> natural) — but organised
// for the GPU:
//
//   * k_tile_pass: a run of up to 9 radix-2 stages over index bits [lo, lo+nb) of every row,
//     on LDS tiles of 2^(nb+logc) elements: 2^nb strided points x 2^logc consecutive columns,
//     2^12 elements (32 KiB) in all, so each wave reads and writes runs of 2^logc words (64 B
//     for the 9-bit passes). 32 KiB tiles of eight columns measured faster than 64 KiB tiles of
//     sixteen (half the workgroups per CU) and than 16 KiB tiles of four. Ten bits run as one
//     pass on 64 KiB tiles of eight columns (1024 threads) rather than two passes of five.
//   * k_encode_mid: the fused middle — the low CT stages of the size-k inverse transform on a
//     contiguous tile of 2^tk values, the spread and mask, and the low GS stages of the size-n
//     forward transform on the 2^(tk+r) symbols that tile expands to. One HBM read of the row
//     tile, one write of the codeword tile.
//
// Twiddles come from tables Z[t] = w^{rev(t)} (t < size/2; ntt.rs:110-125), read through the
// L2/MALL; the group of a butterfly at stage bit b is g = p >> (b + 1) for the low index p.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <algorithm>
#include <memory>
#include <string>
#include <vector>
#include "field.h"
#include "ntt.h"

namespace tsr {

constexpr int TILE_THREADS = 512;
// A strided pass's tile: 2^PASS_TILE_LOG elements (32 KiB), 2^nb strided points x 2^(12 - nb)
// consecutive columns, so a pass of few stage bits still has 512 radix-8 groups per round.
constexpr uint32_t PASS_TILE_LOG = 12;
// Threads of the fused middle's workgroup. Its 64 KiB tile allows two workgroups per CU, so
// 1024 threads give the CU its full 32 waves (16 with 512 threads): 8% faster at the
// 256 x 2^20 rate-1/4 headline, measured side by side in one process.
#ifndef TSR_MID_THREADS
#define TSR_MID_THREADS 1024
#endif
constexpr int MID_THREADS = TSR_MID_THREADS;
constexpr int TILE_MAX_LOG = 13;  // 8192 u64 = 64 KiB of LDS

// ---- stage rounds on an LDS tile ------------------------------------------------------------
// A tile holds 2^(nb+logc) elements, element e = (mid << logc) | col with mid < 2^nb the stage
// index and col < 2^logc a column. The stage at mid bit sb pairs mid and mid + 2^sb (bit sb of
// mid clear) with the twiddle tw[g], g = (hi << (nb - sb - 1)) | (mid >> (sb + 1)): the group
// p >> (b + 1) of the global low index p of the pair, b the stage's global bit, hi the tile's
// index above its nb bits. A round runs R consecutive stage bits s0..s0+R-1 in registers: each
// thread loads the 2^R elements mid = base | (j << s0) (bits s0..s0+R-1 of base clear), applies
// the R stages (GS: ascending bits, CT: descending), and stores them; every butterfly is the
// radix-2 butterfly of the stage it belongs to, so rounds change the order of work only.
template <bool GS, int R>
__device__ __forceinline__ void stage_round(uint64_t *lds, uint32_t nb, uint32_t logc, uint32_t s0,
                                            uint64_t hi, const uint64_t *__restrict__ tw,
                                            uint32_t tid, uint32_t nthreads) {
    constexpr int M = 1 << R;
    const uint32_t groups = 1u << (nb - R + logc);
    const uint32_t cmask = (1u << logc) - 1;
    for (uint32_t q = tid; q < groups; q += nthreads) {
        const uint32_t col = q & cmask, m = q >> logc;
        const uint32_t base = ((m >> s0) << (s0 + R)) | (m & ((1u << s0) - 1));
        uint64_t x[M];
#pragma unroll
        for (int j = 0; j < M; j++) x[j] = lds[((base | ((uint32_t)j << s0)) << logc) | col];
#pragma unroll
        for (int tt = 0; tt < R; tt++) {
            const int t = GS ? tt : R - 1 - tt;
            const uint32_t sb = s0 + t;
            const uint64_t gbase = (hi << (nb - sb - 1)) | (base >> (sb + 1));
            uint64_t w[M / 2 > 0 ? M / 2 : 1];
#pragma unroll
            for (int u = 0; u < (M >> (t + 1)); u++) w[u] = tw[gbase | (uint64_t)u];
#pragma unroll
            for (int j = 0; j < M; j++) {
                if (j & (1 << t)) continue;
                const uint64_t wj = w[j >> (t + 1)];
                const uint64_t u = x[j], v = x[j + (1 << t)];
                if (GS) {
                    x[j] = gl::add(u, v);
                    x[j + (1 << t)] = gl::mul(gl::sub(u, v), wj);
                } else {
                    const uint64_t tv = gl::mul(v, wj);
                    x[j] = gl::add(u, tv);
                    x[j + (1 << t)] = gl::sub(u, tv);
                }
            }
        }
#pragma unroll
        for (int j = 0; j < M; j++) lds[((base | ((uint32_t)j << s0)) << logc) | col] = x[j];
    }
}

// All nb stages of a tile in rounds of up to 3 bits (GS from bit 0 up, CT from the top down),
// with a barrier after each round. The tile must be in LDS, behind a barrier, on entry.
template <bool GS>
__device__ __forceinline__ void tile_stages(uint64_t *lds, uint32_t nb, uint32_t logc, uint64_t hi,
                                            const uint64_t *__restrict__ tw) {
    const uint32_t tid = threadIdx.x, nthreads = blockDim.x;
    uint32_t done = 0;
    while (done < nb) {
        const uint32_t R = nb - done >= 3 ? 3 : nb - done;
        const uint32_t s0 = GS ? done : nb - done - R;
        if (R == 3)
            stage_round<GS, 3>(lds, nb, logc, s0, hi, tw, tid, nthreads);
        else if (R == 2)
            stage_round<GS, 2>(lds, nb, logc, s0, hi, tw, tid, nthreads);
        else
            stage_round<GS, 1>(lds, nb, logc, s0, hi, tw, tid, nthreads);
        __syncthreads();
        done += R;
    }
}

// The same rounds with the tile geometry known at compile time (NB stage bits, LOGC column
// bits, the round's first bit S0): each thread's 2^R elements and the stage's twiddles then sit
// at constant offsets from one address, which the LDS and global loads take as immediates.
// The R stages of one group held in registers (x[j] = element base | (j << S0)).
// TOP: the tile covers the row's top index bits (hi = 0), so its top stage (sb = NB - 1) has the
// group g = 0 and the twiddle Z[0] = 1: its butterflies skip the multiply.
template <bool GS, int R, int NB, int S0, bool TOP = false>
__device__ __forceinline__ void round_regs(uint64_t *x, uint32_t base, uint64_t hi,
                                           const uint64_t *__restrict__ tw) {
    constexpr int M = 1 << R;
#pragma unroll
    for (int tt = 0; tt < R; tt++) {
        const int t = GS ? tt : R - 1 - tt;
        const int sb = S0 + t;
        const bool unit = TOP && sb == NB - 1;
        const uint64_t *wp = tw + ((hi << (NB - sb - 1)) | (base >> (sb + 1)));
        uint64_t w[M / 2 > 0 ? M / 2 : 1];
#pragma unroll
        for (int u = 0; u < (M >> (t + 1)); u++) w[u] = unit ? 1 : wp[u];
#pragma unroll
        for (int j = 0; j < M; j++) {
            if (j & (1 << t)) continue;
            const uint64_t wj = w[j >> (t + 1)];
            const uint64_t u = x[j], v = x[j + (1 << t)];
            if (GS) {
                x[j] = gl::add(u, v);
                x[j + (1 << t)] = unit ? gl::sub(u, v) : gl::mul(gl::sub(u, v), wj);
            } else {
                const uint64_t tv = unit ? v : gl::mul(v, wj);
                x[j] = gl::add(u, tv);
                x[j + (1 << t)] = gl::sub(u, tv);
            }
        }
    }
}

// ---- radix-8 rounds: a constant-root network and seven twiddle products --------------------
// Three stages at global bits b, b+1, b+2 over x[j] = element base + j 2^b use the twiddles
// Z[4G + u] = V w8^{rev2(u)} (u < 4), Z[2G + u] = V^2 w4^u (u < 2) and Z[G] = V^4, with
// G = p >> (b + 3), V = Z[4G] and w8, w4 the table's 8th and 4th roots of unity (rev over the
// table's bits puts u's bits on top). Pulling V out, the GS round is the 8-point network with
// the twiddles w8^{rev2(u)}, w4^u and 1 followed by x[m] *= V^m; the CT round (the transpose)
// is x[m] *= V^m followed by the same network run top stage first. The network's twiddles are
// +-2^24, +-2^48, +-2^72 (2^96 = -1): shifts, not products. A forward table has w8 = 2^120 =
// -2^24, w4 = 2^48, w8^3 = -2^72; an inverse one w8 = 2^72, w4 = -2^48, w8^3 = 2^24. Twelve
// general products per eight values become seven plus five shifts (none at all in a pass's top
// round, where G = 0 and V = 1), the same field values either way. The seven powers of V come
// from a table of eight words per G (vpow_table), read as one 64-byte row.

// The network twiddle K (1..3: w4, w8, w8^3; w4 is also the middle stage's) as (shift, sign).
template <bool INV, int K>
struct RootShift;
template <> struct RootShift<false, 1> { static constexpr int E = 48; static constexpr bool NEG = false; };
template <> struct RootShift<false, 2> { static constexpr int E = 24; static constexpr bool NEG = true; };
template <> struct RootShift<false, 3> { static constexpr int E = 72; static constexpr bool NEG = true; };
template <> struct RootShift<true, 1> { static constexpr int E = 48; static constexpr bool NEG = true; };
template <> struct RootShift<true, 2> { static constexpr int E = 72; static constexpr bool NEG = false; };
template <> struct RootShift<true, 3> { static constexpr int E = 24; static constexpr bool NEG = false; };

// GS butterfly (u + v, (u - v) c) with network twiddle K; ZU / ZV: u / v known to be zero.
template <bool INV, int K, bool ZU, bool ZV>
__device__ __forceinline__ void gs_pair(uint64_t &a, uint64_t &b) {
    if constexpr (ZU && ZV) {
        return;
    } else if constexpr (K == 0) {
        const uint64_t u = a, v = b;
        if constexpr (ZV) {
            b = u;
        } else if constexpr (ZU) {
            a = v;
            b = gl::neg(v);
        } else {
            a = gl::add(u, v);
            b = gl::sub(u, v);
        }
    } else {
        using S = RootShift<INV, K>;
        const uint64_t u = a, v = b;
        if constexpr (ZV) {  // (u, +-u 2^E)
            const uint64_t t = gl::mul_pow2<S::E>(u);
            b = S::NEG ? gl::neg(t) : t;
        } else if constexpr (ZU) {  // (v, -+v 2^E)
            const uint64_t t = gl::mul_pow2<S::E>(v);
            a = v;
            b = S::NEG ? t : gl::neg(t);
        } else {
            a = gl::add(u, v);
            b = gl::mul_pow2<S::E>(S::NEG ? gl::sub(v, u) : gl::sub(u, v));
        }
    }
}

// CT butterfly (u + c v, u - c v) with network twiddle K.
template <bool INV, int K>
__device__ __forceinline__ void ct_pair(uint64_t &a, uint64_t &b) {
    const uint64_t u = a;
    if constexpr (K == 0) {
        const uint64_t t = b;
        a = gl::add(u, t);
        b = gl::sub(u, t);
    } else {
        using S = RootShift<INV, K>;
        const uint64_t t = gl::mul_pow2<S::E>(b);
        a = S::NEG ? gl::sub(u, t) : gl::add(u, t);
        b = S::NEG ? gl::add(u, t) : gl::sub(u, t);
    }
}

// Zero flags after a stage at local bit t: a pair stays zero only if both of its values were.
constexpr uint32_t zero_after(uint32_t z, int t) {
    uint32_t out = 0;
    for (int j = 0; j < 8; j++) {
        if (j & (1 << t)) continue;
        const int k = j + (1 << t);
        if (((z >> j) & 1) && ((z >> k) & 1)) out |= (1u << j) | (1u << k);
    }
    return out;
}
constexpr bool zbit(uint32_t z, int j) { return ((z >> j) & 1) != 0; }

// The GS network; Z0: the inputs known to be zero (bit j for x[j]).
template <bool INV, uint32_t Z0>
__device__ __forceinline__ void gs8_network(uint64_t *x) {
    constexpr uint32_t Z1 = zero_after(Z0, 0), Z2 = zero_after(Z1, 1);
    gs_pair<INV, 0, zbit(Z0, 0), zbit(Z0, 1)>(x[0], x[1]);
    gs_pair<INV, 1, zbit(Z0, 2), zbit(Z0, 3)>(x[2], x[3]);
    gs_pair<INV, 2, zbit(Z0, 4), zbit(Z0, 5)>(x[4], x[5]);
    gs_pair<INV, 3, zbit(Z0, 6), zbit(Z0, 7)>(x[6], x[7]);
    gs_pair<INV, 0, zbit(Z1, 0), zbit(Z1, 2)>(x[0], x[2]);
    gs_pair<INV, 0, zbit(Z1, 1), zbit(Z1, 3)>(x[1], x[3]);
    gs_pair<INV, 1, zbit(Z1, 4), zbit(Z1, 6)>(x[4], x[6]);
    gs_pair<INV, 1, zbit(Z1, 5), zbit(Z1, 7)>(x[5], x[7]);
    gs_pair<INV, 0, zbit(Z2, 0), zbit(Z2, 4)>(x[0], x[4]);
    gs_pair<INV, 0, zbit(Z2, 1), zbit(Z2, 5)>(x[1], x[5]);
    gs_pair<INV, 0, zbit(Z2, 2), zbit(Z2, 6)>(x[2], x[6]);
    gs_pair<INV, 0, zbit(Z2, 3), zbit(Z2, 7)>(x[3], x[7]);
}

template <bool INV>
__device__ __forceinline__ void ct8_network(uint64_t *x) {
    ct_pair<INV, 0>(x[0], x[4]);
    ct_pair<INV, 0>(x[1], x[5]);
    ct_pair<INV, 0>(x[2], x[6]);
    ct_pair<INV, 0>(x[3], x[7]);
    ct_pair<INV, 0>(x[0], x[2]);
    ct_pair<INV, 0>(x[1], x[3]);
    ct_pair<INV, 1>(x[4], x[6]);
    ct_pair<INV, 1>(x[5], x[7]);
    ct_pair<INV, 0>(x[0], x[1]);
    ct_pair<INV, 1>(x[2], x[3]);
    ct_pair<INV, 2>(x[4], x[5]);
    ct_pair<INV, 3>(x[6], x[7]);
}

// The radix-8 round at local bits S0..S0+2 of a tile of NB stage bits (the R = 3 case of
// round_regs, same values). tv: the table's V powers, row G = (hi << (NB - S0 - 3)) |
// (base >> (S0 + 3)). ZM: GS inputs known to be zero.
template <bool GS, bool INV, int NB, int S0, bool TOP = false, uint32_t ZM = 0>
__device__ __forceinline__ void round8_regs(uint64_t *x, uint32_t base, uint64_t hi,
                                            const uint64_t *__restrict__ tv) {
    constexpr bool unit = TOP && S0 + 3 == NB;  // G = 0: V = 1
    uint64_t v[7];
    if constexpr (!unit) {
        const uint64_t G = (hi << (NB - S0 - 3)) | (base >> (S0 + 3));
        const uint64_t *vp = (const uint64_t *)__builtin_assume_aligned(tv + 8 * G, 64);
#pragma unroll
        for (int m = 0; m < 7; m++) v[m] = vp[m];
    }
    if constexpr (!GS) {
        if constexpr (!unit) {
#pragma unroll
            for (int m = 1; m < 8; m++) x[m] = gl::mul(x[m], v[m - 1]);
        }
        ct8_network<INV>(x);
    } else {
        gs8_network<INV, ZM>(x);
        if constexpr (!unit) {
#pragma unroll
            for (int m = 1; m < 8; m++) x[m] = gl::mul(x[m], v[m - 1]);
        }
    }
}

template <bool GS, bool INV, int R, int NB, int LOGC, int S0, bool TOP = false, int THREADS = TILE_THREADS>
__device__ __forceinline__ void stage_round_ct(uint64_t *lds, uint64_t hi, const uint64_t *__restrict__ tw,
                                               const uint64_t *__restrict__ tv, uint32_t tid) {
    constexpr int M = 1 << R;
    constexpr uint32_t groups = 1u << (NB - R + LOGC);
    constexpr uint32_t cmask = (1u << LOGC) - 1;
    constexpr uint32_t stride = 1u << (S0 + LOGC);  // between a group's consecutive elements
#pragma unroll
    for (uint32_t q0 = 0; q0 < groups; q0 += THREADS) {
        const uint32_t q = q0 + tid;
        if (groups % THREADS != 0 && q >= groups) break;
        const uint32_t col = q & cmask, m = q >> LOGC;
        const uint32_t base = ((m >> S0) << (S0 + R)) | (m & ((1u << S0) - 1));
        uint64_t *e0 = lds + ((base << LOGC) | col);
        uint64_t x[M];
#pragma unroll
        for (int j = 0; j < M; j++) x[j] = e0[j * stride];
        if constexpr (R == 3)
            round8_regs<GS, INV, NB, S0, TOP>(x, base, hi, tv);
        else
            round_regs<GS, R, NB, S0, TOP>(x, base, hi, tw);
#pragma unroll
        for (int j = 0; j < M; j++) e0[j * stride] = x[j];
    }
}

template <bool GS, bool INV, int NB, int LOGC, int DONE = 0, bool TOP = false, int THREADS = TILE_THREADS>
__device__ __forceinline__ void tile_stages_ct(uint64_t *lds, uint64_t hi, const uint64_t *__restrict__ tw,
                                               const uint64_t *__restrict__ tv) {
    if constexpr (DONE < NB) {
        constexpr int R = NB - DONE >= 3 ? 3 : NB - DONE;
        constexpr int S0 = GS ? DONE : NB - DONE - R;
        stage_round_ct<GS, INV, R, NB, LOGC, S0, TOP, THREADS>(lds, hi, tw, tv, threadIdx.x);
        __syncthreads();
        tile_stages_ct<GS, INV, NB, LOGC, DONE + R, TOP, THREADS>(lds, hi, tw, tv);
    }
}


// Stages over bits [lo, lo+nb) of every row of length 2^L. Element e of the tile is
// (mid = e >> logc, col = e & (C-1)); global index p = hi*2^(lo+nb) + mid*2^lo + lblk*C + col.
// NB > 0: nb = NB and logc = LOGC, known at compile time. THREADS: the workgroup size. INV: tw
// is an inverse table (w^{-rev(t)}); tv: its V powers (vpow_table).
template <bool GS, bool INV, int NB, int LOGC, bool TOP = false, int THREADS = TILE_THREADS>
__global__ void __launch_bounds__(THREADS)
k_tile_pass(const uint64_t *__restrict__ in, uint64_t *__restrict__ out, uint64_t in_stride,
            uint64_t out_stride, uint32_t L, uint32_t lo, uint32_t nb, uint32_t logc,
            const uint64_t *__restrict__ tw, const uint64_t *__restrict__ tv,
            const uint64_t *const *in_rows, uint64_t *const *out_rows) {
    extern __shared__ uint64_t lds[];
    const uint32_t C = 1u << logc;
    const uint32_t tiles_per_hi = 1u << (lo - logc);
    // Workgroups go to the 8 XCDs round-robin by linear id. Adjacent tiles read the two halves
    // of the same 128-byte lines (runs of 64 bytes); tile t runs on XCD t / (tiles / 8), so each
    // XCD takes a contiguous range of tiles. The 9-bit pass over 256 rows of 2^22 runs 6% faster
    // so; its HBM bytes (FETCH_SIZE, WRITE_SIZE) are the algorithmic 8 + 8 GiB either way.
    uint32_t bx = blockIdx.x;
    if ((gridDim.x & 7) == 0) bx = (bx & 7) * (gridDim.x >> 3) + (bx >> 3);
    const uint64_t hi = bx / tiles_per_hi;
    const uint64_t lblk = bx % tiles_per_hi;
    const uint64_t row = blockIdx.y;
    const uint64_t base = (hi << (lo + nb)) + (lblk << logc);
    const uint32_t tile = 1u << (nb + logc);
    // rows by stride, or (batched matrices) by a per-row pointer table
    const uint64_t *src = in_rows ? in_rows[row] : in + row * in_stride;
    uint64_t *dst = out_rows ? out_rows[row] : out + row * out_stride;

    for (uint32_t e = threadIdx.x; e < tile; e += THREADS) {
        uint64_t mid = e >> logc, col = e & (C - 1);
        lds[e] = src[base + (mid << lo) + col];
    }
    __syncthreads();

    if constexpr (NB > 0)
        tile_stages_ct<GS, INV, NB, LOGC, 0, TOP, THREADS>(lds, hi, tw, tv);
    else
        tile_stages<GS>(lds, nb, logc, hi, tw);

    for (uint32_t e = threadIdx.x; e < tile; e += THREADS) {
        uint64_t mid = e >> logc, col = e & (C - 1);
        dst[base + (mid << lo) + col] = lds[e];
    }
}

__device__ __forceinline__ uint32_t rev_bits(uint64_t x, uint32_t bits) {
    return bits == 0 ? 0u : (uint32_t)(__builtin_bitreverse64(x) >> (64 - bits));
}

// The fused middle of the encoding, one tile of 2^tk row values -> 2^(tk+r) codeword symbols.
// TK > 0: tk = TK and tk + log_r = TN, known at compile time.
template <int TK, int TN>
__global__ void __launch_bounds__(MID_THREADS)
k_encode_mid(const uint64_t *__restrict__ in, uint64_t in_stride, uint64_t *__restrict__ cw,
             uint64_t cw_stride, uint32_t log_k, uint32_t log_r, uint32_t tk,
             const uint64_t *__restrict__ zk_inv, const uint64_t *__restrict__ zn,
             const uint64_t *__restrict__ vk_inv, const uint64_t *__restrict__ vn,
             const uint64_t *__restrict__ spread, const uint64_t *__restrict__ masks,
             uint64_t P, const uint64_t *__restrict__ pow7, uint64_t shift_pow_k,
             const uint64_t *const *in_rows, uint64_t *const *cw_rows, uint64_t c_off) {
    extern __shared__ uint64_t lds[];
    const uint32_t Tk = 1u << tk, Tn = Tk << log_r, s = 1u << log_r;
    // x: the row, y: the tile (+ c_off), so the rows of one tile run together and the tile's
    // twiddles stay in L2.
    const uint64_t row = blockIdx.x, c = blockIdx.y + c_off;
    const uint64_t kbase = c << tk, nbase = c << (tk + log_r);
    const uint64_t *src = in_rows ? in_rows[row] : in + row * in_stride;

    for (uint32_t e = threadIdx.x; e < Tk; e += MID_THREADS) lds[e] = src[kbase + e];
    __syncthreads();

    // Cooley-Tukey stages b = tk-1 .. 0 of the size-k inverse transform. Radix-8 rounds leave
    // half the waves without a group here; they wait at the barrier while the other
    // workgroup's waves use the SIMDs (radix-4 rounds, which busy every wave, measured slower).
    if constexpr (TK > 0)
        tile_stages_ct<false, true, TK, 0, 0, false, MID_THREADS>(lds, c, zk_inv, vk_inv);
    else
        tile_stages<false>(lds, tk, 0, c, zk_inv);

    if constexpr (TK > 0) {
        if (P == 0) {
            // Unmasked rows: the spread puts entry t times 7^{rev_k(t)}/k at position s t and
            // zeros between, so the first GS round (bits 0..2, groups of 8 positions) starts from
            // registers: each group's 8 / s entries, zeros elsewhere, which the butterflies of
            // the round fold away. Every group writes all 8 of its positions.
            constexpr int S = 1 << (TN - TK), PERG = 8 / S;
            // positions j of a group with j % S != 0 hold zeros
            constexpr uint32_t ZEROS = S == 2 ? 0xAAu : S == 4 ? 0xEEu : 0xFEu;
            constexpr uint32_t GROUPS = 1u << (TN - 3);
            constexpr uint32_t GPT = (GROUPS + MID_THREADS - 1) / MID_THREADS;
            static_assert(TN - TK >= 1 && TN - TK <= 3 && TN >= 3, "expansion round shape");
            uint64_t v[GPT][PERG];
#pragma unroll
            for (uint32_t i = 0; i < GPT; i++) {
                const uint32_t g = threadIdx.x + i * MID_THREADS;
#pragma unroll
                for (int jj = 0; jj < PERG; jj++) {
                    const uint32_t t = g * PERG + jj;
                    v[i][jj] = g < GROUPS ? gl::mul(lds[t], spread[kbase + t]) : 0;
                }
            }
            __syncthreads();
#pragma unroll
            for (uint32_t i = 0; i < GPT; i++) {
                const uint32_t g = threadIdx.x + i * MID_THREADS;
                if (g >= GROUPS) break;
                uint64_t x[8];
#pragma unroll
                for (int j = 0; j < 8; j++) x[j] = (j % S == 0) ? v[i][j / S] : 0;
                round8_regs<true, false, TN, 0, false, ZEROS>(x, g << 3, c, vn);
#pragma unroll
                for (int j = 0; j < 8; j++) lds[(g << 3) + j] = x[j];
            }
            __syncthreads();
            tile_stages_ct<true, false, TN, 0, 3, false, MID_THREADS>(lds, c, zn, vn);
            uint64_t *dst = cw_rows ? cw_rows[row] : cw + row * cw_stride;
            for (uint32_t e = threadIdx.x; e < Tn; e += MID_THREADS) dst[nbase + e] = lds[e];
            return;
        }
    }

    // Spread (encoding.rs:323-340) and masks (encoding.rs:303-312): entry t times
    // 7^{rev_k(t)}/k to position s*t, mask terms at s*t and s*t + s/2, zero elsewhere.
    // Registers hold the k values while the n tile is written over them.
    constexpr int PER = (8192 >> 1) / MID_THREADS;  // at most Tk/threads values per thread
    uint64_t held[PER];
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t e = threadIdx.x + i * MID_THREADS;
        held[i] = e < Tk ? lds[e] : 0;
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < Tn; e += MID_THREADS) lds[e] = 0;
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; i++) {
        uint32_t e = threadIdx.x + i * MID_THREADS;
        if (e < Tk) {
            uint64_t t = kbase + e;
            uint64_t v = gl::mul(held[i], spread[t]);
            uint32_t mi = rev_bits(t, log_k);
            if (mi < P) {
                uint64_t scaled = gl::mul(masks[row * P + mi], pow7[mi]);
                v = gl::sub(v, scaled);
                lds[e * s + (s >> 1)] = gl::mul(scaled, shift_pow_k);
            }
            lds[e * s] = v;
        }
    }
    __syncthreads();

    // Gentleman-Sande stages b = 0 .. tk+r-1 of the size-n forward transform.
    if constexpr (TK > 0)
        tile_stages_ct<true, false, TN, 0, 0, false, MID_THREADS>(lds, c, zn, vn);
    else
        tile_stages<true>(lds, tk + log_r, 0, c, zn);

    uint64_t *dst = cw_rows ? cw_rows[row] : cw + row * cw_stride;
    for (uint32_t e = threadIdx.x; e < Tn; e += MID_THREADS) dst[nbase + e] = lds[e];
}

// The unmasked fused middle over RPW = 2^LOGR rows per workgroup (the values of k_encode_mid
// with P = 0). One row's 2^TK-value tile has 2^(TK-3) radix-8 groups per CT round, a quarter of
// the threads at rate 1/4, and the SIMDs idle through those rounds (counted: 75% VALU-busy over
// the kernel). RPW rows' tiles share the LDS tile in the CT stages, interleaved (word
// (t << LOGR) | r holds row r's value t), so RPW times the threads have a group there. The
// spread values of all RPW rows then wait in registers (RPW x 8 / 2^(TN-TK) words per thread)
// while the GS stages run on one row's codeword tile at a time: the current row's in registers,
// the others parked in LDS past the tile (RPW - 1 rows x 8 / 2^(TN-TK) words per thread). Two
// rows: 64 VGPRs and 80 KiB of LDS, two workgroups per CU, faster than one row at 256 x 2^20
// rate 1/4 (holding both rows' values in registers instead spilled 4 VGPRs to scratch memory:
// 2 GiB more HBM traffic per launch); four rows spill 22 registers and are slower.
template <int TK, int TN, int LOGR>
__global__ void __launch_bounds__(MID_THREADS) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_encode_mid_rows(const uint64_t *__restrict__ in, uint64_t in_stride, uint64_t *__restrict__ cw,
                  uint64_t cw_stride, const uint64_t *__restrict__ zk_inv, const uint64_t *__restrict__ zn,
                  const uint64_t *__restrict__ vk_inv, const uint64_t *__restrict__ vn,
                  const uint64_t *__restrict__ spread, const uint64_t *const *in_rows,
                  uint64_t *const *cw_rows, uint64_t c_off) {
    constexpr int RPW = 1 << LOGR, S = 1 << (TN - TK), PERG = 8 / S;
    constexpr uint32_t Tk = 1u << TK, Tn = 1u << TN;
    static_assert(S >= 2 && S <= 4, "rate 1/2 or 1/4");
    static_assert((Tk << LOGR) <= Tn, "the rows' tiles fit the codeword tile");
    static_assert((1u << (TN - 3)) == MID_THREADS, "one first-round GS group per thread");
    extern __shared__ uint64_t lds[];
    const uint64_t row0 = (uint64_t)blockIdx.x * RPW, c = blockIdx.y + c_off;
    const uint64_t kbase = c << TK, nbase = c << TN;
    // each wave reads runs of 64 / RPW consecutive values from each of the RPW rows
    for (uint32_t e = threadIdx.x; e < (Tk << LOGR); e += MID_THREADS) {
        const uint32_t r = e & (RPW - 1), t = e >> LOGR;
        const uint64_t *src = in_rows ? in_rows[row0 + r] : in + (row0 + r) * in_stride;
        lds[e] = src[kbase + t];
    }
    __syncthreads();
    tile_stages_ct<false, true, TK, LOGR, 0, false, MID_THREADS>(lds, c, zk_inv, vk_inv);

    // Thread g's first GS group (positions 8g..8g+7 of the codeword tile) holds the spread values
    // of k-positions t = g PERG + jj at positions S jj, zeros elsewhere. Row 0's stay in v, row
    // r > 0's go to park[((r - 1) PERG + jj) MID_THREADS + g] (each thread reads back only what
    // it wrote: no barrier orders them).
    const uint32_t g = threadIdx.x;
    uint64_t *park = lds + Tn;
    uint64_t v[PERG];
#pragma unroll
    for (int jj = 0; jj < PERG; jj++) {
        const uint32_t t = g * PERG + jj;
        const uint64_t f = spread[kbase + t];
        v[jj] = gl::mul(lds[t << LOGR], f);
#pragma unroll
        for (int r = 1; r < RPW; r++)
            park[((r - 1) * PERG + jj) * MID_THREADS + g] = gl::mul(lds[(t << LOGR) | r], f);
    }
    constexpr uint32_t ZEROS = S == 2 ? 0xAAu : 0xEEu;
    // One row per iteration, not unrolled: the row's values are always v, reloaded from the park
    // at the end of the iteration (so nothing of the next row is live through the GS stages).
#pragma unroll 1
    for (int r = 0; r < RPW; r++) {
        __syncthreads();  // the CT tile, or the previous row's codeword tile, has been read
        // an empty asm statement the compiler must assume changes the table pointer, so the
        // twiddle loads stay inside the iteration instead of being hoisted (and held) for all rows
        const uint64_t *vnr = vn;
        asm volatile("" : "+s"(vnr));
        uint64_t x[8];
#pragma unroll
        for (int j = 0; j < 8; j++) x[j] = (j % S == 0) ? v[j / S] : 0;
        round8_regs<true, false, TN, 0, false, ZEROS>(x, g << 3, c, vnr);
#pragma unroll
        for (int j = 0; j < 8; j++) lds[(g << 3) + j] = x[j];
        __syncthreads();
        tile_stages_ct<true, false, TN, 0, 3, false, MID_THREADS>(lds, c, zn, vnr);
        uint64_t *dst = cw_rows ? cw_rows[row0 + r] : cw + (row0 + r) * cw_stride;
        for (uint32_t e = threadIdx.x; e < Tn; e += MID_THREADS) dst[nbase + e] = lds[e];
        if (r + 1 < RPW) {
            // the thread's park address recomputed here (an empty asm statement hides that it
            // equals g), not held through the GS stages
            uint32_t gg = threadIdx.x;
            asm volatile("" : "+v"(gg));
#pragma unroll
            for (int jj = 0; jj < PERG; jj++) v[jj] = park[(r * PERG + jj) * MID_THREADS + gg];
        }
    }
}

// ---------------------------------------------------------------------------------------------

using TilePassFn = void (*)(const uint64_t *, uint64_t *, uint64_t, uint64_t, uint32_t, uint32_t,
                            uint32_t, uint32_t, const uint64_t *, const uint64_t *,
                            const uint64_t *const *, uint64_t *const *);
using MidFn = void (*)(const uint64_t *, uint64_t, uint64_t *, uint64_t, uint32_t, uint32_t, uint32_t,
                       const uint64_t *, const uint64_t *, const uint64_t *, const uint64_t *,
                       const uint64_t *, const uint64_t *, uint64_t, const uint64_t *, uint64_t,
                       const uint64_t *const *, uint64_t *const *, uint64_t);

// The compile-time tile pass for (nb, 2^(12 - nb) columns), else the general one. top: the pass
// covers the row's top bits (its tiles have hi = 0).
template <bool GS, bool INV, bool TOP>
static TilePassFn tile_pass_nb(uint32_t nb) {
    switch (nb) {
        case 1: return k_tile_pass<GS, INV, 1, PASS_TILE_LOG - 1, TOP>;
        case 2: return k_tile_pass<GS, INV, 2, PASS_TILE_LOG - 2, TOP>;
        case 3: return k_tile_pass<GS, INV, 3, PASS_TILE_LOG - 3, TOP>;
        case 4: return k_tile_pass<GS, INV, 4, PASS_TILE_LOG - 4, TOP>;
        case 5: return k_tile_pass<GS, INV, 5, PASS_TILE_LOG - 5, TOP>;
        case 6: return k_tile_pass<GS, INV, 6, PASS_TILE_LOG - 6, TOP>;
        case 7: return k_tile_pass<GS, INV, 7, PASS_TILE_LOG - 7, TOP>;
        case 8: return k_tile_pass<GS, INV, 8, PASS_TILE_LOG - 8, TOP>;
        case 9: return k_tile_pass<GS, INV, 9, PASS_TILE_LOG - 9, TOP>;
        default: return nullptr;
    }
}

// A 10-bit pass: 2^10 strided points x 8 columns, 64 KiB of LDS (two workgroups per CU), so
// 1024 threads give the CU its 32 waves. Ten stage bits then cost one read and one write of
// the rows instead of two of each.
constexpr uint32_t WIDE_PASS_NB = 10, WIDE_PASS_LOGC = 3, WIDE_PASS_THREADS = 1024;

static bool wide_pass(uint32_t nb, uint32_t logc) {
    return nb == WIDE_PASS_NB && logc == WIDE_PASS_LOGC;
}

template <bool GS, bool INV>
static TilePassFn tile_pass_fn(uint32_t nb, uint32_t logc, bool top) {
    if (wide_pass(nb, logc))
        return top ? k_tile_pass<GS, INV, WIDE_PASS_NB, WIDE_PASS_LOGC, true, WIDE_PASS_THREADS>
                   : k_tile_pass<GS, INV, WIDE_PASS_NB, WIDE_PASS_LOGC, false, WIDE_PASS_THREADS>;
    TilePassFn f = nullptr;
    if (nb <= 9 && logc == PASS_TILE_LOG - nb)
        f = top ? tile_pass_nb<GS, INV, true>(nb) : tile_pass_nb<GS, INV, false>(nb);
    return f ? f : k_tile_pass<GS, INV, 0, 0>;
}

static uint32_t tile_pass_threads(uint32_t nb, uint32_t logc) {
    return wide_pass(nb, logc) ? WIDE_PASS_THREADS : TILE_THREADS;
}

// Column bits of a pass over bits [lo, lo + nb): the tile's 2^12 elements (2^13 for a 10-bit
// pass), or as many columns as lie below lo.
static uint32_t pass_logc(uint32_t lo, uint32_t nb) {
    const uint32_t want = nb == WIDE_PASS_NB ? WIDE_PASS_LOGC : nb < PASS_TILE_LOG ? PASS_TILE_LOG - nb : 0;
    return lo < want ? lo : want;
}

// The compile-time fused middle for the full tiles of tile_log 12 and 13 at rates 1/2 and 1/4.
static MidFn mid_fn(uint32_t tk, uint32_t log_r) {
    const uint32_t tn = tk + log_r;
    if (tk == 11 && tn == 13) return k_encode_mid<11, 13>;
    if (tk == 12 && tn == 13) return k_encode_mid<12, 13>;
    if (tk == 10 && tn == 12) return k_encode_mid<10, 12>;
    if (tk == 11 && tn == 12) return k_encode_mid<11, 12>;
    return k_encode_mid<0, 0>;
}

using MidRowsFn = void (*)(const uint64_t *, uint64_t, uint64_t *, uint64_t, const uint64_t *,
                           const uint64_t *, const uint64_t *, const uint64_t *, const uint64_t *,
                           const uint64_t *const *, uint64_t *const *, uint64_t);

#ifndef TSR_MID_ROWS_LOG
#define TSR_MID_ROWS_LOG 1
#endif
// The several-rows fused middle (2^TSR_MID_ROWS_LOG rows per workgroup), where one exists:
// unmasked, tile_log 13, rate 1/4 (at rate 1/2 a thread would hold twice the spread values).
// LDS past the codeword tile that k_encode_mid_rows parks the later rows' spread values in.
static size_t mid_rows_park_bytes(uint32_t log_r, uint32_t rows_log) {
    const uint32_t perg = 8u >> log_r;  // spread values per thread and row
    return sizeof(uint64_t) * (((size_t)1 << rows_log) - 1) * perg * MID_THREADS;
}

static MidRowsFn mid_rows_fn(uint32_t tk, uint32_t log_r, uint64_t P, uint32_t *rows_log) {
    *rows_log = TSR_MID_ROWS_LOG;
#if TSR_MID_THREADS == 1024 && TSR_MID_ROWS_LOG > 0
    if (P == 0 && tk == 11 && log_r == 2) return k_encode_mid_rows<11, 13, TSR_MID_ROWS_LOG>;
#endif
    return nullptr;
}

static void bit_reversed_powers(uint64_t base, uint32_t bits, uint64_t scale,
                                std::vector<uint64_t> &out) {
    std::vector<uint64_t> squares(bits);
    uint64_t sq = base;
    for (uint32_t i = 0; i < bits; i++) {
        squares[i] = sq;
        sq = gl::mul(sq, sq);
    }
    out.assign(1, scale);
    out.reserve((size_t)1 << bits);
    for (int e = (int)bits - 1; e >= 0; e--) {
        size_t len = out.size();
        for (size_t t = 0; t < len; t++) out.push_back(gl::mul(out[t], squares[e]));
    }
}

// Z[t] = w^{rev(t)} for t < n/2 (ntt.rs:110-125).
static std::vector<uint64_t> twiddle_table(uint32_t log_n, bool inverse) {
    std::vector<uint64_t> out;
    if (log_n == 0) return std::vector<uint64_t>(1, 1);
    uint64_t root = gl::root(log_n);
    uint64_t base = inverse ? gl::pow(root, (1ull << log_n) - 1) : root;
    bit_reversed_powers(base, log_n - 1, 1, out);
    return out;
}

// The radix-8 rounds' V powers of a twiddle table Z of 2^(L-1) entries: row G < 2^(L-3) holds
// Z[4G]^m for m = 1..7 and a zero, eight words (64 bytes) per row. One row for L < 3 (no radix-8
// round then).
static std::vector<uint64_t> vpow_table(const std::vector<uint64_t> &Z, uint32_t L) {
    const uint64_t rows = L >= 3 ? 1ull << (L - 3) : 1;
    std::vector<uint64_t> out(8 * rows, 0);
    for (uint64_t G = 0; G < rows; G++) {
        const uint64_t V = L >= 3 ? Z[4 * G] : 1;
        uint64_t acc = V;
        for (int m = 0; m < 7; m++) {
            out[8 * G + m] = acc;
            acc = gl::mul(acc, V);
        }
    }
    return out;
}

static std::vector<std::pair<uint32_t, uint32_t>> split_bits(uint32_t lo, uint32_t hi,
                                                             uint32_t max_nb) {
    std::vector<std::pair<uint32_t, uint32_t>> passes;
    uint32_t bits = hi > lo ? hi - lo : 0;
    if (!bits) return passes;
    uint32_t count = (bits + max_nb - 1) / max_nb;
    uint32_t at = lo;
    for (uint32_t i = 0; i < count; i++) {
        uint32_t nb = bits / count + (i < bits % count ? 1 : 0);
        passes.push_back({at, nb});
        at += nb;
    }
    return passes;
}

Status Encoder::init(Ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P, uint32_t tile_log,
                     uint32_t max_nb_in) {
    this->ctx = ctx;
    if (tile_log < log_r + 1) tile_log = log_r + 1;
    if (tile_log > TILE_MAX_LOG) tile_log = TILE_MAX_LOG;
    this->tile_log = tile_log;
    this->log_k = log_k;
    this->log_r = log_r;
    this->P = P;
    const uint32_t log_n = log_k + log_r;
    const uint64_t k = 1ull << log_k;
    tk = log_k < tile_log - log_r ? log_k : tile_log - log_r;
    // CT passes over k-bits [tk, log_k), run from the top bits down; up to 10 bits per pass
    // with 64 KiB tiles (tile_log 13), else tile_log - 4
    const uint32_t max_nb_default = tile_log >= 13 ? WIDE_PASS_NB : tile_log - 4;
    const uint32_t max_nb = max_nb_in && max_nb_in < max_nb_default ? max_nb_in : max_nb_default;
    auto ct = split_bits(tk, log_k, max_nb);
    ct_passes.assign(ct.rbegin(), ct.rend());
    gs_passes = split_bits(tk + log_r, log_n, max_nb);

    std::vector<uint64_t> zki = twiddle_table(log_k, true);
    std::vector<uint64_t> zn = twiddle_table(log_n, false);
    std::vector<uint64_t> spread;
    bit_reversed_powers(7, log_k, gl::pow(gl::P / 2 + 1, log_k), spread);
    std::vector<uint64_t> pow7(P ? P : 1, 1);
    for (uint64_t i = 1; i < P; i++) pow7[i] = gl::mul(pow7[i - 1], 7);
    shift_pow_k = gl::pow(7, k);

    Status st;
    if ((st = zk_inv.upload(ctx, zki))) return st;
    if ((st = zn_fwd.upload(ctx, zn))) return st;
    if ((st = vk_inv.upload(ctx, vpow_table(zki, log_k)))) return st;
    if ((st = vn_fwd.upload(ctx, vpow_table(zn, log_n)))) return st;
    if ((st = spread_tab.upload(ctx, spread))) return st;
    if ((st = pow7_tab.upload(ctx, pow7))) return st;
    return Status::ok();
}

uint64_t Encoder::scratch_elems(uint64_t count) const {
    return ct_passes.empty() ? 0 : count << log_k;
}

Status Encoder::encode(const uint64_t *rows, const uint64_t *masks, uint64_t count,
                       uint64_t *codewords, uint64_t *scratch, hipStream_t stream) const {
    const uint64_t k = 1ull << log_k, n = k << log_r;
    for (uint64_t r0 = 0; r0 < count; r0 += 65535) {
        const uint64_t rc = count - r0 < 65535 ? count - r0 : 65535;
        launch(rows + r0 * k, k, nullptr, scratch + r0 * k, nullptr, codewords + r0 * n, nullptr,
               masks ? masks + r0 * P : nullptr, rc, stream);
    }
    return check_launch("encode_rows");
}

Status Encoder::encode_rowmap(const uint64_t *const *src_rows, uint64_t *const *tmp_rows,
                              uint64_t *const *cw_rows, uint64_t count, hipStream_t stream) const {
    for (uint64_t r0 = 0; r0 < count; r0 += 65535) {
        const uint64_t rc = count - r0 < 65535 ? count - r0 : 65535;
        launch(nullptr, 0, src_rows + r0, nullptr, tmp_rows + r0, nullptr, cw_rows + r0, nullptr,
               rc, stream);
    }
    return check_launch("encode_rows (batched)");
}

// One group of up to 65535 rows: the strided CT passes (rows -> scratch, then in place), the
// fused middle (-> codewords), the strided GS passes (in place). Rows are addressed by stride
// (src, tmp, cw) or by per-row pointer tables (src_rows, tmp_rows, cw_rows).
void Encoder::launch(const uint64_t *src, uint64_t src_stride, const uint64_t *const *src_rows,
                     uint64_t *tmp, uint64_t *const *tmp_rows, uint64_t *cw,
                     uint64_t *const *cw_rows, const uint64_t *masks, uint64_t rc,
                     hipStream_t stream) const {
    const uint64_t k = 1ull << log_k, n = k << log_r;
    for (auto &ps : ct_passes) {
        uint32_t lo = ps.first, nb = ps.second, logc = pass_logc(lo, nb);
        dim3 grid((unsigned)(k >> (nb + logc)), (unsigned)rc);
        size_t shm = sizeof(uint64_t) << (nb + logc);
        KernelTimer timer(ctx, "k_tile_pass<ct>", stream);
        const TilePassFn f = tile_pass_fn<false, true>(nb, logc, lo + nb == log_k);
        hipLaunchKernelGGL(f, grid,
                           dim3(tile_pass_threads(nb, logc)), shm, stream, src, tmp, src_stride, k, log_k,
                           lo, nb, logc, zk_inv.ptr, vk_inv.ptr, src_rows, tmp_rows);
        src = tmp;
        src_stride = k;
        src_rows = (const uint64_t *const *)tmp_rows;
    }
    {
        const uint64_t tiles = k >> tk;
        size_t shm = sizeof(uint64_t) << (tk + log_r);
        KernelTimer timer(ctx, "k_encode_mid", stream);
        // whole groups of rows by the several-rows middle, the rest one row per workgroup
        uint32_t rows_log = 0;
        const MidRowsFn fr = mid_rows_fn(tk, log_r, P, &rows_log);
        const uint64_t rpw = 1ull << rows_log, grouped = fr ? rc - rc % rpw : 0;
        for (uint64_t c0 = 0; c0 < tiles; c0 += 65535) {
            const uint64_t tc = tiles - c0 < 65535 ? tiles - c0 : 65535;
            if (grouped)
                hipLaunchKernelGGL(fr, dim3((unsigned)(grouped / rpw), (unsigned)tc), dim3(MID_THREADS),
                                   shm + mid_rows_park_bytes(log_r, rows_log),
                                   stream, src, src_stride, cw, n, zk_inv.ptr, zn_fwd.ptr, vk_inv.ptr,
                                   vn_fwd.ptr, spread_tab.ptr, src_rows, cw_rows, c0);
            if (grouped < rc)
                hipLaunchKernelGGL(mid_fn(tk, log_r), dim3((unsigned)(rc - grouped), (unsigned)tc),
                                   dim3(MID_THREADS), shm, stream,
                                   src_rows ? src : src + grouped * src_stride, src_stride,
                                   cw_rows ? cw : cw + grouped * n, n, log_k, log_r, tk, zk_inv.ptr,
                                   zn_fwd.ptr, vk_inv.ptr, vn_fwd.ptr, spread_tab.ptr,
                                   masks ? masks + grouped * P : masks, P, pow7_tab.ptr, shift_pow_k,
                                   src_rows ? src_rows + grouped : src_rows,
                                   cw_rows ? cw_rows + grouped : cw_rows, c0);
        }
    }
    for (auto &ps : gs_passes) {
        uint32_t lo = ps.first, nb = ps.second, logc = pass_logc(lo, nb);
        uint32_t log_n = log_k + log_r;
        dim3 grid((unsigned)(n >> (nb + logc)), (unsigned)rc);
        size_t shm = sizeof(uint64_t) << (nb + logc);
        KernelTimer timer(ctx, "k_tile_pass<gs>", stream);
        const TilePassFn f = tile_pass_fn<true, false>(nb, logc, lo + nb == log_n);
        hipLaunchKernelGGL(f, grid,
                           dim3(tile_pass_threads(nb, logc)), shm, stream, cw, cw, n, n, log_n, lo, nb,
                           logc, zn_fwd.ptr, vn_fwd.ptr, (const uint64_t *const *)cw_rows, cw_rows);
    }
}

Status get_encoder(Ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P, Encoder **out,
                   uint32_t tile_log, uint32_t max_nb) {
    auto key = std::make_tuple(log_k, log_r, P, tile_log | (max_nb << 8));
    auto it = ctx->encoders.find(key);
    if (it != ctx->encoders.end()) {
        *out = it->second.get();
        return Status::ok();
    }
    std::shared_ptr<Encoder> e = std::make_shared<Encoder>();
    Status st = e->init(ctx, log_k, log_r, P, tile_log, max_nb);
    if (st) return st;
    *out = e.get();
    ctx->encoders.emplace(key, std::move(e));
    return Status::ok();
}

}  // namespace tsr

// ---------------------------------------------------------------------------------------------
// NttPlan (crates/tessera-pcs/src/ntt.rs:397-516) on the device: transforms of size n = 2^L, in
// place, natural order on both sides, over a batch of rows. The same CT / GS stage passes as the
// encoder (k_tile_pass), the bit reversal as tiles through LDS, the coset shift as a scaling by
// s^i:
//   ntt        = CT(forward table), then bit reversal
//   intt       = bit reversal times 1/n, then GS(inverse table)
//   coset_ntt  = values[i] *= s^i, then ntt          (f(s w^j) = sum_i (f_i s^i) w^{ij})
//   coset_intt = intt, then values[i] *= s^{-i}      (the coefficients of f(sX) are f_i s^i)
// equal, value for value, to the reference's Coset twiddles (ntt.rs:157-180), which apply the
// same powers of s inside the butterflies.
namespace tsr {

constexpr uint32_t BR_BITS = 5, BR_T = 1u << BR_BITS;  // bit-reversal tiles of 32 x 32

__device__ __forceinline__ uint64_t rev_low(uint64_t x, uint32_t bits) {
    return bits ? (__brevll(x) >> (64 - bits)) : 0;
}

// Rows of 2^L >= 2^(2 BR_BITS) values, in place. Index i = (h, m, l) with h and l of BR_BITS bits
// around a middle m of M = L - 2 BR_BITS bits: rev(i) = (rev(l), rev(m), rev(h)), so the 32 x 32
// block of middle m (rows h, columns l: runs of 32 consecutive words) lands transposed, both
// parts reversed, in the block of rev(m). Workgroup m < 2^M handles the pair (m, rev(m)) when
// m <= rev(m); the pair owns both blocks, so no other workgroup reads or writes them.
__global__ void __launch_bounds__(256)
k_bitrev_tiles(uint64_t *__restrict__ v, uint64_t row_stride, uint32_t L, uint64_t scale, int scaled) {
    const uint32_t M = L - 2 * BR_BITS;
    const uint64_t m = blockIdx.x, rm = rev_low(m, M);
    if (rm < m) return;
    uint64_t *row = v + (uint64_t)blockIdx.y * row_stride;
    __shared__ uint64_t A[BR_T][BR_T + 1], Bt[BR_T][BR_T + 1];
    const uint32_t hs = M + BR_BITS;  // the shift of h (and of the output row r)
    for (uint32_t e = threadIdx.x; e < BR_T * BR_T; e += blockDim.x) {
        const uint32_t h = e >> BR_BITS, l = e & (BR_T - 1);
        A[h][l] = row[((uint64_t)h << hs) | (m << BR_BITS) | l];
        if (rm != m) Bt[h][l] = row[((uint64_t)h << hs) | (rm << BR_BITS) | l];
    }
    __syncthreads();
    for (uint32_t e = threadIdx.x; e < BR_T * BR_T; e += blockDim.x) {
        const uint32_t r = e >> BR_BITS, c = e & (BR_T - 1);  // output row rev(l), column rev(h)
        const uint32_t l = (uint32_t)rev_low(r, BR_BITS), h = (uint32_t)rev_low(c, BR_BITS);
        const uint64_t x = A[h][l];
        row[((uint64_t)r << hs) | (rm << BR_BITS) | c] = scaled ? gl::mul(x, scale) : x;
        if (rm != m) {
            const uint64_t y = Bt[h][l];
            row[((uint64_t)r << hs) | (m << BR_BITS) | c] = scaled ? gl::mul(y, scale) : y;
        }
    }
}

// Rows of 2^L < 2^(2 BR_BITS) values: one workgroup per row through LDS.
__global__ void __launch_bounds__(256)
k_bitrev_small(uint64_t *__restrict__ v, uint64_t row_stride, uint32_t L, uint64_t scale, int scaled) {
    __shared__ uint64_t s[BR_T * BR_T];
    uint64_t *row = v + (uint64_t)blockIdx.x * row_stride;
    const uint32_t n = 1u << L;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) s[i] = row[i];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        const uint64_t x = s[rev_low(i, L)];
        row[i] = scaled ? gl::mul(x, scale) : x;
    }
}

// values[i] *= s^i within each row of 2^L: s^i = lo[i mod 2^a] * hi[i >> a].
__global__ void __launch_bounds__(256)
k_scale_powers(uint64_t *__restrict__ v, uint64_t total, uint32_t L, const uint64_t *__restrict__ lo,
               const uint64_t *__restrict__ hi, uint32_t a) {
    const uint64_t mask_n = (1ull << L) - 1, mask_a = (1ull << a) - 1;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < total;
         i += (uint64_t)gridDim.x * blockDim.x) {
        const uint64_t j = i & mask_n;
        v[i] = gl::mul(v[i], gl::mul(lo[j & mask_a], hi[j >> a]));
    }
}

struct NttPlanDev {
    Ctx *ctx = nullptr;
    uint32_t L = 0;
    DevBuf<uint64_t> fwd, inv;  // Z[t] = w^{rev(t)}, w^{-rev(t)}, t < n/2
    DevBuf<uint64_t> vfwd, vinv;  // their V powers (vpow_table)
    uint64_t size_inv = 1;
    std::vector<std::pair<uint32_t, uint32_t>> passes;  // stage-bit ranges, ascending

    Status init(Ctx *c, uint32_t log_n) {
        ctx = c;
        L = log_n;
        Status st;
        {
            const std::vector<uint64_t> zf = twiddle_table(L, false), zi = twiddle_table(L, true);
            if ((st = fwd.upload(ctx, zf)) || (st = inv.upload(ctx, zi))) return st;
            if ((st = vfwd.upload(ctx, vpow_table(zf, L))) || (st = vinv.upload(ctx, vpow_table(zi, L))))
                return st;
        }
        size_inv = gl::pow(gl::P / 2 + 1, L);
        // the low bits in one pass of contiguous tiles (2^12 values), the rest in passes of up
        // to 10 strided bits
        const uint32_t low = L < 12 ? L : 12;
        if (low) passes.push_back({0, low});
        for (auto &p : split_bits(low, L, WIDE_PASS_NB)) passes.push_back(p);
        return Status::ok();
    }

    // CT stages top-down (GS = false) or GS stages bottom-up (GS = true) over `rows` rows.
    void stages(bool gs, uint64_t *v, uint64_t rows, hipStream_t s) const {
        const uint64_t n = 1ull << L;
        for (size_t i = 0; i < passes.size(); i++) {
            const auto &ps = gs ? passes[i] : passes[passes.size() - 1 - i];
            const uint32_t lo = ps.first, nb = ps.second, logc = pass_logc(lo, nb);
            const size_t shm = sizeof(uint64_t) << (nb + logc);
            const bool top = lo + nb == L;
            // ntt: CT with the forward table; intt: GS with the inverse table
            TilePassFn f = gs ? tile_pass_fn<true, true>(nb, logc, top) : tile_pass_fn<false, false>(nb, logc, top);
            for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
                const uint64_t rc = rows - r0 < 65535 ? rows - r0 : 65535;
                KernelTimer timer(ctx, gs ? "k_tile_pass<gs>" : "k_tile_pass<ct>", s);
                hipLaunchKernelGGL(f, dim3((unsigned)(n >> (nb + logc)), (unsigned)rc),
                                   dim3(tile_pass_threads(nb, logc)), shm, s, v + r0 * n, v + r0 * n, n, n,
                                   L, lo, nb, logc, gs ? inv.ptr : fwd.ptr, gs ? vinv.ptr : vfwd.ptr,
                                   nullptr, nullptr);
            }
        }
    }

    void bitrev(uint64_t *v, uint64_t rows, bool scaled, hipStream_t s) const {
        const uint64_t n = 1ull << L;
        for (uint64_t r0 = 0; r0 < rows; r0 += 65535) {
            const uint64_t rc = rows - r0 < 65535 ? rows - r0 : 65535;
            KernelTimer timer(ctx, "k_bitrev", s);
            if (L >= 2 * BR_BITS)
                hipLaunchKernelGGL(k_bitrev_tiles, dim3((unsigned)(1ull << (L - 2 * BR_BITS)), (unsigned)rc),
                                   dim3(256), 0, s, v + r0 * n, n, L, size_inv, scaled ? 1 : 0);
            else
                hipLaunchKernelGGL(k_bitrev_small, dim3((unsigned)rc), dim3(256), 0, s, v + r0 * n, n, L,
                                   size_inv, scaled ? 1 : 0);
        }
    }

    // values[i] *= shift^i in every row; the two power tables go up with the call
    Status scale_powers(uint64_t *v, uint64_t rows, uint64_t shift, hipStream_t s) const {
        const uint32_t a = (L + 1) / 2;
        std::vector<uint64_t> lo(1ull << a), hi(1ull << (L - a));
        lo[0] = 1;
        for (size_t i = 1; i < lo.size(); i++) lo[i] = gl::mul(lo[i - 1], shift);
        const uint64_t step = gl::pow(shift, 1ull << a);
        hi[0] = 1;
        for (size_t i = 1; i < hi.size(); i++) hi[i] = gl::mul(hi[i - 1], step);
        DevBuf<uint64_t> dlo, dhi;
        Status st;
        if ((st = dlo.upload(ctx, lo)) || (st = dhi.upload(ctx, hi))) return st;
        const uint64_t total = rows << L;
        const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 1u << 16);
        KernelTimer timer(ctx, "k_scale_powers", s);
        hipLaunchKernelGGL(k_scale_powers, dim3(blocks), dim3(256), 0, s, v, total, L, dlo.ptr, dhi.ptr, a);
        if ((st = check_launch("scale powers"))) return st;
        // the tables go back to the context's pool, which other streams may draw from
        TSR_HIP(tsr::stream_sync(s));
        return Status::ok();
    }
};

}  // namespace tsr

struct tsr_ntt_plan {
    tsr::NttPlanDev p;
};

using namespace tsr;

namespace {
enum NttKind { NTT_FWD, NTT_INV, NTT_COSET_FWD, NTT_COSET_INV };

tsr_status run_ntt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len, NttKind kind, uint64_t shift) {
    if (!plan || (len && !values)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    const NttPlanDev &p = plan->p;
    const uint64_t n = 1ull << p.L;
    // check (ntt.rs:441-450), then the shift (ZeroShift), before the values are touched
    if (len == 0 || len % n)
        return Status::fail(TSR_BAD_SHAPE, "TransformLength: expected a non-zero multiple of " +
                                               std::to_string(n) + " values, found " + std::to_string(len)).code;
    if ((kind == NTT_COSET_FWD || kind == NTT_COSET_INV) && shift == 0)
        return Status::fail(TSR_BAD_SHAPE, "ZeroShift").code;
    if (shift >= gl::P) return Status::fail(TSR_BAD_ARG, "the shift is not canonical").code;
    Ctx *ctx = p.ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    hipStream_t s = ctx->stream;
    const uint64_t rows = len / n;
    uint64_t *v = values;
    DevBuf<uint64_t> staged;
    Status st;
    const bool host = !is_device_ptr(values);
    if (host) {
        if ((st = staged.alloc(ctx, len))) return st.code;
        e = hipMemcpyAsync(staged.ptr, values, 8 * len, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy").code;
        v = staged.ptr;
    }
    if (p.L > 0) {
        switch (kind) {
            case NTT_COSET_FWD:
                if ((st = p.scale_powers(v, rows, shift, s))) return st.code;
                // fall through
            case NTT_FWD:
                p.stages(false, v, rows, s);
                p.bitrev(v, rows, false, s);
                break;
            case NTT_INV:
            case NTT_COSET_INV:
                p.bitrev(v, rows, true, s);
                p.stages(true, v, rows, s);
                if (kind == NTT_COSET_INV && (st = p.scale_powers(v, rows, gl::inv(shift), s))) return st.code;
                break;
        }
        if ((st = check_launch("ntt"))) return st.code;
    }
    if (host) {
        e = hipMemcpyAsync(values, staged.ptr, 8 * len, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "copy").code;
    }
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    return TSR_OK;
}
}  // namespace

extern "C" {

tsr_status tsr_ntt_plan_create(tsr_ctx *ctx_, uint64_t size, tsr_ntt_plan **out) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    // NttPlan::new (ntt.rs:430-437): log2_len
    if (size == 0 || (size & (size - 1)))
        return Status::fail(TSR_BAD_SHAPE, "NotPowerOfTwo: " + std::to_string(size)).code;
    if (size > (1ull << 32)) return Status::fail(TSR_BAD_SHAPE, "TooLong: " + std::to_string(size)).code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    std::unique_ptr<tsr_ntt_plan> p(new tsr_ntt_plan());
    Status st = p->p.init(ctx, (uint32_t)__builtin_ctzll(size));
    if (st) return st.code;
    *out = p.release();
    return TSR_OK;
}

uint64_t tsr_ntt_plan_size(const tsr_ntt_plan *plan) { return plan ? 1ull << plan->p.L : 0; }

tsr_status tsr_ntt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len) {
    return run_ntt(plan, values, len, NTT_FWD, 0);
}
tsr_status tsr_intt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len) {
    return run_ntt(plan, values, len, NTT_INV, 0);
}
tsr_status tsr_coset_ntt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len, uint64_t shift) {
    return run_ntt(plan, values, len, NTT_COSET_FWD, shift);
}
tsr_status tsr_coset_intt(tsr_ntt_plan *plan, uint64_t *values, uint64_t len, uint64_t shift) {
    return run_ntt(plan, values, len, NTT_COSET_INV, shift);
}

void tsr_ntt_plan_free(tsr_ntt_plan *plan) {
    if (!plan) return;
    (void)hipSetDevice(plan->p.ctx->device);
    (void)tsr::stream_sync(plan->p.ctx->stream);
    delete plan;
}

}  // extern "C"
