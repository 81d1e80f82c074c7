// LogUp-GKR on gfx950: fraction trees built on the device, proven level by level with the
// device-resident round loop of sumcheck.hip.
//
// Semantics: crates/tessera-logup/src/prover.rs:201-469 (prove_logup: alpha, beta, leaves
// (1, alpha - (key + beta value)) padded with row 0, table leaves (-m_j, alpha - t_j), roots
// absorbed under logup/roots, the levels of prove_trees, values under logup/values),
// crates/tessera-logup/src/instance.rs:225-253 (multiplicities) and
// crates/tessera-sumcheck/src/trees.rs:199-345,761-930 (FractionTree: node x of level l is
// (nL dR + nR dL, dL dR) of entries 2x, 2x+1 of level l+1; per level: gamma, lambda, the
// rounds of eq(r_i, x)(N0 D1 + N1 D0 + lambda D0 D1), the evals, rho under tree/l/line, next
// claims at rho || r'_i).
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "commit.h"
#include "common.h"
#include "field.h"
#include "sumcheck.h"
#include "transcript.h"

namespace tsr {

using gl::fp2;

// alpha - (key + beta value), or alpha - key without values (compress, prover.rs:201-207)
__device__ __forceinline__ fp2 leaf_den(fp2 alpha, fp2 beta, uint64_t key, const uint64_t *vals,
                                        uint64_t x) {
    fp2 t = fp2{key, 0};
    if (vals) t = gl::add(gl::mul_base(beta, vals[x]), t);
    return gl::sub(alpha, t);
}

// Leaves of a tree of 2^h entries over `count` pairs, padded with `pad`. When mult is given
// (a table tree) the numerators -m_x are written (base) too.
__global__ void k_leaves(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                         uint64_t count, uint64_t leaves, fp2 alpha, fp2 beta, fp2 pad,
                         const uint64_t *__restrict__ mult, uint64_t *__restrict__ nout,
                         fp2 *__restrict__ dout) {
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= leaves) return;
    dout[x] = x < count ? leaf_den(alpha, beta, keys[x], vals, x) : pad;
    if (mult) nout[x] = gl::neg(mult[x]);
}

// Up to three tree levels per pass. A workgroup of 512 threads takes 256 * 2^L consecutive
// nodes of a level into LDS (loads and stores coalesced: thread t touches t, t + 512, ...) and
// writes the L levels above them, so a level is read once for every three built.
constexpr int LEVELS_PER_PASS = 3;
constexpr int LEVEL_THREADS = 512;
constexpr uint32_t level_block(int L) { return 256u << L; }  // inputs per workgroup
struct LevelsOutBase {
    uint64_t *v[3];  // level k above the inputs: v[k-1]
};
struct LevelsOut {
    fp2 *n[LEVELS_PER_PASS], *d[LEVELS_PER_PASS];  // level k above the inputs: n[k-1], d[k-1]
};

// The block's inputs are sn/sd[0, cnt) (input index base + i); writes levels 1..L above.
// UNIT: the inputs' numerators are all 1 (level 1 is D_l + D_r over D_l D_r; sn unused).
template <int L, bool UNIT>
__device__ __forceinline__ void block_levels(fp2 *sn, fp2 *sd, uint64_t base, uint32_t cnt,
                                             const LevelsOut &out) {
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int k = 1; k <= L; k++) {
        const uint32_t m = cnt >> k;  // nodes of level k in this block
        // level-1 nodes per thread (at least one; higher levels have fewer, j < m guards them)
        constexpr int PER = (int)(level_block(L) / 2 / LEVEL_THREADS) > 0 ? (int)(level_block(L) / 2 / LEVEL_THREADS) : 1;
        fp2 nn[PER], dd[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t j = t + LEVEL_THREADS * i;
            if (j < m) {
                const fp2 dl = sd[2 * j], dr = sd[2 * j + 1];
                nn[i] = (UNIT && k == 1) ? gl::add(dl, dr)
                                         : gl::add(gl::mul(sn[2 * j], dr), gl::mul(sn[2 * j + 1], dl));
                dd[i] = gl::mul(dl, dr);
            }
        }
        __syncthreads();  // every child read before any parent overwrites it
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t j = t + LEVEL_THREADS * i;
            if (j < m) {
                sn[j] = nn[i];
                sd[j] = dd[i];
                out.n[k - 1][(base >> k) + j] = nn[i];
                out.d[k - 1][(base >> k) + j] = dd[i];
            }
        }
        __syncthreads();
    }
}

// The leaves of a tree (as k_leaves) and the L levels above them: workgroup b owns leaves
// [b 256 2^L, (b + 1) 256 2^L) of the 2^h. Unit leaves keep no numerators: sn holds level 1 up.
template <int L, bool UNIT>
__global__ void __launch_bounds__(LEVEL_THREADS)
k_leaves_above(const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
               uint64_t count, uint64_t leaves, fp2 alpha, fp2 beta, fp2 pad,
               const uint64_t *__restrict__ mult, uint64_t *__restrict__ nout,
               fp2 *__restrict__ dout, LevelsOut out) {
    __shared__ fp2 sn[UNIT ? level_block(L) / 2 : level_block(L)], sd[level_block(L)];
    const uint64_t base = (uint64_t)blockIdx.x * level_block(L);
    const uint32_t cnt = (uint32_t)std::min<uint64_t>(leaves - base, level_block(L));
#pragma unroll
    for (int i = 0; i < (int)(level_block(L) / LEVEL_THREADS); i++) {
        const uint32_t j = threadIdx.x + LEVEL_THREADS * i;
        if (j >= cnt) break;
        const uint64_t y = base + j;
        const fp2 d = y < count ? leaf_den(alpha, beta, keys[y], vals, y) : pad;
        dout[y] = d;
        sd[j] = d;
        if (!UNIT) {
            const uint64_t m = gl::neg(mult[y]);
            nout[y] = m;
            sn[j] = fp2{m, 0};
        }
    }
    __syncthreads();
    block_levels<L, UNIT>(sn, sd, base, cnt, out);
}

// The L levels above a stored level of 2^below Fp2 numerators and denominators.
template <int L>
__global__ void __launch_bounds__(LEVEL_THREADS)
k_levels_above(const fp2 *__restrict__ nb, const fp2 *__restrict__ db, uint64_t inputs,
               LevelsOut out) {
    __shared__ fp2 sn[level_block(L)], sd[level_block(L)];
    const uint64_t base = (uint64_t)blockIdx.x * level_block(L);
    const uint32_t cnt = (uint32_t)std::min<uint64_t>(inputs - base, level_block(L));
#pragma unroll
    for (int i = 0; i < (int)(level_block(L) / LEVEL_THREADS); i++) {
        const uint32_t j = threadIdx.x + LEVEL_THREADS * i;
        if (j >= cnt) break;
        sn[j] = nb[base + j];
        sd[j] = db[base + j];
    }
    __syncthreads();
    block_levels<L, false>(sn, sd, base, cnt, out);
}


// eq(r, x) = prod_i (x_i ? r_i : 1 - r_i), x_1 the lowest bit (tessera-poly eq.rs:23).
__global__ void k_eq_table(const fp2 *__restrict__ point, uint32_t n, fp2 *__restrict__ out) {
    uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= (1ull << n)) return;
    fp2 acc{1, 0};
    for (uint32_t i = 0; i < n; i++) {
        fp2 r = point[i];
        acc = gl::mul(acc, ((x >> i) & 1) ? r : gl::sub(fp2{1, 0}, r));
    }
    out[x] = acc;
}

// eq(r, x) = eq(r_lo, x_lo) eq(r_hi, x_hi): out[x] = lo[x mod 2^a] hi[x >> a], one Fp2 product
// per entry from two tables of about 2^(n/2) entries each.
__global__ void k_eq_outer(const fp2 *__restrict__ lo, uint32_t a, const fp2 *__restrict__ hi,
                           uint64_t len, fp2 *__restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t x = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; x < len; x += stride)
        out[x] = gl::mul(lo[x & ((1ull << a) - 1)], hi[x >> a]);
}

// Partial sums of eq(point, x) col[x] over x < groups 2^a (col padded with `pad` from `count`
// on), one per block, with eq(point, x) = lo[x mod 2^a] hi[x >> a]: a block takes whole groups
// g of 2^a consecutive x, sums lo[l] col[g 2^a + l] (two base products per entry) and weights
// the group's sum by hi[g] once. No eq table of 2^n entries is built or read.
__global__ void __launch_bounds__(256)
k_eq_dot_split(const fp2 *__restrict__ lo, uint32_t a, const fp2 *__restrict__ hi,
               uint64_t groups, const uint64_t *__restrict__ col, uint64_t count, uint64_t pad,
               fp2 *__restrict__ partials) {
    fp2 acc{0, 0};
    const uint64_t glen = 1ull << a;
    for (uint64_t g = blockIdx.x; g < groups; g += gridDim.x) {
        fp2 sg{0, 0};
        for (uint64_t l = threadIdx.x; l < glen; l += 256) {
            const uint64_t x = (g << a) + l;
            sg = gl::add(sg, gl::mul_base(lo[l], x < count ? col[x] : pad));
        }
        acc = gl::add(acc, gl::mul(sg, hi[g]));
    }
    __shared__ fp2 red[256];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] = gl::add(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partials[blockIdx.x] = red[0];
}

// The sum of n partials (one block of 256 threads; the loads of a thread are independent).
__global__ void __launch_bounds__(256)
k_sum_partials(const fp2 *__restrict__ partials, uint32_t n, fp2 *__restrict__ out) {
    fp2 s{0, 0};
    for (uint32_t i = threadIdx.x; i < n; i += 256) s = gl::add(s, partials[i]);
    __shared__ fp2 red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] = gl::add(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// multiplicities (instance.rs:225-253): exact counts over 2^{h_t} rows.
// counts[0] += x (the padding lookups of row 0), on the device.
__global__ void k_add_first(uint64_t *__restrict__ counts, uint64_t x) {
    if (threadIdx.x == 0) counts[0] += x;
}

__global__ void k_count_rows(const uint64_t *__restrict__ rows, uint64_t n, uint64_t limit,
                             unsigned long long *__restrict__ counts, unsigned *__restrict__ bad) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint64_t r = rows[i];
    if (r < limit) atomicAdd(&counts[r], 1ull);
    else atomicOr(bad, 1u);
}

// The same counts for tables of at most 2^17 rows, without global contention on hot rows: the
// rows are cut into slices of 2^14; workgroup (chunk c, slice s) counts the lookups of chunk c
// that fall in slice s in an LDS histogram, then adds each nonzero bin to the global count once.
// Integer counting is exact in any order.
constexpr uint32_t HIST_SLICE_LOG = 14, HIST_THREADS = 512;
constexpr uint64_t HIST_CHUNK = 1ull << 16;
__global__ void __launch_bounds__(HIST_THREADS)
k_count_rows_lds(const uint64_t *__restrict__ rows, uint64_t n, uint64_t limit,
                 unsigned long long *__restrict__ counts, unsigned *__restrict__ bad) {
    __shared__ uint32_t hist[1u << HIST_SLICE_LOG];
    const uint64_t slice_lo = (uint64_t)blockIdx.y << HIST_SLICE_LOG;
    for (uint32_t i = threadIdx.x; i < (1u << HIST_SLICE_LOG); i += HIST_THREADS) hist[i] = 0;
    __syncthreads();
    const uint64_t c0 = (uint64_t)blockIdx.x * HIST_CHUNK;
    const uint64_t c1 = c0 + HIST_CHUNK < n ? c0 + HIST_CHUNK : n;
    for (uint64_t i = c0 + threadIdx.x; i < c1; i += HIST_THREADS) {
        const uint64_t r = rows[i];
        if (r >= limit) {
            if (blockIdx.y == 0) atomicOr(bad, 1u);
        } else if (r - slice_lo < (1ull << HIST_SLICE_LOG) && r >= slice_lo) {
            atomicAdd(&hist[r - slice_lo], 1u);
        }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < (1u << HIST_SLICE_LOG); i += HIST_THREADS) {
        const uint32_t v = hist[i];
        if (v && slice_lo + i < limit) atomicAdd(&counts[slice_lo + i], (unsigned long long)v);
    }
}

static unsigned nblocks(uint64_t n, unsigned per) {
    return (unsigned)std::max<uint64_t>(1, (n + per - 1) / per);
}

static uint32_t ceil_log2(uint64_t x) {
    uint32_t h = 0;
    while ((1ull << h) < x) h++;
    return h;
}

// A product forest's level above (forest.rs:306-340): V[t][z] = V[t+1][2z] V[t+1][2z+1], base
// field throughout since the leaves are base.
// L levels per pass as the fraction trees': a 512-thread workgroup stages 256 2^L consecutive
// values in LDS and writes their ancestors on each level (out[k-1]: level k above the input).
template <int L>
__global__ void __launch_bounds__(LEVEL_THREADS)
k_products_above(const uint64_t *__restrict__ below, uint64_t inputs, LevelsOutBase out) {
    __shared__ uint64_t sv[level_block(L)];
    const uint64_t base = (uint64_t)blockIdx.x * level_block(L);
    const uint32_t cnt = (uint32_t)std::min<uint64_t>(inputs - base, level_block(L));
    const uint32_t t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < (int)(level_block(L) / LEVEL_THREADS); i++) {
        const uint32_t j = t + LEVEL_THREADS * i;
        if (j < cnt) sv[j] = below[base + j];
    }
    __syncthreads();
#pragma unroll
    for (int k = 1; k <= L; k++) {
        const uint32_t m = cnt >> k;
        constexpr int PER = (int)(level_block(L) / 2 / LEVEL_THREADS) > 0 ? (int)(level_block(L) / 2 / LEVEL_THREADS) : 1;
        uint64_t v[PER];
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t j = t + LEVEL_THREADS * i;
            if (j < m) v[i] = gl::mul(sv[2 * j], sv[2 * j + 1]);
        }
        __syncthreads();
#pragma unroll
        for (int i = 0; i < PER; i++) {
            const uint32_t j = t + LEVEL_THREADS * i;
            if (j < m) {
                sv[j] = v[i];
                out.v[k - 1][(base >> k) + j] = v[i];
            }
        }
        __syncthreads();
    }
}

// A tree of the level loop on the device.
//   fraction tree: level t (0..h) holds N[t], D[t] of 2^t entries; the leaf level's numerators
//     are unit (lookup tree) or base (table tree, nbase).
//   product forest: 2^s stacked trees of height h; level t holds V[t] of 2^(s+t) base values,
//     V[h] the caller's leaves.
struct DevTree {
    uint32_t h = 0, join = 0, s = 0;
    bool forest = false;
    bool unit_leaves = true;
    std::vector<DevBuf<fp2>> N, D;
    DevBuf<uint64_t> nbase;
    std::vector<DevBuf<uint64_t>> V;  // levels 0..h-1 of a forest
    const uint64_t *leaves = nullptr;
    // current claim: point and (num, den), or the product value in num; a forest has none
    // until the level it joins
    std::vector<fp2> point;
    fp2 num{0, 0}, den{0, 0};
    bool has_claim = true;
    const uint64_t *level_values(uint32_t t) const { return t == h ? leaves : V[t].ptr; }
};

static Status build_forest(Ctx *ctx, DevTree &t, const uint64_t *leaves, uint64_t len,
                           uint32_t height) {
    hipStream_t s = ctx->stream;
    uint32_t vars = 0;
    while ((1ull << vars) < len) vars++;
    t.forest = true;
    t.has_claim = false;
    t.h = height;
    t.s = vars - height;
    t.leaves = leaves;
    t.V = std::vector<DevBuf<uint64_t>>(height);
    Status st;
    for (int l = (int)height - 1; l >= 0; l--)
        if ((st = t.V[l].alloc(ctx, 1ull << (t.s + l)))) return st;
    KernelTimer timer(ctx, "k_products_above", s);
    for (uint32_t below = height; below > 0;) {
        const uint32_t L = std::min<uint32_t>(below, LEVELS_PER_PASS);
        const uint64_t inputs = 1ull << (t.s + below);
        LevelsOutBase o{};
        for (uint32_t k = 1; k <= L; k++) o.v[k - 1] = t.V[below - k].ptr;
        const dim3 grid(nblocks(inputs, level_block(L)));
        if (L == 1)
            hipLaunchKernelGGL(k_products_above<1>, grid, dim3(LEVEL_THREADS), 0, s, t.level_values(below), inputs, o);
        else if (L == 2)
            hipLaunchKernelGGL(k_products_above<2>, grid, dim3(LEVEL_THREADS), 0, s, t.level_values(below), inputs, o);
        else
            hipLaunchKernelGGL(k_products_above<3>, grid, dim3(LEVEL_THREADS), 0, s, t.level_values(below), inputs, o);
        below -= L;
    }
    return check_launch("product forest");
}

static Status build_tree(Ctx *ctx, DevTree &t, uint32_t h, const uint64_t *keys,
                         const uint64_t *vals, uint64_t count, const uint64_t *mult, fp2 alpha,
                         fp2 beta, fp2 pad) {
    hipStream_t s = ctx->stream;
    t.h = h;
    t.unit_leaves = mult == nullptr;
    t.N = std::vector<DevBuf<fp2>>(h + 1);
    t.D = std::vector<DevBuf<fp2>>(h + 1);
    const uint64_t leaves = 1ull << h;
    Status st;
    if ((st = t.D[h].alloc(ctx, leaves))) return st;
    if (mult && (st = t.nbase.alloc(ctx, leaves))) return st;
    for (int l = (int)h - 1; l >= 0; l--)
        if ((st = t.N[l].alloc(ctx, 1ull << l)) || (st = t.D[l].alloc(ctx, 1ull << l))) return st;
    if (h == 0) {
        KernelTimer timer(ctx, "k_leaves", s);
        hipLaunchKernelGGL(k_leaves, dim3(1), dim3(256), 0, s, keys, vals, count, leaves, alpha,
                           beta, pad, mult, t.nbase.ptr, t.D[0].ptr);
        return check_launch("fraction tree");
    }
    // the leaves and up to three levels above them, then three levels per pass
    auto outs = [&](uint32_t below, uint32_t L) {
        LevelsOut o{};
        for (uint32_t k = 1; k <= L; k++) {
            o.n[k - 1] = t.N[below - k].ptr;
            o.d[k - 1] = t.D[below - k].ptr;
        }
        return o;
    };
    {
        KernelTimer timer(ctx, "k_leaves_above", s);
        const uint32_t L = std::min<uint32_t>(h, LEVELS_PER_PASS);
        const LevelsOut o = outs(h, L);
        const dim3 grid(nblocks(leaves, level_block(L)));
#define LEAVES(LL)                                                                               \
    if (L == LL) {                                                                               \
        if (mult)                                                                                \
            hipLaunchKernelGGL((k_leaves_above<LL, false>), grid, dim3(LEVEL_THREADS), 0, s,      \
                               keys, vals, count, leaves, alpha, beta, pad, mult, t.nbase.ptr,   \
                               t.D[h].ptr, o);                                                   \
        else                                                                                     \
            hipLaunchKernelGGL((k_leaves_above<LL, true>), grid, dim3(LEVEL_THREADS), 0, s, keys, \
                               vals, count, leaves, alpha, beta, pad, mult, nullptr, t.D[h].ptr, \
                               o);                                                               \
    }
        LEAVES(1) LEAVES(2) LEAVES(3)
#undef LEAVES
    }
    KernelTimer timer(ctx, "k_levels_above", s);
    for (uint32_t below = h - std::min<uint32_t>(h, LEVELS_PER_PASS); below > 0;) {
        const uint32_t L = std::min<uint32_t>(below, LEVELS_PER_PASS);
        const uint64_t inputs = 1ull << below;
        const LevelsOut o = outs(below, L);
        const dim3 grid(nblocks(inputs, level_block(L)));
        if (L == 1)
            hipLaunchKernelGGL(k_levels_above<1>, grid, dim3(LEVEL_THREADS), 0, s, t.N[below].ptr,
                               t.D[below].ptr, inputs, o);
        else if (L == 2)
            hipLaunchKernelGGL(k_levels_above<2>, grid, dim3(LEVEL_THREADS), 0, s, t.N[below].ptr,
                               t.D[below].ptr, inputs, o);
        else
            hipLaunchKernelGGL(k_levels_above<3>, grid, dim3(LEVEL_THREADS), 0, s, t.N[below].ptr,
                               t.D[below].ptr, inputs, o);
        below -= L;
    }
    return check_launch("fraction tree");
}

static Status read_root(Ctx *ctx, DevTree &t) {
    hipStream_t s = ctx->stream;
    TSR_HIP(hipMemcpyAsync(&t.den, t.D[0].ptr, 16, hipMemcpyDeviceToHost, s));
    if (t.h == 0) {
        if (t.unit_leaves) {
            t.num = fp2{1, 0};
        } else {
            uint64_t v = 0;
            TSR_HIP(hipMemcpyAsync(&v, t.nbase.ptr, 8, hipMemcpyDeviceToHost, s));
            TSR_HIP(tsr::stream_sync(s));
            t.num = fp2{v, 0};
        }
    } else {
        TSR_HIP(hipMemcpyAsync(&t.num, t.N[0].ptr, 16, hipMemcpyDeviceToHost, s));
    }
    TSR_HIP(tsr::stream_sync(s));
    return Status::ok();
}

struct Log {
    uint64_t *buf;
    uint64_t cap, len = 0;
    void put(uint64_t v) {
        if (buf && len < cap) buf[len] = v;
        len++;
    }
    void f2(const fp2 &v) {
        put(v.c0);
        put(v.c1);
    }
};

// Sum_x eq(point, x) col[x] over x < 2^n (col base, padded with `pad` from `count` on), read back
// to the host; point: n Fp2 on the device. The eq half tables over point[0, a) and point[a, n)
// take 2^a + 2^(n-a) entries.
static Status eq_dot(Ctx *ctx, const fp2 *point, uint32_t n, const uint64_t *col, uint64_t count,
                     uint64_t pad, fp2 *part, fp2 *val, fp2 &out) {
    hipStream_t s = ctx->stream;
    const uint32_t a = n <= 12 ? n : n / 2;
    const uint64_t groups = 1ull << (n - a);
    DevBuf<fp2> lo, hi;
    Status st;
    if ((st = lo.alloc(ctx, 1ull << a)) || (st = hi.alloc(ctx, groups))) return st;
    hipLaunchKernelGGL(k_eq_table, dim3(nblocks(1ull << a, 256)), dim3(256), 0, s, point, a, lo.ptr);
    hipLaunchKernelGGL(k_eq_table, dim3(nblocks(groups, 256)), dim3(256), 0, s, point + a, n - a, hi.ptr);
    const unsigned nb = (unsigned)std::min<uint64_t>(1024, groups);
    hipLaunchKernelGGL(k_eq_dot_split, dim3(nb), dim3(256), 0, s, lo.ptr, a, hi.ptr, groups, col,
                       count, pad, part);
    hipLaunchKernelGGL(k_sum_partials, dim3(1), dim3(256), 0, s, part, nb, val);
    if ((st = check_launch("eq dot"))) return st;
    TSR_HIP(hipMemcpyAsync(&out, val, 16, hipMemcpyDeviceToHost, s));
    TSR_HIP(tsr::stream_sync(s));
    return Status::ok();
}

// The level loop of prove_trees_in (trees.rs:779-930) over the scheduled trees: each joins at
// tallest - height. Per level: the root challenges tree/l/root/{c} of the forests joining here
// (each takes the last s of them as its point, its claim Sum_z eq(point, z) V[0][z]), gamma,
// lambda when a fraction tree is active, the batched rounds of eq (N0 D1 + N1 D0 + lambda D0 D1)
// per fraction tree and eq V0 V1 per forest (degree 3, padding rounds from the claims), the
// evals, rho under tree/l/line and the next claims at rho || own point. The log record per
// level: s_max, the root challenges, has_lambda, n_max, gamma, lambda (0 without fractions),
// the rounds (n_max x 3), the challenges, the count of evals and the evals, rho, the state,
// the counter.
static Status prove_levels(Ctx *ctx, Transcript &tr, std::vector<DevTree> &trees,
                           const std::vector<uint32_t> &schedule, Log &log) {
    hipStream_t s = ctx->stream;
    uint32_t tallest = 0;
    for (uint32_t id : schedule) tallest = std::max(tallest, trees[id].h);
    for (uint32_t id : schedule) trees[id].join = tallest - trees[id].h;

    DevBuf<fp2> d_point, d_part, d_val;
    Status st;
    if ((st = d_point.alloc(ctx, 64)) || (st = d_part.alloc(ctx, 1024)) || (st = d_val.alloc(ctx, 1)))
        return st;
    for (uint32_t level = 0; level < tallest; level++) {
        const std::string prefix = "tree/" + std::to_string(level);
        uint32_t s_max = 0;
        for (uint32_t id : schedule)
            if (trees[id].forest && trees[id].join == level) s_max = std::max(s_max, trees[id].s);
        std::vector<fp2> rootc(s_max);
        for (uint32_t c = 0; c < s_max; c++)
            rootc[c] = tr.challenge_fp2((prefix + "/root/" + std::to_string(c)).c_str(), nullptr);
        for (uint32_t id : schedule) {
            DevTree &t = trees[id];
            if (t.forest && t.join == level)
                t.point.assign(rootc.begin() + (s_max - t.s), rootc.end());
        }
        std::vector<uint32_t> act;
        bool fractions = false;
        for (uint32_t id : schedule)
            if (trees[id].join <= level) {
                act.push_back(id);
                fractions |= !trees[id].forest;
            }
        const fp2 gamma = tr.challenge_fp2((prefix + "/gamma").c_str(), nullptr);
        const fp2 lambda = fractions ? tr.challenge_fp2((prefix + "/lambda").c_str(), nullptr) : fp2{0, 0};
        std::vector<GenInst> insts(act.size());
        for (size_t i = 0; i < act.size(); i++) {
            DevTree &t = trees[act[i]];
            const uint32_t below = level - t.join + 1, n = (uint32_t)t.point.size();
            if (n > 64) return Status::fail(TSR_DECLINED, "a claim point of more than 64 variables");
            GenInst &g = insts[i];
            g.vars = n;
            g.claim_given = true;
            g.eq_factor = true;
            g.eq_point = t.point;
            if (t.forest) {
                if (!t.has_claim) {
                    if (n) {
                        TSR_HIP(hipMemcpyAsync(d_point.ptr, t.point.data(), 16 * n, hipMemcpyHostToDevice, s));
                    }
                    // the joining claim: V0 V1 at z is the root V[0][z] (trees.rs:638-640); eq_dot
                    // ends synchronized, so d_point is free for the next point
                    if ((st = eq_dot(ctx, d_point.ptr, n, t.V[0].ptr, 1ull << n, 0, d_part.ptr,
                                     d_val.ptr, t.num)))
                        return st;
                    t.has_claim = true;
                }
                g.rel = REL_PRODUCT;
                g.claim = t.num;
                GenTable v0, v1;
                v0.src = v1.src = t.level_values(below);
                v0.base = v1.base = true;
                v0.shift = v1.shift = 1;
                v1.off = 1;
                g.tabs = {v0, v1};
            } else {
                g.rel = REL_FRACTION;
                g.claim = gl::add(t.num, gl::mul(lambda, t.den));
                GenTable n0, n1, d0, d1;
                const bool leaf = below == t.h;
                if (leaf && t.unit_leaves) {
                    n0.src = n1.src = nullptr;
                } else if (leaf) {
                    n0.src = n1.src = t.nbase.ptr;
                    n0.base = n1.base = true;
                } else {
                    n0.src = n1.src = t.N[below].ptr;
                }
                d0.src = d1.src = t.D[below].ptr;
                n0.shift = n1.shift = d0.shift = d1.shift = 1;
                n1.off = d1.off = 1;
                g.tabs = {n0, n1, d0, d1};
                // the level this instance proves: its first round's sums at z = 0, 1
                if (n0.src && t.N[below - 1].ptr && t.D[below - 1].ptr) {
                    g.parent_n = t.N[below - 1].ptr;
                    g.parent_d = t.D[below - 1].ptr;
                }
            }
        }
        SumcheckOut out;
        if ((st = run_batch(ctx, tr, prefix, 3, insts, gamma, lambda, out))) return st;
        const fp2 rho = tr.challenge_fp2((prefix + "/line").c_str(), nullptr);
        const uint32_t n_max = (uint32_t)out.challenges.size();
        log.put(s_max);
        for (auto &v : rootc) log.f2(v);
        log.put(fractions ? 1 : 0);
        log.put(n_max);
        log.f2(gamma);
        log.f2(lambda);
        for (auto &v : out.rounds) log.f2(v);
        for (auto &v : out.challenges) log.f2(v);
        log.put(out.finals.size());
        for (auto &v : out.finals) log.f2(v);
        log.f2(rho);
        for (int b = 0; b < 4; b++) {
            uint64_t w = 0;
            for (int j = 0; j < 8; j++) w |= (uint64_t)tr.state[8 * b + j] << (8 * j);
            log.put(w);
        }
        log.put(tr.counter);
        size_t at = 0;
        for (size_t i = 0; i < act.size(); i++) {
            DevTree &t = trees[act[i]];
            const uint32_t n = (uint32_t)t.point.size();
            std::vector<fp2> np;
            np.push_back(rho);
            for (uint32_t k = 0; k < n; k++) np.push_back(out.challenges[n_max - n + k]);
            t.point = np;
            const fp2 *f = &out.finals[at];
            const uint32_t below = level - t.join + 1;
            if (t.forest) {
                t.num = gl::add(f[0], gl::mul(rho, gl::sub(f[1], f[0])));
                at += 2;
                // levels at and above this one are no longer needed (trees.rs drop_level)
                if (below == 1) t.V[0].release();
                if (below < t.h) t.V[below].release();
            } else {
                t.num = gl::add(f[0], gl::mul(rho, gl::sub(f[1], f[0])));
                t.den = gl::add(f[2], gl::mul(rho, gl::sub(f[3], f[2])));
                at += 4;
                // the level above this one is done with; this one stays as the next level's
                // parents, unless the next is a leaf level with unit numerators (no parents)
                if (below >= 2) {
                    t.N[below - 1].release();
                    t.D[below - 1].release();
                }
                if (below + 1 == t.h && t.unit_leaves) {
                    t.N[below].release();
                    t.D[below].release();
                }
            }
        }
    }
    return Status::ok();
}

// Uploads a host column into `up` (device columns pass through); nullptr stays nullptr.
struct Uploads {
    Ctx *ctx;
    std::vector<DevBuf<uint64_t>> bufs;
    Status operator()(const uint64_t *p, uint64_t n, const uint64_t **out) {
        if (!p || is_device_ptr(p)) {
            *out = p;
            return Status::ok();
        }
        bufs.emplace_back();
        Status st = bufs.back().alloc(ctx, n);
        if (st) return st;
        TSR_HIP(hipMemcpyAsync(bufs.back().ptr, p, 8 * n, hipMemcpyHostToDevice, ctx->stream));
        *out = bufs.back().ptr;
        return Status::ok();
    }
};

// prove_trees (trees.rs:761-776) over product forests alone, in order: each forest's leaf claim
// value in forest_claims (its point is the last level's rho || challenges).
Status prove_trees(Ctx *ctx, Transcript &tr, const tsr_forest *forests, uint32_t nforests,
                   std::vector<fp2> &forest_claims, Log &log) {
    Uploads dev{ctx, {}};
    dev.bufs.reserve(nforests + 1);
    std::vector<DevTree> trees(nforests);
    std::vector<uint32_t> sched(nforests);
    Status st;
    for (uint32_t f = 0; f < nforests; f++) {
        const uint64_t *lv;
        if ((st = dev(forests[f].leaves, forests[f].len, &lv)) ||
            (st = build_forest(ctx, trees[f], lv, forests[f].len, forests[f].height)))
            return st;
        sched[f] = f;
    }
    if ((st = prove_levels(ctx, tr, trees, sched, log))) return st;
    forest_claims.clear();
    for (auto &t : trees) forest_claims.push_back(t.num);
    return Status::ok();
}

Status prove_logup(Ctx *ctx, Transcript &tr, const tsr_logup_instance *inst, uint32_t count,
                   const tsr_forest *forests, uint32_t nforests,
                   const std::vector<uint32_t> &schedule, fp2 &alpha, fp2 &beta,
                   std::vector<fp2> &roots, std::vector<fp2> &values, std::vector<fp2> &claims,
                   std::vector<fp2> &forest_claims, Log &log) {
    hipStream_t s = ctx->stream;
    alpha = tr.challenge_fp2("logup/alpha", nullptr);
    beta = tr.challenge_fp2("logup/beta", nullptr);

    // device copies of the columns
    Uploads up{ctx, {}};
    up.bufs.reserve(8 * count + nforests + 1);
    std::vector<DevTree> trees(2 * count + nforests);
    std::vector<const uint64_t *> lvals(count), tvals(count);
    roots.clear();
    Status st;
    for (uint32_t k = 0; k < count; k++) {
        const tsr_logup_instance &in = inst[k];
        const uint32_t ha = ceil_log2(in.lookups), ht = ceil_log2(in.rows);
        const uint64_t *keys, *vals, *lk, *lv, *m;
        if ((st = up(in.keys, in.rows, &keys)) || (st = up(in.values, in.rows, &vals)) ||
            (st = up(in.lookup_keys, in.lookups, &lk)) ||
            (st = up(in.lookup_values, in.lookups, &lv)) ||
            (st = up(in.multiplicities, 1ull << ht, &m)))
            return st;
        lvals[k] = lv;
        tvals[k] = vals;
        // alpha - t_0, row 0's denominator (prover.rs:247-256), from the host's copy of row 0
        uint64_t key0 = 0, val0 = 0;
        TSR_HIP(hipMemcpyAsync(&key0, keys, 8, hipMemcpyDeviceToHost, s));
        if (vals) TSR_HIP(hipMemcpyAsync(&val0, vals, 8, hipMemcpyDeviceToHost, s));
        TSR_HIP(tsr::stream_sync(s));
        fp2 t0 = fp2{key0, 0};
        if (vals) t0 = gl::add(gl::mul_base(beta, val0), t0);
        const fp2 pad = gl::sub(alpha, t0);
        if ((st = build_tree(ctx, trees[2 * k], ha, lk, lv, in.lookups, nullptr, alpha, beta, pad)) ||
            (st = build_tree(ctx, trees[2 * k + 1], ht, keys, vals, in.rows, m, alpha, beta, pad)))
            return st;
        if ((st = read_root(ctx, trees[2 * k])) || (st = read_root(ctx, trees[2 * k + 1]))) return st;
        roots.push_back(trees[2 * k].num);
        roots.push_back(trees[2 * k].den);
        roots.push_back(trees[2 * k + 1].num);
        roots.push_back(trees[2 * k + 1].den);
    }
    tr.absorb_fp2s("logup/roots", (const uint64_t *)roots.data(), roots.size());
    for (uint32_t f = 0; f < nforests; f++) {
        const uint64_t *lv;
        if ((st = up(forests[f].leaves, forests[f].len, &lv)) ||
            (st = build_forest(ctx, trees[2 * count + f], lv, forests[f].len, forests[f].height)))
            return st;
    }
    if ((st = prove_levels(ctx, tr, trees, schedule, log))) return st;

    // values: the padded lookup value column at rho_a (prover.rs:301-318)
    values.clear();
    claims.clear();
    forest_claims.clear();
    for (uint32_t f = 0; f < nforests; f++) forest_claims.push_back(trees[2 * count + f].num);
    DevBuf<fp2> d_part, d_val, d_point;
    if ((st = d_part.alloc(ctx, 1024)) || (st = d_val.alloc(ctx, 1)) || (st = d_point.alloc(ctx, 64)))
        return st;
    for (uint32_t k = 0; k < count; k++) {
        DevTree &la = trees[2 * k], &ta = trees[2 * k + 1];
        if (inst[k].values) {
            const uint32_t n = la.h;
            if (n) TSR_HIP(hipMemcpyAsync(d_point.ptr, la.point.data(), 16 * n, hipMemcpyHostToDevice, s));
            uint64_t pad = 0;  // row 0's value fills the padding slots
            TSR_HIP(hipMemcpyAsync(&pad, tvals[k], 8, hipMemcpyDeviceToHost, s));
            TSR_HIP(tsr::stream_sync(s));
            fp2 v{0, 0};
            if ((st = eq_dot(ctx, d_point.ptr, n, lvals[k], inst[k].lookups, pad, d_part.ptr, d_val.ptr, v)))
                return st;
            values.push_back(v);
        }
        claims.push_back(la.num);
        claims.push_back(la.den);
        claims.push_back(ta.num);
        claims.push_back(ta.den);
    }
    tr.absorb_fp2s("logup/values", (const uint64_t *)values.data(), values.size());
    return Status::ok();
}

}  // namespace tsr

using namespace tsr;

extern "C" {

// Forest shapes (forest.rs:269-286), forest by forest: LeafCount, then Height.
static Status check_forests(const tsr_forest *forests, uint32_t nforests) {
    for (uint32_t f = 0; f < nforests; f++) {
        const uint64_t len = forests[f].len;
        if (len == 0 || (len & (len - 1)))
            return Status::fail(TSR_BAD_SHAPE, "LeafCount: forest " + std::to_string(f));
        if (forests[f].height == 0 || (1ull << forests[f].height) > len)
            return Status::fail(TSR_BAD_SHAPE, "Height: forest " + std::to_string(f));
        if (len > (1ull << 32)) return Status::fail(TSR_DECLINED, "a forest above 2^32 leaves");
        if (!forests[f].leaves) return Status::fail(TSR_BAD_ARG, "null forest leaves");
    }
    return Status::ok();
}

static void put_f2s(uint64_t *out, const std::vector<fp2> &v) {
    if (out && !v.empty()) memcpy(out, v.data(), 16 * v.size());
}

tsr_status tsr_logup_prove(tsr_ctx *ctx_, tsr_transcript *t, const tsr_logup_instance *inst,
                           uint32_t count, const tsr_forest *forests, uint32_t nforests,
                           const uint32_t *schedule, uint64_t alpha_out[2], uint64_t beta_out[2],
                           uint64_t *roots_out, uint64_t *values_out, uint64_t *n_values,
                           uint64_t *claims_out, uint64_t *forest_claims_out, uint64_t *log,
                           uint64_t log_cap, uint64_t *log_len) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !t || (count && !inst) || (nforests && !forests))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    // LogUpError order (prover.rs:361-375): per instance the columns, the forests' shapes, then
    // the schedule.
    for (uint32_t k = 0; k < count; k++) {
        const tsr_logup_instance &in = inst[k];
        if (in.rows == 0 || !in.keys)
            return Status::fail(TSR_BAD_SHAPE, "NoRows: instance " + std::to_string(k)).code;
        if (in.lookups == 0 || !in.lookup_keys)
            return Status::fail(TSR_BAD_SHAPE, "NoLookups: instance " + std::to_string(k)).code;
        if ((in.values == nullptr) != (in.lookup_values == nullptr))
            return Status::fail(TSR_BAD_SHAPE, "ValuesMismatch: instance " + std::to_string(k)).code;
        if (!in.multiplicities)
            return Status::fail(TSR_BAD_SHAPE, "Multiplicities: instance " + std::to_string(k)).code;
        if (ceil_log2(in.lookups) > 32 || ceil_log2(in.rows) > 32)
            return Status::fail(TSR_DECLINED, "a tree above 2^32 leaves").code;
    }
    Status st = check_forests(forests, nforests);
    if (st) return st.code;
    const uint32_t ntrees = 2 * count + nforests;
    std::vector<uint32_t> sched(ntrees);
    if (schedule) {
        std::vector<int> seen(ntrees, 0);
        for (uint32_t i = 0; i < ntrees; i++) {
            if (schedule[i] >= ntrees || seen[schedule[i]]++)
                return Status::fail(TSR_BAD_SHAPE, "ScheduleTree: position " + std::to_string(i)).code;
            sched[i] = schedule[i];
        }
    } else {
        for (uint32_t i = 0; i < ntrees; i++) sched[i] = i;  // TreeId::paired, then the forests
    }
    if (forest_claims_out == nullptr && nforests) return Status::fail(TSR_BAD_ARG, "null forest claims").code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    fp2 alpha, beta;
    std::vector<fp2> roots, values, claims, fclaims;
    Log lg{log, log ? log_cap : 0};
    // on a copy of the transcript, kept only when the whole proof succeeded
    Transcript work = t->t;
    st = prove_logup(ctx, work, inst, count, forests, nforests, sched, alpha, beta, roots, values,
                     claims, fclaims, lg);
    if (st) return st.code;
    t->t = work;
    if (alpha_out) {
        alpha_out[0] = alpha.c0;
        alpha_out[1] = alpha.c1;
    }
    if (beta_out) {
        beta_out[0] = beta.c0;
        beta_out[1] = beta.c1;
    }
    put_f2s(roots_out, roots);
    put_f2s(values_out, values);
    if (n_values) *n_values = values.size();
    put_f2s(claims_out, claims);
    put_f2s(forest_claims_out, fclaims);
    if (log_len) *log_len = lg.len;
    if (log && lg.len > log_cap) return Status::fail(TSR_BAD_SHAPE, "log buffer too small").code;
    return TSR_OK;
}

tsr_status tsr_prove_trees(tsr_ctx *ctx_, tsr_transcript *t, const tsr_forest *forests,
                           uint32_t nforests, uint64_t *forest_claims_out, uint64_t *log,
                           uint64_t log_cap, uint64_t *log_len) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !t || (nforests && (!forests || !forest_claims_out)))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    Status st = check_forests(forests, nforests);
    if (st) return st.code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    std::vector<fp2> fclaims;
    Log lg{log, log ? log_cap : 0};
    Transcript work = t->t;  // kept only when the whole proof succeeded
    if ((st = prove_trees(ctx, work, forests, nforests, fclaims, lg))) return st.code;
    t->t = work;
    put_f2s(forest_claims_out, fclaims);
    if (log_len) *log_len = lg.len;
    if (log && lg.len > log_cap) return Status::fail(TSR_BAD_SHAPE, "log buffer too small").code;
    return TSR_OK;
}

// The lookup rows and the counts are secret-class (instance.rs:217-219: SecretBuf, cleared on
// drop): the device copies are cleared before their blocks go back to the context's pool.
struct ClearOnExit {
    uint64_t *p;
    uint64_t n;
    hipStream_t s;
    ~ClearOnExit() {
        if (p && n) {
            zero_words(p, n, s);
            (void)tsr::stream_sync(s);
        }
    }
};

tsr_status tsr_logup_multiplicities(tsr_ctx *ctx_, uint64_t rows, const uint64_t *lookup_rows,
                                    uint64_t lookups, uint64_t *out) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || (lookups && !lookup_rows) || (rows && !out))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    // A table of no row has no leaf, so nothing is written; any lookup then names no row
    // (instance.rs:228-241).
    if (rows == 0) {
        if (lookups)
            return Status::fail(TSR_BAD_SHAPE, "LookupRow: a lookup names no row of the table").code;
        return TSR_OK;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    hipStream_t s = ctx->stream;
    const uint64_t leaves = 1ull << ceil_log2(rows);
    DevBuf<uint64_t> counts, in;
    DevBuf<uint32_t> bad;
    Status st;
    if ((st = counts.alloc(ctx, leaves))) return st.code;
    ClearOnExit clear_counts{counts.ptr, leaves, s};
    const uint64_t *d = lookup_rows;
    if ((st = bad.alloc(ctx, 1))) return st.code;
    if (lookups && !is_device_ptr(lookup_rows)) {
        if ((st = in.alloc(ctx, lookups))) return st.code;
        e = hipMemcpyAsync(in.ptr, lookup_rows, 8 * lookups, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy").code;
        d = in.ptr;
    }
    ClearOnExit clear_in{in.ptr, in.ptr ? lookups : 0, s};
    zero_words(counts.ptr, leaves, s);
    e = hipMemsetAsync(bad.ptr, 0, 4, s);
    if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync").code;
    const uint64_t slices = (rows + (1ull << HIST_SLICE_LOG) - 1) >> HIST_SLICE_LOG;
    const uint64_t chunks = (lookups + HIST_CHUNK - 1) / HIST_CHUNK;
    if (lookups && slices <= 8 && chunks <= 65535)
        hipLaunchKernelGGL(k_count_rows_lds, dim3((unsigned)chunks, (unsigned)slices), dim3(HIST_THREADS), 0, s,
                           d, lookups, rows, (unsigned long long *)counts.ptr, bad.ptr);
    else if (lookups)
        hipLaunchKernelGGL(k_count_rows, dim3(nblocks(lookups, 256)), dim3(256), 0, s, d, lookups,
                           rows, (unsigned long long *)counts.ptr, bad.ptr);
    if ((st = check_launch("multiplicities"))) return st.code;
    uint32_t hbad = 0;
    e = hipMemcpyAsync(&hbad, bad.ptr, 4, hipMemcpyDeviceToHost, s);
    if (e != hipSuccess) return hip_fail(e, "copy").code;
    // the row check comes back first, so that a bad lookup leaves `out` untouched
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    if (hbad) return Status::fail(TSR_BAD_SHAPE, "LookupRow: a lookup names no row of the table").code;
    if (lookups)
        hipLaunchKernelGGL(k_add_first, dim3(1), dim3(64), 0, s, counts.ptr,
                           (1ull << ceil_log2(lookups)) - lookups);
    if ((st = check_launch("multiplicities"))) return st.code;
    // device or host output alike: one copy of the counts
    e = hipMemcpyAsync(out, counts.ptr, 8 * leaves, hipMemcpyDefault, s);
    if (e != hipSuccess) return hip_fail(e, "copy").code;
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    return TSR_OK;
}

}  // extern "C"

namespace tsr {
// eq(point, .) over n variables. Up to 2^12 entries each entry is its own product; above, the
// table is the outer product of the tables over the low n/2 and high n - n/2 variables.
Status eq_table(Ctx *ctx, const gl::fp2 *point, uint32_t n, gl::fp2 *out, hipStream_t s) {
    KernelTimer timer(ctx, "k_eq_table", s);
    if (n <= 12) {
        hipLaunchKernelGGL(k_eq_table, dim3(nblocks(1ull << n, 256)), dim3(256), 0, s, point, n, out);
        return check_launch("eq table");
    }
    const uint32_t a = n / 2;
    DevBuf<gl::fp2> lo, hi;
    Status st;
    if ((st = lo.alloc(ctx, 1ull << a)) || (st = hi.alloc(ctx, 1ull << (n - a)))) return st;
    hipLaunchKernelGGL(k_eq_table, dim3(nblocks(1ull << a, 256)), dim3(256), 0, s, point, a, lo.ptr);
    hipLaunchKernelGGL(k_eq_table, dim3(nblocks(1ull << (n - a), 256)), dim3(256), 0, s, point + a,
                       n - a, hi.ptr);
    const uint64_t len = 1ull << n;
    hipLaunchKernelGGL(k_eq_outer, dim3((unsigned)std::min<uint64_t>(nblocks(len, 256), 1u << 20)), dim3(256),
                       0, s, lo.ptr, a, hi.ptr, len, out);
    if ((st = check_launch("eq table"))) return st;
    // lo and hi go back to the context's pool, which other streams may draw from
    TSR_HIP(tsr::stream_sync(s));
    return Status::ok();
}
}  // namespace tsr

extern "C" tsr_status tsr_eq_table(tsr_ctx *ctx_, const uint64_t *point, uint32_t n, uint64_t *out) {
    using namespace tsr;
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !out || (n && !point)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (n > 30) return Status::fail(TSR_BAD_SHAPE, "eq table of more than 30 variables").code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    hipStream_t s = ctx->stream;
    const uint64_t len = 1ull << n;
    DevBuf<gl::fp2> d_point, tmp;
    Status st;
    if ((st = d_point.alloc(ctx, n ? n : 1))) return st.code;
    if (n) {
        e = hipMemcpyAsync(d_point.ptr, point, 16ull * n, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "copy").code;
    }
    gl::fp2 *dst = (gl::fp2 *)out;
    const bool dev_out = is_device_ptr(out);
    if (!dev_out) {
        if ((st = tmp.alloc(ctx, len))) return st.code;
        dst = tmp.ptr;
    }
    if ((st = eq_table(ctx, d_point.ptr, n, dst, s))) return st.code;
    if (!dev_out) {
        e = hipMemcpyAsync(out, tmp.ptr, 16 * len, hipMemcpyDeviceToHost, s);
        if (e != hipSuccess) return hip_fail(e, "copy").code;
    }
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    return TSR_OK;
}
