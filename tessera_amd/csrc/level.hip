// The reference's level offload as a C object (LevelOffload::load and DeviceLevel::{sums, bind,
// firsts}, crates/tessera-sumcheck/src/offload.rs:17-79): one level instance's tables on the
// device, driven round by round by the host's level prover (crates/tessera-sumcheck/src/level.rs).
//
//   sums(low, high, low_bits) = Sum_p w(p) P(line_x(p)) for x = 0, 2, 3 over the pairs p of the
//     current tables, w(p) = low[p mod 2^low_bits] high[p >> low_bits],
//     line_x(p) = t[2p] + x (t[2p+1] - t[2p]) per table, and P the shape's polynomial:
//     Fraction: N0 D1 + N1 D0 + lambda D0 D1 over [N0, N1, D0, D1]; Product: V0 V1 over [V0, V1];
//   bind(r): t'[i] = t[2i] + r (t[2i+1] - t[2i]) for every table (into the other buffer);
//   firsts(): each table's first entry.
// Sums are per-block partials added by one more block: field addition is exact, so the order
// of the additions does not change the result.
#include <hip/hip_runtime.h>
#include <memory>
#include <stdint.h>
#include <string>
#include <vector>
#include "common.h"
#include "field.h"

struct tsr_level {
    tsr::Ctx *ctx = nullptr;
    int shape = 0;  // TSR_LEVEL_FRACTION or TSR_LEVEL_PRODUCT
    gl::fp2 lambda{0, 0};
    uint32_t tables = 0;
    uint64_t len = 0;  // current entries per table
    int current = 0;
    std::vector<tsr::DevBuf<gl::fp2>> buf[2];
    tsr::DevBuf<gl::fp2> low, high, partials;
};

namespace tsr {

constexpr int LVL_THREADS = 256;
constexpr uint32_t LVL_MAX_BLOCKS = 1024;

struct LevelTabs {
    const gl::fp2 *t[4];
};

template <int SHAPE>
__device__ __forceinline__ gl::fp2 shape_poly(const gl::fp2 *v, gl::fp2 lambda) {
    if (SHAPE == 0) {  // N0 D1 + N1 D0 + lambda D0 D1
        const gl::fp2 d01 = gl::mul(v[2], v[3]);
        return gl::add(gl::add(gl::mul(v[0], v[3]), gl::mul(v[1], v[2])), gl::mul(lambda, d01));
    }
    return gl::mul(v[0], v[1]);
}

template <int SHAPE>
__global__ void __launch_bounds__(LVL_THREADS)
k_level_sums(LevelTabs tabs, uint64_t pairs, const gl::fp2 *__restrict__ low, uint32_t low_bits,
             const gl::fp2 *__restrict__ high, gl::fp2 lambda, gl::fp2 *__restrict__ partials) {
    constexpr int NT = SHAPE == 0 ? 4 : 2;
    __shared__ gl::fp2 red[3][LVL_THREADS];
    gl::fp2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};
    const uint64_t lmask = (1ull << low_bits) - 1;
    for (uint64_t p = (uint64_t)blockIdx.x * LVL_THREADS + threadIdx.x; p < pairs;
         p += (uint64_t)gridDim.x * LVL_THREADS) {
        const gl::fp2 w = gl::mul(low[p & lmask], high[p >> low_bits]);
        gl::fp2 a[NT], d[NT];
#pragma unroll
        for (int i = 0; i < NT; i++) {
            a[i] = tabs.t[i][2 * p];
            d[i] = gl::sub(tabs.t[i][2 * p + 1], a[i]);
        }
        gl::fp2 v[NT];
        // x = 0
#pragma unroll
        for (int i = 0; i < NT; i++) v[i] = a[i];
        acc[0] = gl::add(acc[0], gl::mul(w, shape_poly<SHAPE>(v, lambda)));
        // x = 2
#pragma unroll
        for (int i = 0; i < NT; i++) v[i] = gl::add(gl::add(a[i], d[i]), d[i]);
        acc[1] = gl::add(acc[1], gl::mul(w, shape_poly<SHAPE>(v, lambda)));
        // x = 3
#pragma unroll
        for (int i = 0; i < NT; i++) v[i] = gl::add(v[i], d[i]);
        acc[2] = gl::add(acc[2], gl::mul(w, shape_poly<SHAPE>(v, lambda)));
    }
#pragma unroll
    for (int j = 0; j < 3; j++) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int s = LVL_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int j = 0; j < 3; j++) red[j][threadIdx.x] = gl::add(red[j][threadIdx.x], red[j][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) partials[3 * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// partials[0..2] = the sums of `blocks` triples (one block of threads).
__global__ void __launch_bounds__(LVL_THREADS) k_level_total(gl::fp2 *partials, uint32_t blocks) {
    __shared__ gl::fp2 red[3][LVL_THREADS];
    gl::fp2 acc[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (uint32_t b = threadIdx.x; b < blocks; b += LVL_THREADS)
#pragma unroll
        for (int j = 0; j < 3; j++) acc[j] = gl::add(acc[j], partials[3 * b + j]);
#pragma unroll
    for (int j = 0; j < 3; j++) red[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int s = LVL_THREADS / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
#pragma unroll
            for (int j = 0; j < 3; j++) red[j][threadIdx.x] = gl::add(red[j][threadIdx.x], red[j][threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x < 3) partials[threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_level_bind(const gl::fp2 *__restrict__ t, gl::fp2 *__restrict__ out, uint64_t half,
                             gl::fp2 r) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= half) return;
    const gl::fp2 a = t[2 * i], b = t[2 * i + 1];
    out[i] = gl::add(a, gl::mul(r, gl::sub(b, a)));
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_level_load(tsr_ctx *ctx_, int shape, const uint64_t lambda[2],
                          const uint64_t *const *tables, uint32_t n_tables, uint32_t m,
                          tsr_level **out) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !tables || !out || (shape == TSR_LEVEL_FRACTION && !lambda))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (shape != TSR_LEVEL_FRACTION && shape != TSR_LEVEL_PRODUCT)
        return Status::fail(TSR_BAD_SHAPE, "unknown level shape").code;
    const uint32_t want = shape == TSR_LEVEL_FRACTION ? 4 : 2;
    if (n_tables != want) return Status::fail(TSR_BAD_SHAPE, "a fraction level has 4 tables, a product level 2").code;
    if (m < 1 || m > 32) return Status::fail(TSR_BAD_SHAPE, "tables of 2^m entries, 1 <= m <= 32").code;
    for (uint32_t i = 0; i < n_tables; i++)
        if (!tables[i]) return Status::fail(TSR_BAD_ARG, "null table").code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    std::unique_ptr<tsr_level> L(new tsr_level());
    L->ctx = ctx;
    L->shape = shape;
    if (shape == TSR_LEVEL_FRACTION) L->lambda = gl::fp2{lambda[0], lambda[1]};
    L->tables = n_tables;
    L->len = 1ull << m;
    Status st;
    for (uint32_t i = 0; i < n_tables; i++) {
        L->buf[0].emplace_back();
        L->buf[1].emplace_back();
        if ((st = L->buf[0][i].alloc(ctx, L->len)) || (st = L->buf[1][i].alloc(ctx, L->len / 2)))
            return st.code;
        e = hipMemcpyAsync(L->buf[0][i].ptr, tables[i], 16 * L->len, hipMemcpyDefault, ctx->stream);
        if (e != hipSuccess) return hip_fail(e, "level upload").code;
    }
    if ((st = L->partials.alloc(ctx, 3 * LVL_MAX_BLOCKS))) return st.code;
    e = tsr::stream_sync(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "level upload").code;
    *out = L.release();
    return TSR_OK;
}

tsr_status tsr_level_sums(tsr_level *L, const uint64_t *low, uint32_t low_bits, const uint64_t *high,
                          uint64_t high_len, uint64_t out[6]) {
    if (!L || !low || !high || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    const uint64_t pairs = L->len / 2;
    if (pairs == 0 || low_bits > 32 || ((unsigned __int128)high_len << low_bits) < pairs)
        return Status::fail(TSR_BAD_SHAPE, "level weights cover the pairs").code;
    Ctx *ctx = L->ctx;
    hipStream_t s = ctx->stream;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    const uint64_t nlow = 1ull << low_bits;
    Status st;
    if ((st = L->low.reserve(ctx, nlow)) || (st = L->high.reserve(ctx, high_len))) return st.code;
    if ((e = hipMemcpyAsync(L->low.ptr, low, 16 * nlow, hipMemcpyDefault, s)) != hipSuccess ||
        (e = hipMemcpyAsync(L->high.ptr, high, 16 * high_len, hipMemcpyDefault, s)) != hipSuccess)
        return hip_fail(e, "level weights").code;
    LevelTabs tabs{};
    for (uint32_t i = 0; i < L->tables; i++) tabs.t[i] = L->buf[L->current][i].ptr;
    const uint64_t want = (pairs + LVL_THREADS - 1) / LVL_THREADS;
    const uint32_t blocks = (uint32_t)(want < LVL_MAX_BLOCKS ? want : LVL_MAX_BLOCKS);
    if (L->shape == TSR_LEVEL_FRACTION)
        hipLaunchKernelGGL(k_level_sums<0>, dim3(blocks), dim3(LVL_THREADS), 0, s, tabs, pairs, L->low.ptr,
                           low_bits, L->high.ptr, L->lambda, L->partials.ptr);
    else
        hipLaunchKernelGGL(k_level_sums<1>, dim3(blocks), dim3(LVL_THREADS), 0, s, tabs, pairs, L->low.ptr,
                           low_bits, L->high.ptr, L->lambda, L->partials.ptr);
    hipLaunchKernelGGL(k_level_total, dim3(1), dim3(LVL_THREADS), 0, s, L->partials.ptr, blocks);
    if ((st = check_launch("level sums"))) return st.code;
    if ((e = hipMemcpyAsync(out, L->partials.ptr, 48, hipMemcpyDeviceToHost, s)) != hipSuccess ||
        (e = tsr::stream_sync(s)) != hipSuccess)
        return hip_fail(e, "level sums").code;
    return TSR_OK;
}

tsr_status tsr_level_bind(tsr_level *L, const uint64_t r[2]) {
    if (!L || !r) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (L->len < 2) return Status::fail(TSR_BAD_SHAPE, "every variable is bound").code;
    Ctx *ctx = L->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    const uint64_t half = L->len / 2;
    const int next = 1 - L->current;
    for (uint32_t i = 0; i < L->tables; i++) {
        if (L->buf[next][i].len < half) {
            Status st = L->buf[next][i].alloc(ctx, half);
            if (st) return st.code;
        }
        hipLaunchKernelGGL(k_level_bind, dim3((unsigned)((half + 255) / 256)), dim3(256), 0, ctx->stream,
                           L->buf[L->current][i].ptr, L->buf[next][i].ptr, half, gl::fp2{r[0], r[1]});
    }
    Status st = check_launch("level bind");
    if (st) return st.code;
    if ((e = tsr::stream_sync(ctx->stream)) != hipSuccess) return hip_fail(e, "level bind").code;
    L->current = next;
    L->len = half;
    return TSR_OK;
}

tsr_status tsr_level_firsts(tsr_level *L, uint64_t *out) {
    if (!L || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    Ctx *ctx = L->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    for (uint32_t i = 0; i < L->tables; i++)
        if ((e = hipMemcpyAsync(out + 2 * i, L->buf[L->current][i].ptr, 16, hipMemcpyDeviceToHost,
                                ctx->stream)) != hipSuccess)
            return hip_fail(e, "level firsts").code;
    if ((e = tsr::stream_sync(ctx->stream)) != hipSuccess) return hip_fail(e, "level firsts").code;
    return TSR_OK;
}

void tsr_level_free(tsr_level *L) {
    if (!L) return;
    (void)hipSetDevice(L->ctx->device);
    delete L;
}

}  // extern "C"
