// Advice openings on the device (RFC-0011 steps 2 to 5; crates/tessera-pcs/src/opening.rs):
// the row coefficients of the claim groups and the proximity point, and the combined rows
// u_g = Sum_row c_g(row) A_row and M_g = Sum_row c_g(row) m_row over a commitment's resident
// rows and masks. Field sums are exact, so the split of the row sums over blocks does not
// change the result.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <stdint.h>
#include <string.h>
#include <vector>
#include "common.h"
#include "commit.h"
#include "field.h"

namespace tsr {

using gl::fp2;

// out[g * N + i] = Sum_{claims c of g} gamma_c eq(point_c, rows[i]) for g < groups, and
// eq(xi, rows[i]) for g = groups (RowLayout::row_coefficients, opening.rs:116-158).
__global__ void k_row_coefficients(uint32_t row_vars, const uint64_t *__restrict__ rows, uint64_t N,
                                   const fp2 *__restrict__ points, const fp2 *__restrict__ gammas,
                                   const uint32_t *__restrict__ first, uint32_t groups,
                                   const fp2 *__restrict__ xi, fp2 *__restrict__ out) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t g = blockIdx.y;
    if (i >= N) return;
    const uint64_t row = rows[i];
    const fp2 one{1, 0};
    fp2 acc{0, 0};
    const uint32_t c0 = g < groups ? first[g] : 0, c1 = g < groups ? first[g + 1] : 1;
    for (uint32_t c = c0; c < c1; c++) {
        const fp2 *pt = g < groups ? points + (uint64_t)c * row_vars : xi;
        fp2 e = g < groups ? gammas[c] : one;
        for (uint32_t t = 0; t < row_vars; t++) e = gl::mul(e, (row >> t) & 1 ? pt[t] : gl::sub(one, pt[t]));
        acc = gl::add(acc, e);
    }
    out[(uint64_t)g * N + i] = acc;
}

constexpr int CMB_THREADS = 256, CMB_MAXG = 16;
constexpr uint64_t CMB_ROWS = 256;  // rows per block of the partial sums

// partials[(chunk * count + g) * width + i] = Sum_{r in chunk} coef[g * N + r] m[r * width + i]
template <int COUNT>
__global__ void __launch_bounds__(CMB_THREADS)
k_combine(const fp2 *__restrict__ coef, const uint64_t *__restrict__ m, uint64_t N, uint64_t width,
          fp2 *__restrict__ partials) {
    const uint64_t i = (uint64_t)blockIdx.x * CMB_THREADS + threadIdx.x;
    const uint64_t chunk = blockIdx.y, r0 = chunk * CMB_ROWS;
    if (i >= width) return;
    const uint64_t r1 = r0 + CMB_ROWS < N ? r0 + CMB_ROWS : N;
    fp2 acc[COUNT];
#pragma unroll
    for (int g = 0; g < COUNT; g++) acc[g] = fp2{0, 0};
    for (uint64_t r = r0; r < r1; r++) {
        const uint64_t x = m[r * width + i];
#pragma unroll
        for (int g = 0; g < COUNT; g++) acc[g] = gl::add(acc[g], gl::mul_base(coef[(uint64_t)g * N + r], x));
    }
#pragma unroll
    for (int g = 0; g < COUNT; g++) partials[(chunk * COUNT + g) * width + i] = acc[g];
}

// out[g * width + i] = Sum_chunks partials[(chunk * count + g) * width + i]
__global__ void k_combine_total(const fp2 *__restrict__ partials, uint64_t chunks, uint32_t count,
                                uint64_t width, fp2 *__restrict__ out) {
    const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count * width) return;
    fp2 acc{0, 0};
    for (uint64_t c = 0; c < chunks; c++) acc = gl::add(acc, partials[c * count * width + t]);
    out[t] = acc;
}

// The combined rows of a row-major N x width Fp matrix on the device into `out` (device).
Status combine_rows(Ctx *ctx, const fp2 *coef, uint32_t count, const uint64_t *m, uint64_t N,
                    uint64_t width, fp2 *out, hipStream_t s) {
    if (count == 0 || width == 0) return Status::ok();
    const uint64_t chunks = std::max<uint64_t>(1, (N + CMB_ROWS - 1) / CMB_ROWS);
    DevBuf<fp2> partials;
    Status st;
    if ((st = partials.alloc(ctx, chunks * count * width))) return st;
    const dim3 grid((unsigned)((width + CMB_THREADS - 1) / CMB_THREADS), (unsigned)chunks);
    for (uint32_t g0 = 0; g0 < count; g0 += CMB_MAXG) {
        // groups g0.. in batches of up to CMB_MAXG (the accumulators stay in registers)
        const uint32_t gc = std::min<uint32_t>(CMB_MAXG, count - g0);
        const fp2 *cg = coef + (uint64_t)g0 * N;
        switch (gc) {
#define CMB_CASE(G)                                                                                   \
    case G:                                                                                           \
        hipLaunchKernelGGL(k_combine<G>, grid, dim3(CMB_THREADS), 0, s, cg, m, N, width, partials.ptr); \
        break;
            CMB_CASE(1) CMB_CASE(2) CMB_CASE(3) CMB_CASE(4) CMB_CASE(5) CMB_CASE(6) CMB_CASE(7) CMB_CASE(8)
            CMB_CASE(9) CMB_CASE(10) CMB_CASE(11) CMB_CASE(12) CMB_CASE(13) CMB_CASE(14) CMB_CASE(15) CMB_CASE(16)
#undef CMB_CASE
        }
        const uint64_t total = (uint64_t)gc * width;
        hipLaunchKernelGGL(k_combine_total, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s,
                           partials.ptr, chunks, gc, width, out + (uint64_t)g0 * width);
    }
    return check_launch("combine");
}

Status row_coefficients(Ctx *ctx, uint32_t row_vars, const uint64_t *rows, uint64_t N,
                        const fp2 *points, const fp2 *gammas, const uint32_t *first, uint32_t groups,
                        const fp2 *xi, fp2 *out, hipStream_t s) {
    if (N == 0) return Status::ok();
    hipLaunchKernelGGL(k_row_coefficients, dim3((unsigned)((N + 255) / 256), groups + 1), dim3(256), 0, s,
                       row_vars, rows, N, points, gammas, first, groups, xi, out);
    return check_launch("row coefficients");
}

__global__ void k_zero_u64(uint64_t *__restrict__ p, uint64_t n) {
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

void zero_words(uint64_t *p, uint64_t n, hipStream_t s) {
    if (p && n)
        hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n);
}

}  // namespace tsr
