// The register-resident Goldilocks butterfly rate of the device: the VALU ceiling of the NTT
// kernels, measured rather than quoted. Every lane runs independent chains of the library's own
// GS butterfly (u + v, (u - v) w, with gl::add / gl::sub / gl::mul) on 8 values held in
// registers, with no memory traffic inside the timed loop.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "common.h"
#include "field.h"

namespace tsr {

__global__ void __launch_bounds__(256) k_butterfly_peak(uint64_t *out, uint64_t seed, int iters) {
    const uint64_t gid = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    uint64_t x[8];
#pragma unroll
    for (int i = 0; i < 8; i++) x[i] = gl::canon(seed * (gid + 1) * (i + 3) + i);
    const uint64_t w = gl::canon(seed ^ (gid * 0x9E3779B97F4A7C15ull));
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const uint64_t u = x[2 * j], v = x[2 * j + 1];
            x[2 * j] = gl::add(u, v);
            x[2 * j + 1] = gl::mul(gl::sub(u, v), w);
        }
        const uint64_t t = x[1];  // mix the chains
        x[1] = x[2];
        x[2] = x[5];
        x[5] = x[6];
        x[6] = t;
    }
    uint64_t acc = 0;
#pragma unroll
    for (int i = 0; i < 8; i++) acc ^= x[i];
    out[gid] = acc;  // keeps the chains live
}

}  // namespace tsr

extern "C" tsr_status tsr_butterfly_peak(tsr_ctx *ctx_, double *butterflies_per_s) {
    using namespace tsr;
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !butterflies_per_s) return Status::fail(TSR_BAD_ARG, "null argument").code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    const int threads = 256, blocks = 256 * 16, iters = 2048;
    const uint64_t lanes = (uint64_t)threads * blocks;
    DevBuf<uint64_t> out;
    Status st = out.alloc(ctx, lanes);
    if (st) return st.code;
    hipEvent_t a, b;
    if ((e = hipEventCreate(&a)) != hipSuccess) return hip_fail(e, "event").code;
    if ((e = hipEventCreate(&b)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(e, "event").code;
    }
    hipLaunchKernelGGL(k_butterfly_peak, dim3(blocks), dim3(threads), 0, ctx->stream, out.ptr, 12345ull, iters);
    (void)hipEventRecord(a, ctx->stream);
    hipLaunchKernelGGL(k_butterfly_peak, dim3(blocks), dim3(threads), 0, ctx->stream, out.ptr, 12345ull, iters);
    (void)hipEventRecord(b, ctx->stream);
    e = hipEventSynchronize(b);
    float ms = 0;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (e != hipSuccess) return hip_fail(e, "butterfly peak").code;
    if ((st = check_launch("butterfly peak"))) return st.code;
    *butterflies_per_s = (double)lanes * iters * 4 / (ms * 1e-3);
    return TSR_OK;
}
