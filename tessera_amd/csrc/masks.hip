// The advice mask stream on the device: the reference's per-proof mask source (Masks, a MaskSource,
// crates/tessera-prover/src/cpu/random.rs:123-229,275-320) is the BLAKE3 extended output of a
// 32-byte secret seed in derive-key mode under the context "tessera/v1/advice-masks"
// (crates/tessera-hash/src/domain.rs:90-127), read as u64 little-endian words; a word >= p is
// discarded whole and the next one is tried. Output block b of that stream is one compression:
// chaining value = the context key, message = seed || 0^32, counter b, block length 32, flags
// CHUNK_START | CHUNK_END | ROOT | DERIVE_KEY_MATERIAL, all 16 output words (8 u64).
//
// A draw of `count` masks from word position w0 runs three kernels over candidate words
// [w0, w0 + C), C = count plus a slack: each thread computes one output block and counts its
// accepted words (k_mask_count), one workgroup scans the per-workgroup counts (k_mask_scan), and
// each thread recomputes its block and writes its accepted words at their place in the draw
// (k_mask_write), recording the position after the count-th accepted word. When fewer than
// `count` candidates were accepted (about one rejection in 2^32 words makes that all but
// impossible) the draw runs again over twice the candidates: the result is exact either way.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "blake3.h"
#include "commit.h"
#include "common.h"
#include "field.h"

struct tsr_mask_stream {
    tsr::Ctx *ctx = nullptr;
    uint32_t key[8] = {0};   // public context key
    uint32_t seed[8] = {0};  // the secret seed, cleared on free
    uint64_t pos = 0;        // next candidate word of the stream
};

namespace tsr {

constexpr uint32_t MASK_THREADS = 256;
constexpr uint32_t MASK_SCAN_THREADS = 1024;

struct MaskKey {
    uint32_t key[8];
    uint32_t seed[8];
};

__device__ __forceinline__ void xof_block(const MaskKey &mk, uint64_t b, uint64_t w[8]) {
    uint32_t blk[16];
#pragma unroll
    for (int i = 0; i < 8; i++) {
        blk[i] = mk.seed[i];
        blk[8 + i] = 0;
    }
    uint32_t cv[8], full[16];
    b3::compress(mk.key, blk, b, 32,
                 b3::CHUNK_START | b3::CHUNK_END | b3::ROOT | b3::DERIVE_KEY_MATERIAL, cv, full);
#pragma unroll
    for (int j = 0; j < 8; j++) w[j] = (uint64_t)full[2 * j] | ((uint64_t)full[2 * j + 1] << 32);
}

// Bit j set: word j of block b0 + t lies in [w0, w0 + C) and is below p.
__device__ __forceinline__ uint32_t accepted_mask(const uint64_t w[8], uint64_t b, uint64_t w0,
                                                  uint64_t C) {
    uint32_t m = 0;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        const uint64_t idx = 8 * b + j;
        if (idx >= w0 && idx - w0 < C && w[j] < gl::P) m |= 1u << j;
    }
    return m;
}

__global__ void __launch_bounds__(MASK_THREADS)
k_mask_count(MaskKey mk, uint64_t w0, uint64_t C, uint64_t b0, uint64_t nb,
             uint32_t *__restrict__ counts) {
    __shared__ uint32_t total;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    const uint64_t t = (uint64_t)blockIdx.x * MASK_THREADS + threadIdx.x;
    if (t < nb) {
        uint64_t w[8];
        xof_block(mk, b0 + t, w);
        atomicAdd(&total, (uint32_t)__popc(accepted_mask(w, b0 + t, w0, C)));
    }
    __syncthreads();
    if (threadIdx.x == 0) counts[blockIdx.x] = total;
}

// offsets[g] = counts[0] + ... + counts[g-1] for g <= G (offsets[G] the total); one workgroup,
// each thread summing a contiguous run of counts, then a scan of the run sums in LDS.
__global__ void __launch_bounds__(MASK_SCAN_THREADS)
k_mask_scan(const uint32_t *__restrict__ counts, uint64_t G, uint64_t *__restrict__ offsets) {
    __shared__ uint64_t sums[2][MASK_SCAN_THREADS];
    const uint32_t i = threadIdx.x;
    const uint64_t run = (G + MASK_SCAN_THREADS - 1) / MASK_SCAN_THREADS;
    const uint64_t lo = i * run < G ? i * run : G, hi = lo + run < G ? lo + run : G;
    uint64_t own = 0;
    for (uint64_t g = lo; g < hi; g++) own += counts[g];
    int cur = 0;
    sums[cur][i] = own;
    __syncthreads();
    for (uint32_t d = 1; d < MASK_SCAN_THREADS; d <<= 1) {
        sums[cur ^ 1][i] = sums[cur][i] + (i >= d ? sums[cur][i - d] : 0);
        cur ^= 1;
        __syncthreads();
    }
    uint64_t at = sums[cur][i] - own;  // exclusive prefix of this run
    for (uint64_t g = lo; g < hi; g++) {
        offsets[g] = at;
        at += counts[g];
    }
    if (i == MASK_SCAN_THREADS - 1) offsets[G] = sums[cur][i];
}

__global__ void __launch_bounds__(MASK_THREADS)
k_mask_write(MaskKey mk, uint64_t w0, uint64_t C, uint64_t b0, uint64_t nb,
             const uint64_t *__restrict__ offsets, uint64_t count, uint64_t *__restrict__ out,
             uint64_t *__restrict__ endpos) {
    __shared__ uint32_t sc[2][MASK_THREADS];
    const uint64_t t = (uint64_t)blockIdx.x * MASK_THREADS + threadIdx.x;
    uint64_t w[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t acc = 0;
    if (t < nb) {
        xof_block(mk, b0 + t, w);
        acc = accepted_mask(w, b0 + t, w0, C);
    }
    const uint32_t own = (uint32_t)__popc(acc);
    int cur = 0;
    sc[cur][threadIdx.x] = own;
    __syncthreads();
    for (uint32_t d = 1; d < MASK_THREADS; d <<= 1) {
        sc[cur ^ 1][threadIdx.x] = sc[cur][threadIdx.x] + (threadIdx.x >= d ? sc[cur][threadIdx.x - d] : 0);
        cur ^= 1;
        __syncthreads();
    }
    uint64_t idx = offsets[blockIdx.x] + sc[cur][threadIdx.x] - own;
#pragma unroll
    for (int j = 0; j < 8; j++) {
        if (!((acc >> j) & 1)) continue;
        if (idx < count) out[idx] = w[j];
        if (idx == count - 1) *endpos = 8 * (b0 + t) + j + 1;
        idx++;
    }
}

// Draws `count` masks into the device buffer `out`, advancing the stream.
static Status mask_draw(tsr_mask_stream *m, uint64_t count, uint64_t *out) {
    Ctx *ctx = m->ctx;
    hipStream_t s = ctx->stream;
    if (count == 0) return Status::ok();
    struct Cleared : MaskKey {
        ~Cleared() {
            volatile uint32_t *p = seed;
            for (int i = 0; i < 8; i++) p[i] = 0;
        }
    } mk;
    memcpy(mk.key, m->key, 32);
    memcpy(mk.seed, m->seed, 32);
    uint64_t C = count + 64 + (count >> 20);
    Status st;
    for (;;) {
        const uint64_t w0 = m->pos, b0 = w0 / 8, nb = (w0 + C - 1) / 8 - b0 + 1;
        const uint64_t G = (nb + MASK_THREADS - 1) / MASK_THREADS;
        if (G > 0x7fffffffull) return Status::fail(TSR_DECLINED, "a mask draw beyond one launch");
        DevBuf<uint32_t> counts;
        DevBuf<uint64_t> offsets, endpos;
        if ((st = counts.alloc(ctx, G)) || (st = offsets.alloc(ctx, G + 1)) || (st = endpos.alloc(ctx, 1)))
            return st;
        {
            KernelTimer timer(ctx, "k_mask_count", s);
            hipLaunchKernelGGL(k_mask_count, dim3((unsigned)G), dim3(MASK_THREADS), 0, s, (MaskKey)mk, w0, C, b0,
                               nb, counts.ptr);
        }
        hipLaunchKernelGGL(k_mask_scan, dim3(1), dim3(MASK_SCAN_THREADS), 0, s, counts.ptr, G, offsets.ptr);
        {
            KernelTimer timer(ctx, "k_mask_write", s);
            hipLaunchKernelGGL(k_mask_write, dim3((unsigned)G), dim3(MASK_THREADS), 0, s, (MaskKey)mk, w0, C, b0,
                               nb, offsets.ptr, count, out, endpos.ptr);
        }
        if ((st = check_launch("mask draw"))) return st;
        uint64_t total = 0, end = 0;
        TSR_HIP(hipMemcpyAsync(&total, offsets.ptr + G, 8, hipMemcpyDeviceToHost, s));
        TSR_HIP(hipMemcpyAsync(&end, endpos.ptr, 8, hipMemcpyDeviceToHost, s));
        TSR_HIP(tsr::stream_sync(s));
        if (total >= count) {
            m->pos = end;
            return Status::ok();
        }
        C *= 2;
    }
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_mask_stream_create(tsr_ctx *ctx, const uint8_t seed[32], tsr_mask_stream **out) {
    if (!ctx || !seed || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    std::unique_ptr<tsr_mask_stream> m(new tsr_mask_stream());
    m->ctx = (Ctx *)ctx;
    derive_context_key("tessera/v1/advice-masks", m->key);
    for (int i = 0; i < 8; i++)
        m->seed[i] = (uint32_t)seed[4 * i] | ((uint32_t)seed[4 * i + 1] << 8) |
                     ((uint32_t)seed[4 * i + 2] << 16) | ((uint32_t)seed[4 * i + 3] << 24);
    *out = m.release();
    return TSR_OK;
}

tsr_status tsr_mask_stream_draw(tsr_mask_stream *m, uint64_t count, uint64_t *out) {
    if (!m || (count && !out)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (count > (1ull << 40)) return Status::fail(TSR_BAD_SHAPE, "a draw of more than 2^40 masks").code;
    Ctx *ctx = m->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    if (count == 0) return TSR_OK;
    if (is_device_ptr(out)) return mask_draw(m, count, out).code;
    // a host destination: draw into device memory, copy, clear the device copy
    DevBuf<uint64_t> tmp;
    Status st = tmp.alloc(ctx, count);
    if (!st) st = mask_draw(m, count, tmp.ptr);
    if (!st) {
        hipError_t c = hipMemcpyAsync(out, tmp.ptr, 8 * count, hipMemcpyDeviceToHost, ctx->stream);
        if (c != hipSuccess) st = hip_fail(c, "copy");
    }
    if (tmp.ptr) zero_words(tmp.ptr, count, ctx->stream);
    e = tsr::stream_sync(ctx->stream);
    if (!st && e != hipSuccess) st = hip_fail(e, "sync");
    return st.code;
}

uint64_t tsr_mask_stream_position(const tsr_mask_stream *m) { return m ? m->pos : 0; }

void tsr_mask_stream_free(tsr_mask_stream *m) {
    if (!m) return;
    volatile uint32_t *p = m->seed;
    for (int i = 0; i < 8; i++) p[i] = 0;
    delete m;
}

}  // extern "C"
