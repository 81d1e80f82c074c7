// tsr_merkle_root_noalloc: the Merkle root of tsr_merkle_root with no allocation at all, for
// callers that cannot allocate device memory (code under HIP graph capture, tight serving loops,
// harnesses that forbid hipMalloc). The tree's 2n - 1 digests live in a file-scope device array
// of a fixed size instead of a buffer from the context's pool.
//
// The array's name is a host-side stand-in, not a device address: its device address is taken
// once per device by hipGetSymbolAddress when a context is created on that device (with the
// device current), kept per device, and only that address is handed to kernels and copies.
// Every call on one GPU shares that GPU's array and leaves its last tree in it; calls that
// overlap on one GPU (two streams, two threads) race on it. Nothing in it needs to start at zero:
// a call writes every digest it reads.
#include <hip/hip_runtime.h>
#include <mutex>
#include <stdint.h>
#include <string>
#include "commit.h"
#include "common.h"

namespace tsr {

// The largest tree: 2^16 leaves, 2^17 - 1 digests of 8 words (4 MiB per GPU, held until the
// process ends).
constexpr uint64_t NOALLOC_MAX_LEAVES = 1ull << 16;
constexpr uint64_t NOALLOC_WORDS = 8 * (2 * NOALLOC_MAX_LEAVES - 1);

static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static uint32_t *g_scratch_ptr[NOALLOC_MAX_DEVICES];       // the array's address on each device
static hipError_t g_scratch_err[NOALLOC_MAX_DEVICES];      // the lookup's result, once done
static bool g_scratch_done[NOALLOC_MAX_DEVICES];

// Looks up the array's device address on `device`, which must be current. Called when a context
// is created on the device; the lookup runs once per device and its result is kept.
void noalloc_setup(int device) {
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_done[device]) return;
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
        (void)hipGetLastError();
        p = nullptr;
    }
    g_scratch_ptr[device] = (uint32_t *)p;
    g_scratch_err[device] = e;
    g_scratch_done[device] = true;
}

static Status noalloc_scratch(int device, uint32_t **out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (device < 0 || device >= NOALLOC_MAX_DEVICES || !g_scratch_done[device])
        return Status::fail(TSR_DEVICE, "the no-allocation scratch of this device was never set up");
    if (g_scratch_err[device] != hipSuccess || !g_scratch_ptr[device])
        return hip_fail(g_scratch_err[device], "hipGetSymbolAddress(scratch)");
    *out = g_scratch_ptr[device];
    return Status::ok();
}

}  // namespace tsr

using namespace tsr;

extern "C" tsr_status tsr_merkle_root_noalloc(tsr_ctx *ctx_, const uint8_t *leaves, uint64_t n,
                                              uint8_t root[32]) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !leaves || !root) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (n == 0 || (n & (n - 1)))
        return Status::fail(TSR_BAD_SHAPE,
                            "LeafCount: a Merkle tree needs a power-of-two number of leaves").code;
    if (n > NOALLOC_MAX_LEAVES)
        return Status::fail(TSR_DECLINED, "tsr_merkle_root_noalloc: " + std::to_string(n) +
                                              " leaves exceed its maximum of 2^16 (use tsr_merkle_root)")
            .code;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    uint32_t *nodes = nullptr;
    Status st = noalloc_scratch(ctx->device, &nodes);
    if (st) return st.code;
    // leaves (host or device) into the first n digests, then every level above them
    e = hipMemcpyAsync(nodes, leaves, 32 * n, hipMemcpyDefault, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "copy leaves").code;
    if ((st = merkle_build(ctx, nodes, n, ctx->stream, false))) return st.code;
    e = hipMemcpyAsync(root, nodes + 8 * (2 * n - 2), 32, hipMemcpyDefault, ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "copy root").code;
    e = tsr::stream_sync(ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    return TSR_OK;
}
