// The C ABI (include/tessera_mi355x.h): contexts, buffers, encoding, column chains, Merkle
// roots and the unmasked commitment. Shapes are checked before any launch, in the reference's
// order; failures return a status and set the thread's error string.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <mutex>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "blake3.h"
#include "commit.h"
#include "common.h"
#include "field.h"
#include "ntt.h"

struct tsr_ctx : tsr::Ctx {};

struct tsr_chains {
    tsr::Ctx *ctx = nullptr;
    uint64_t columns = 0, block = 0;
    bool closed = false;
    tsr::DevBuf<uint32_t> state;
};

namespace tsr {

static thread_local std::string g_error;

void set_error(const std::string &msg) { g_error = msg; }

Status hip_fail(hipError_t e, const char *what) {
    return Status::fail(e == hipErrorOutOfMemory ? TSR_OOM : TSR_DEVICE,
                        std::string(what) + ": " + hipGetErrorString(e));
}

Status check_launch(const char *what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, what);
    return Status::ok();
}

static void pool_trim_locked(Ctx *ctx) {
    for (auto &kv : ctx->pool) (void)hipFree(kv.second);
    ctx->pool.clear();
    ctx->pooled = 0;
}

void pool_trim(Ctx *ctx) {
    std::lock_guard<std::mutex> lk(ctx->pool_mu);
    pool_trim_locked(ctx);
}

// Allocation size classes: powers of two and their midpoints (at most a third wasted), so that
// buffers of many nearby sizes (served requests of different lengths) reuse pooled blocks.
static uint64_t size_class(uint64_t bytes) {
    if (bytes <= 4096) return 4096;
    uint64_t p = 1ull << (63 - __builtin_clzll(bytes));
    if (bytes == p) return p;
    if (bytes <= p + p / 2) return p + p / 2;
    return 2 * p;
}

Status dev_alloc(Ctx *ctx, uint64_t bytes, void **out) {
    std::lock_guard<std::mutex> lk(ctx->pool_mu);
    bytes = size_class(bytes);
    if (ctx->budget && ctx->in_use + bytes > ctx->budget)
        return Status::fail(TSR_OOM, "device memory budget exceeded: " +
                                         std::to_string(ctx->in_use + bytes) + " > " +
                                         std::to_string(ctx->budget) + " bytes");
    auto it = ctx->pool.find(bytes);
    if (it != ctx->pool.end()) {
        *out = it->second;
        ctx->pool.erase(it);
        ctx->pooled -= bytes;
        ctx->in_use += bytes;
        return Status::ok();
    }
    if (ctx->budget && ctx->in_use + ctx->pooled + bytes > ctx->budget) pool_trim_locked(ctx);
    hipError_t e = hipMalloc(out, bytes);
    if (e == hipErrorOutOfMemory && ctx->pooled) {
        (void)hipGetLastError();
        (void)tsr::stream_sync(ctx->stream);
        pool_trim_locked(ctx);
        e = hipMalloc(out, bytes);
    }
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return hip_fail(e, "hipMalloc");
    }
    ctx->in_use += bytes;
    return Status::ok();
}

Status pinned_host(Ctx *ctx, size_t bytes, uint8_t **out) {
    if (bytes > ctx->pinned_cap) {
        if (ctx->pinned) {
            (void)tsr::stream_sync(ctx->stream);
            (void)hipHostFree(ctx->pinned);
            ctx->pinned = nullptr;
            ctx->pinned_cap = 0;
        }
        const size_t cap = std::max<size_t>(bytes, 64 * 1024);
        hipError_t e = hipHostMalloc(&ctx->pinned, cap, hipHostMallocDefault);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            ctx->pinned = nullptr;
            return hip_fail(e, "hipHostMalloc");
        }
        ctx->pinned_cap = cap;
    }
    *out = (uint8_t *)ctx->pinned;
    return Status::ok();
}

void dev_free(Ctx *ctx, void *p, uint64_t bytes) {
    if (!p) return;
    if (!ctx) {
        (void)hipFree(p);
        return;
    }
    std::lock_guard<std::mutex> lk(ctx->pool_mu);
    bytes = size_class(bytes);
    ctx->in_use -= bytes;
    ctx->pool.emplace(bytes, p);
    ctx->pooled += bytes;
}

static hipEvent_t take_event(Ctx *ctx) {
    if (!ctx->spare.empty()) {
        hipEvent_t e = ctx->spare.back();
        ctx->spare.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) {
        (void)hipGetLastError();
        return nullptr;
    }
    return e;
}

KernelTimer::KernelTimer(Ctx *c, const char *n, hipStream_t s) : ctx(c), name(n), stream(s) {
    if (ctx && ctx->profile) {
        start = take_event(ctx);
        if (start) (void)hipEventRecord(start, stream);
    }
}

KernelTimer::~KernelTimer() {
    if (!start) return;
    hipEvent_t stop = take_event(ctx);
    if (!stop) return;
    (void)hipEventRecord(stop, stream);
    ctx->timed.push_back(Ctx::Timed{name, start, stop});
}

bool is_device_ptr(const void *p) {
    if (!p) return false;
    hipPointerAttribute_t attr;
    hipError_t e = hipPointerGetAttributes(&attr, p);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return attr.type == hipMemoryTypeDevice;
}

static const char *const DOMAINS[DOM_COUNT] = {
    "tessera/v1/envelope",         "tessera/v1/commitment-id",     "tessera/v1/column-init",
    "tessera/v1/column-step",      "tessera/v1/merkle-node",       "tessera/v1/column-merge",
    "tessera/v1/weights-root",     "tessera/v1/transcript-init",   "tessera/v1/transcript-absorb",
    "tessera/v1/transcript-squeeze", "tessera/v1/transcript-ratchet", "tessera/v1/pow",
    "tessera/v1/coin-commit",      "tessera/v1/coin-toss",         "tessera/v1/table",
    "tessera/v1/audit-sample",     "tessera/v1/audit-select",      "tessera/v1/claim-list",
    "tessera/v1/statement",
};

const char *domain_context(int d) { return d >= 0 && d < DOM_COUNT ? DOMAINS[d] : nullptr; }

// Context key: BLAKE3 of the context string in DERIVE_KEY_CONTEXT mode. Every context string
// here is shorter than 64 bytes, so it is one block of one chunk.
void derive_context_key(const char *context, uint32_t key[8]) {
    static const uint32_t iv[8] = {B3_IV0, B3_IV1, B3_IV2, B3_IV3,
                                   B3_IV4, B3_IV5, B3_IV6, B3_IV7};
    uint8_t bytes[64] = {0};
    size_t len = strlen(context);
    memcpy(bytes, context, len < 64 ? len : 63);
    uint32_t blk[16];
    for (int i = 0; i < 16; i++)
        blk[i] = (uint32_t)bytes[4 * i] | ((uint32_t)bytes[4 * i + 1] << 8) |
                 ((uint32_t)bytes[4 * i + 2] << 16) | ((uint32_t)bytes[4 * i + 3] << 24);
    b3::compress(iv, blk, 0, (uint32_t)len,
                 b3::CHUNK_START | b3::CHUNK_END | b3::ROOT | b3::DERIVE_KEY_CONTEXT, key);
}

void derive_domain_keys(uint32_t keys[DOM_COUNT][8]) {
    for (int d = 0; d < DOM_COUNT; d++) derive_context_key(DOMAINS[d], keys[d]);
}

// SplitMix64 stream value i (state seed + (i+1) * 0x9E3779B97F4A7C15), before rejection.
__device__ __forceinline__ uint64_t splitmix_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Stream values with words >= p skipped: out[i] = the i-th accepted value. `rejects` lists the
// (rare, probability 2^-32 each) rejected stream positions below the output's reach, sorted.
__global__ void k_fill_uniform(uint64_t *__restrict__ out, uint64_t count, uint64_t seed,
                               const uint64_t *__restrict__ rejects, uint32_t n_rejects) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    uint32_t j = 0;
    while (j < n_rejects && rejects[j] <= i + j) j++;
    out[i] = splitmix_at(seed, i + j);
}

__global__ void k_find_rejects(uint64_t span, uint64_t seed, uint64_t *__restrict__ pos,
                               uint32_t *__restrict__ n, uint32_t cap) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= span) return;
    if (splitmix_at(seed, i) >= gl::P) {
        uint32_t slot = atomicAdd(n, 1u);
        if (slot < cap) pos[slot] = i;
    }
}

}  // namespace tsr

using namespace tsr;

#define CHECK_ARG(cond, msg)                                            \
    do {                                                                \
        if (!(cond)) return Status::fail(TSR_BAD_ARG, msg).code;        \
    } while (0)
#define CHECK_SHAPE(cond, msg)                                          \
    do {                                                                \
        if (!(cond)) return Status::fail(TSR_BAD_SHAPE, msg).code;      \
    } while (0)
#define TRY(expr)                                                       \
    do {                                                                \
        Status st__ = (expr);                                           \
        if (st__) return st__.code;                                     \
    } while (0)
#define TRY_HIP(expr)                                                   \
    do {                                                                \
        hipError_t e__ = (expr);                                        \
        if (e__ != hipSuccess) return hip_fail(e__, #expr).code;        \
    } while (0)

extern "C" {

const char *tsr_last_error(void) { return g_error.c_str(); }
const char *tsr_version(void) { return "tessera_mi355x 0.1.0 (gfx950)"; }

tsr_status tsr_ctx_create(int device, tsr_ctx **out) {
    CHECK_ARG(out, "null output pointer");
    int count = 0;
    TRY_HIP(hipGetDeviceCount(&count));
    if (device < 0 || device >= count)
        return Status::fail(TSR_DECLINED, "no GPU " + std::to_string(device)).code;
    TRY_HIP(hipSetDevice(device));
    std::unique_ptr<tsr_ctx> c(new tsr_ctx());
    c->device = device;
    TRY_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    derive_domain_keys(c->keys);
    // the no-allocation entry points' scratch address on this device (the lookup's result is
    // kept and reported by those entry points)
    noalloc_setup(device);
    *out = c.release();
    g_error.clear();
    return TSR_OK;
}

void tsr_ctx_destroy(tsr_ctx *ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)tsr::stream_sync(ctx->stream);
    ctx->encoders.clear();
    for (auto &t : ctx->timed) {
        (void)hipEventDestroy(t.start);
        (void)hipEventDestroy(t.stop);
    }
    for (auto e : ctx->spare) (void)hipEventDestroy(e);
    ctx->scratch.release();
    ctx->staging_in.release();
    ctx->staging_out.release();
    if (ctx->pinned) (void)hipHostFree(ctx->pinned);
    for (auto &kv : ctx->user_allocs) (void)hipFree(kv.first);  // blocks never given back
    ctx->user_allocs.clear();
    pool_trim(ctx);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

tsr_status tsr_ctx_synchronize(tsr_ctx *ctx) {
    CHECK_ARG(ctx, "null context");
    TRY_HIP(hipSetDevice(ctx->device));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    return TSR_OK;
}

tsr_status tsr_ctx_set_round_blocks(tsr_ctx *ctx, uint32_t blocks) {
    CHECK_ARG(ctx, "null context");
    CHECK_ARG(blocks <= 1024, "at most 1024 workgroups per round launch");
    ctx->round_blocks = blocks;
    return TSR_OK;
}

tsr_status tsr_ctx_profile(tsr_ctx *ctx, int enable) {
    CHECK_ARG(ctx, "null context");
    if (enable == 2) {
        TRY_HIP(hipSetDevice(ctx->device));
        TRY_HIP(tsr::stream_sync(ctx->stream));
        for (auto &t : ctx->timed) {
            ctx->spare.push_back(t.start);
            ctx->spare.push_back(t.stop);
        }
        ctx->timed.clear();
        return TSR_OK;
    }
    ctx->profile = enable != 0;
    return TSR_OK;
}

tsr_status tsr_ctx_kernel_time(tsr_ctx *ctx, const char *kernel, double *total_ms,
                               uint64_t *launches) {
    CHECK_ARG(ctx && kernel, "null argument");
    TRY_HIP(hipSetDevice(ctx->device));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    double total = 0;
    uint64_t count = 0;
    for (auto &t : ctx->timed) {
        if (strcmp(t.name, kernel) != 0) continue;
        float ms = 0;
        TRY_HIP(hipEventElapsedTime(&ms, t.start, t.stop));
        total += ms;
        count++;
    }
    if (total_ms) *total_ms = total;
    if (launches) *launches = count;
    return TSR_OK;
}

tsr_status tsr_ctx_set_budget(tsr_ctx *ctx, uint64_t bytes) {
    CHECK_ARG(ctx, "null context");
    ctx->budget = bytes;
    return TSR_OK;
}

uint64_t tsr_ctx_memory(const tsr_ctx *ctx) { return ctx ? ctx->in_use : 0; }

tsr_status tsr_alloc(tsr_ctx *ctx, uint64_t bytes, void **out) {
    CHECK_ARG(ctx && out, "null argument");
    TRY_HIP(hipSetDevice(ctx->device));
    TRY(dev_alloc(ctx, bytes ? bytes : 1, out));
    std::lock_guard<std::mutex> lk(ctx->pool_mu);
    ctx->user_allocs[*out] = size_class(bytes ? bytes : 1);
    return TSR_OK;
}

// The block goes back to the context's pool, like the library's own buffers, with the bytes
// dev_alloc charged for it; a pointer tsr_alloc did not return is refused untouched.
tsr_status tsr_free(tsr_ctx *ctx, void *ptr) {
    CHECK_ARG(ctx, "null context");
    if (!ptr) return TSR_OK;
    uint64_t charged = 0;
    {
        std::lock_guard<std::mutex> lk(ctx->pool_mu);
        auto it = ctx->user_allocs.find(ptr);
        if (it == ctx->user_allocs.end())
            return Status::fail(TSR_BAD_ARG, "tsr_free: not a pointer from tsr_alloc").code;
        charged = it->second;
        ctx->user_allocs.erase(it);
    }
    TRY_HIP(hipSetDevice(ctx->device));
    // work queued on the stream may still read or write the block
    TRY_HIP(tsr::stream_sync(ctx->stream));
    dev_free(ctx, ptr, charged);
    return TSR_OK;
}

tsr_status tsr_copy(tsr_ctx *ctx, void *dst, const void *src, uint64_t bytes) {
    CHECK_ARG(ctx && (bytes == 0 || (dst && src)), "null argument");
    TRY_HIP(hipSetDevice(ctx->device));
    TRY_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDefault, ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    return TSR_OK;
}

tsr_status tsr_fill_uniform(tsr_ctx *ctx, uint64_t *dst, uint64_t count, uint64_t seed) {
    CHECK_ARG(ctx && (count == 0 || dst), "null argument");
    CHECK_ARG(count == 0 || is_device_ptr(dst), "tsr_fill_uniform writes device memory only");
    if (count == 0) return TSR_OK;
    TRY_HIP(hipSetDevice(ctx->device));
    const uint32_t cap = 64;
    DevBuf<uint64_t> pos;
    DevBuf<uint32_t> cnt;
    TRY(pos.alloc(ctx, cap));
    TRY(cnt.alloc(ctx, 1));
    TRY_HIP(hipMemsetAsync(cnt.ptr, 0, 4, ctx->stream));
    uint64_t span = count + cap;
    hipLaunchKernelGGL(k_find_rejects, dim3((unsigned)((span + 255) / 256)), dim3(256), 0,
                       ctx->stream, span, seed, pos.ptr, cnt.ptr, cap);
    TRY(check_launch("find rejects"));
    uint32_t n_rej = 0;
    TRY_HIP(hipMemcpyAsync(&n_rej, cnt.ptr, 4, hipMemcpyDeviceToHost, ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    if (n_rej > cap) return Status::fail(TSR_DECLINED, "too many rejected words").code;
    std::vector<uint64_t> p(n_rej);
    if (n_rej) {
        TRY_HIP(hipMemcpy(p.data(), pos.ptr, 8 * n_rej, hipMemcpyDeviceToHost));
        std::sort(p.begin(), p.end());
        TRY_HIP(hipMemcpy(pos.ptr, p.data(), 8 * n_rej, hipMemcpyHostToDevice));
    }
    hipLaunchKernelGGL(k_fill_uniform, dim3((unsigned)((count + 255) / 256)), dim3(256), 0,
                       ctx->stream, dst, count, seed, pos.ptr, n_rej);
    TRY(check_launch("fill uniform"));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    return TSR_OK;
}

// ---- encoding ----

static tsr_status check_code(uint32_t log_k, uint32_t log_r) {
    if (log_r == 0)
        return Status::fail(TSR_BAD_SHAPE, "InverseRate: log2_inv_rate must be at least 1").code;
    if ((uint64_t)log_k + log_r > 32)
        return Status::fail(TSR_BAD_SHAPE, "TooLong: n = 2^(log_k + r) exceeds 2^32").code;
    return TSR_OK;
}

tsr_status tsr_encode_rows(tsr_ctx *ctx, uint32_t log_k, uint32_t log_r, uint64_t P,
                           const uint64_t *rows, const uint64_t *masks, uint64_t count,
                           uint64_t *codewords) {
    CHECK_ARG(ctx, "null context");
    tsr_status s = check_code(log_k, log_r);
    if (s) return s;
    CHECK_SHAPE(P <= (1ull << log_k), "MaskTooLarge: mask_coeffs exceeds k");
    if (count == 0) return TSR_OK;
    CHECK_ARG(rows && codewords && (P == 0 || masks), "null argument");
    TRY_HIP(hipSetDevice(ctx->device));
    const uint64_t k = 1ull << log_k, n = k << log_r;
    Encoder *encp = nullptr;
    TRY(get_encoder(ctx, log_k, log_r, P, &encp));
    const Encoder &enc = *encp;
    const uint64_t *drows = rows, *dmasks = masks;
    uint64_t *dcw = codewords;
    DevBuf<uint64_t> in_buf, mask_buf, out_buf;
    if (!is_device_ptr(rows)) {
        TRY(in_buf.alloc(ctx, count * k));
        TRY_HIP(hipMemcpyAsync(in_buf.ptr, rows, 8 * count * k, hipMemcpyHostToDevice,
                               ctx->stream));
        drows = in_buf.ptr;
    }
    if (P && !is_device_ptr(masks)) {
        TRY(mask_buf.alloc(ctx, count * P));
        TRY_HIP(hipMemcpyAsync(mask_buf.ptr, masks, 8 * count * P, hipMemcpyHostToDevice,
                               ctx->stream));
        dmasks = mask_buf.ptr;
    }
    bool host_out = !is_device_ptr(codewords);
    if (host_out) {
        TRY(out_buf.alloc(ctx, count * n));
        dcw = out_buf.ptr;
    }
    TRY(ctx->scratch.reserve(ctx, enc.scratch_elems(count)));
    TRY(enc.encode(drows, dmasks, count, dcw, ctx->scratch.ptr, ctx->stream));
    if (host_out)
        TRY_HIP(hipMemcpyAsync(codewords, dcw, 8 * count * n, hipMemcpyDeviceToHost,
                               ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    return TSR_OK;
}

// ---- column chains ----

tsr_status tsr_chains_create(tsr_ctx *ctx, const uint8_t id[32], uint64_t columns,
                             uint64_t block, tsr_chains **out) {
    CHECK_ARG(ctx && id && out, "null argument");
    CHECK_SHAPE(columns >= 1, "a commitment has at least one column");
    CHECK_SHAPE(block >= 1, "the row block B is at least 1");
    TRY_HIP(hipSetDevice(ctx->device));
    std::unique_ptr<tsr_chains> c(new tsr_chains());
    c->ctx = ctx;
    c->columns = columns;
    c->block = block;
    TRY(c->state.alloc(ctx, 8 * columns));
    TRY(chains_init(ctx, id, columns, c->state.ptr, ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    *out = c.release();
    return TSR_OK;
}

tsr_status tsr_chains_absorb(tsr_chains *c, const uint64_t *rows, uint64_t count) {
    CHECK_ARG(c, "null chains");
    if (c->closed) return Status::fail(TSR_BAD_SHAPE, "AfterFinalBlock: a block followed the short final block").code;
    if (count == 0 || count > c->block || count > 124)
        return Status::fail(TSR_BAD_SHAPE, "BlockSize: a block holds 1 to " +
                                               std::to_string(c->block) + " rows, not " +
                                               std::to_string(count)).code;
    CHECK_ARG(rows, "null rows");
    Ctx *ctx = c->ctx;
    TRY_HIP(hipSetDevice(ctx->device));
    const uint64_t *d = rows;
    DevBuf<uint64_t> buf;
    if (!is_device_ptr(rows)) {
        TRY(buf.alloc(ctx, count * c->columns));
        TRY_HIP(hipMemcpyAsync(buf.ptr, rows, 8 * count * c->columns, hipMemcpyHostToDevice,
                               ctx->stream));
        d = buf.ptr;
    }
    TRY(chains_absorb(ctx, d, c->columns, (uint32_t)count, c->state.ptr, ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    c->closed = count < c->block;
    return TSR_OK;
}

tsr_status tsr_chains_read(tsr_chains *c, uint8_t *out) {
    CHECK_ARG(c && out, "null argument");
    TRY_HIP(hipSetDevice(c->ctx->device));
    TRY_HIP(hipMemcpyAsync(out, c->state.ptr, 32 * c->columns, hipMemcpyDefault,
                           c->ctx->stream));
    TRY_HIP(tsr::stream_sync(c->ctx->stream));
    return TSR_OK;
}

void tsr_chains_free(tsr_chains *c) {
    if (!c) return;
    (void)hipSetDevice(c->ctx->device);
    delete c;
}

// ---- Merkle ----

tsr_status tsr_merkle_root(tsr_ctx *ctx, const uint8_t *leaves, uint64_t n, uint8_t root[32]) {
    CHECK_ARG(ctx && leaves && root, "null argument");
    CHECK_SHAPE(n >= 1 && (n & (n - 1)) == 0,
                "LeafCount: a Merkle tree needs a power-of-two number of leaves");
    TRY_HIP(hipSetDevice(ctx->device));
    DevBuf<uint32_t> nodes;
    TRY(nodes.alloc(ctx, 8 * (2 * n - 1)));
    TRY_HIP(hipMemcpyAsync(nodes.ptr, leaves, 32 * n, hipMemcpyDefault, ctx->stream));
    TRY(merkle_build(ctx, nodes.ptr, n, ctx->stream));
    TRY_HIP(hipMemcpyAsync(root, nodes.ptr + 8 * (2 * n - 2), 32, hipMemcpyDeviceToHost,
                           ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    return TSR_OK;
}

// ---- the unmasked commitment ----

tsr_status tsr_commit_unmasked(tsr_ctx *ctx, const tsr_commitment_params *params,
                               const uint64_t *rows, uint64_t count, tsr_commitment **out) {
    CHECK_ARG(ctx && params && out, "null argument");
    tsr_status s = check_code(params->log2_k, params->log2_inv_rate);
    if (s) return s;
    if (params->mask_coeffs != 0)
        return Status::fail(TSR_BAD_SHAPE, "UnexpectedMask: an unmasked commitment has P = 0").code;
    CHECK_SHAPE(params->row_block >= 1, "the row block B is at least 1");
    CHECK_ARG(count == 0 || rows, "null rows");
    TRY_HIP(hipSetDevice(ctx->device));
    const uint32_t log_k = params->log2_k, log_r = params->log2_inv_rate;
    const uint64_t k = 1ull << log_k, n = k << log_r;
    const uint32_t block = (uint32_t)std::min<uint64_t>(params->row_block, 124);
    if (params->row_block > 124)
        return Status::fail(TSR_DECLINED, "row blocks above 124 rows are hashed on the CPU").code;

    std::unique_ptr<tsr_commitment> c(new tsr_commitment());
    c->ctx = ctx;
    c->params = *params;
    c->count = count;
    c->n = n;
    TRY(c->nodes.alloc(ctx, 8 * (2 * n - 1)));
    TRY(c->codewords.alloc(ctx, count * n));
    const uint64_t *drows = rows;
    DevBuf<uint64_t> in_buf;
    if (count) {
        if (!is_device_ptr(rows)) {
            TRY(in_buf.alloc(ctx, count * k));
            TRY_HIP(hipMemcpyAsync(in_buf.ptr, rows, 8 * count * k, hipMemcpyHostToDevice,
                                   ctx->stream));
            drows = in_buf.ptr;
        }
        Encoder *enc = nullptr;
        TRY(get_encoder(ctx, log_k, log_r, 0, &enc));
        TRY(ctx->scratch.reserve(ctx, enc->scratch_elems(count)));
        TRY(enc->encode(drows, nullptr, count, c->codewords.ptr, ctx->scratch.ptr, ctx->stream));
    }
    TRY(chain_leaves(ctx, c->codewords.ptr, n, count, block, params->commitment_id,
                     c->nodes.ptr, ctx->stream));
    TRY(merkle_build(ctx, c->nodes.ptr, n, ctx->stream));
    TRY_HIP(hipMemcpyAsync(c->root, c->nodes.ptr + 8 * (2 * n - 2), 32, hipMemcpyDeviceToHost,
                           ctx->stream));
    TRY_HIP(tsr::stream_sync(ctx->stream));
    *out = c.release();
    return TSR_OK;
}

}  // extern "C"

// Steps 2 to 4 of a commitment from its resident codewords (ReferenceCommitment::from_encoded,
// commitment.rs:418-450): chained leaves (one segment, leaf = the chain's last digest) or merged
// leaves (one chain set per segment, leaf_j = H(column-merge, d_j^(1) || ... || d_j^(S))), the
// Merkle tree, the root.
static Status finish_commitment(tsr_commitment *c, const uint64_t *segment_rows, uint32_t segments,
                                bool merged) {
    Ctx *ctx = c->ctx;
    const uint64_t n = c->n;
    const uint32_t block = (uint32_t)c->params.row_block;
    Status st;
    if (!merged) {
        if ((st = chain_leaves(ctx, c->codewords.ptr, n, c->count, block, c->params.commitment_id,
                               c->nodes.ptr, ctx->stream)))
            return st;
    } else {
        DevBuf<uint32_t> dig;
        if ((st = dig.alloc(ctx, 8 * n * (uint64_t)segments))) return st;
        uint64_t off = 0;
        for (uint32_t sgm = 0; sgm < segments; sgm++) {
            if ((st = chain_leaves(ctx, c->codewords.ptr + off * n, n, segment_rows[sgm], block,
                                   c->params.commitment_id, dig.ptr + 8 * n * (uint64_t)sgm, ctx->stream)))
                return st;
            off += segment_rows[sgm];
        }
        if ((st = merge_leaves(ctx, dig.ptr, segments, n, c->nodes.ptr, ctx->stream))) return st;
    }
    if ((st = merkle_build(ctx, c->nodes.ptr, n, ctx->stream))) return st;
    TSR_HIP(hipMemcpyAsync(c->root, c->nodes.ptr + 8 * (2 * n - 2), 32, hipMemcpyDeviceToHost, ctx->stream));
    TSR_HIP(tsr::stream_sync(ctx->stream));
    return Status::ok();
}

// The shape checks of a commitment with `segments` segments, in the reference's order
// (CommitmentParams::shape / unmasked_shape, commitment.rs:145-185); the rows they sum to.
static Status commitment_shape(const tsr_commitment_params *params, bool masked,
                               const uint64_t *segment_rows, uint32_t segments, uint64_t *count) {
    tsr_status s = check_code(params->log2_k, params->log2_inv_rate);
    if (s) return Status{s};
    const uint64_t k = 1ull << params->log2_k;
    if (masked && params->mask_coeffs == 0)
        return Status::fail(TSR_BAD_SHAPE, "NoMask: a masked commitment has P >= 1");
    if (masked && params->mask_coeffs > k)
        return Status::fail(TSR_BAD_SHAPE, "MaskTooLarge: mask_coeffs exceeds k");
    if (!masked && params->mask_coeffs != 0)
        return Status::fail(TSR_BAD_SHAPE, "UnexpectedMask: an unmasked commitment has P = 0");
    if (segments == 0) return Status::fail(TSR_BAD_SHAPE, "NoSegments: a layout has a segment");
    if (params->row_block == 0) return Status::fail(TSR_BAD_SHAPE, "the row block B is at least 1");
    if (params->row_block > 124)
        return Status::fail(TSR_DECLINED, "row blocks above 124 rows are hashed on the CPU");
    uint64_t total = 0;
    for (uint32_t i = 0; i < segments; i++) total += segment_rows[i];
    *count = total;
    return Status::ok();
}

extern "C" {

tsr_status tsr_commit(tsr_ctx *ctx, const tsr_commitment_params *params, const uint64_t *segment_rows,
                      uint32_t segments, const uint64_t *rows, const uint64_t *masks,
                      tsr_commitment **out) {
    CHECK_ARG(ctx && params && out && (segments == 0 || segment_rows), "null argument");
    uint64_t count = 0;
    TRY(commitment_shape(params, true, segment_rows, segments, &count));
    CHECK_ARG(count == 0 || (rows && masks), "null rows or masks");
    TRY_HIP(hipSetDevice(ctx->device));
    const uint32_t log_k = params->log2_k, log_r = params->log2_inv_rate;
    const uint64_t k = 1ull << log_k, n = k << log_r, P = params->mask_coeffs;
    std::unique_ptr<tsr_commitment> c(new tsr_commitment());
    c->ctx = ctx;
    c->params = *params;
    c->count = count;
    c->n = n;
    TRY(c->nodes.alloc(ctx, 8 * (2 * n - 1)));
    TRY(c->codewords.alloc(ctx, count * n));
    c->secret = true;
    TRY(c->rows.alloc(ctx, count * k));
    TRY(c->masks.alloc(ctx, count * P));
    if (count) {
        TRY_HIP(hipMemcpyAsync(c->rows.ptr, rows, 8 * count * k, hipMemcpyDefault, ctx->stream));
        TRY_HIP(hipMemcpyAsync(c->masks.ptr, masks, 8 * count * P, hipMemcpyDefault, ctx->stream));
        Encoder *enc = nullptr;
        TRY(get_encoder(ctx, log_k, log_r, P, &enc));
        TRY(ctx->scratch.reserve(ctx, enc->scratch_elems(count)));
        TRY(enc->encode(c->rows.ptr, c->masks.ptr, count, c->codewords.ptr, ctx->scratch.ptr, ctx->stream));
    }
    TRY(finish_commitment(c.get(), segment_rows, segments, true));
    *out = c.release();
    return TSR_OK;
}

tsr_status tsr_commit_from_encoded(tsr_ctx *ctx, const tsr_commitment_params *params,
                                   const uint64_t *segment_rows, uint32_t segments, int merged,
                                   const uint64_t *codewords, tsr_commitment **out) {
    CHECK_ARG(ctx && params && out && (segments == 0 || segment_rows), "null argument");
    uint64_t count = 0;
    TRY(commitment_shape(params, merged != 0, segment_rows, segments, &count));
    CHECK_SHAPE(merged || segments == 1, "a commitment of chained leaves has one segment");
    CHECK_ARG(count == 0 || codewords, "null codewords");
    TRY_HIP(hipSetDevice(ctx->device));
    const uint64_t n = 1ull << (params->log2_k + params->log2_inv_rate);
    std::unique_ptr<tsr_commitment> c(new tsr_commitment());
    c->ctx = ctx;
    c->params = *params;
    c->count = count;
    c->n = n;
    TRY(c->nodes.alloc(ctx, 8 * (2 * n - 1)));
    TRY(c->codewords.alloc(ctx, count * n));
    if (count)
        TRY_HIP(hipMemcpyAsync(c->codewords.ptr, codewords, 8 * count * n, hipMemcpyDefault, ctx->stream));
    TRY(finish_commitment(c.get(), segment_rows, segments, merged != 0));
    *out = c.release();
    return TSR_OK;
}

tsr_status tsr_commitment_root(const tsr_commitment *c, uint8_t root[32]) {
    CHECK_ARG(c && root, "null argument");
    memcpy(root, c->root, 32);
    return TSR_OK;
}

tsr_status tsr_commitment_leaves(const tsr_commitment *c, uint8_t *out) {
    CHECK_ARG(c && out, "null argument");
    TRY_HIP(hipSetDevice(c->ctx->device));
    TRY_HIP(hipMemcpyAsync(out, c->nodes.ptr, 32 * c->n, hipMemcpyDefault, c->ctx->stream));
    TRY_HIP(tsr::stream_sync(c->ctx->stream));
    return TSR_OK;
}

tsr_status tsr_commitment_codewords(const tsr_commitment *c, const uint64_t **dev,
                                    uint64_t *count, uint64_t *n) {
    CHECK_ARG(c, "null commitment");
    if (dev) *dev = c->codewords.ptr;
    if (count) *count = c->count;
    if (n) *n = c->n;
    return TSR_OK;
}

tsr_status tsr_commitment_answer(const tsr_commitment *c, const uint64_t *queries, uint64_t q,
                                 uint64_t *distinct_out, uint64_t distinct_cap,
                                 uint64_t *columns_out, uint8_t *siblings_out,
                                 uint64_t siblings_cap, uint64_t *n_distinct,
                                 uint64_t *n_siblings) {
    CHECK_ARG(c, "null commitment");
    if (q == 0) return Status::fail(TSR_BAD_SHAPE, "NoQueries: no column was queried").code;
    CHECK_ARG(queries, "null queries");
    for (uint64_t i = 0; i < q; i++)
        if (queries[i] >= c->n)
            return Status::fail(TSR_BAD_SHAPE, "QueryRange: query " + std::to_string(i) + " = " +
                                                   std::to_string(queries[i]) + " >= " +
                                                   std::to_string(c->n)).code;
    std::vector<uint64_t> d(queries, queries + q);
    std::sort(d.begin(), d.end());
    d.erase(std::unique(d.begin(), d.end()), d.end());
    std::vector<uint64_t> pos = multiproof_positions(c->n, d);
    if (n_distinct) *n_distinct = d.size();
    if (n_siblings) *n_siblings = pos.size();
    if (d.size() > distinct_cap || pos.size() > siblings_cap)
        return Status::fail(TSR_BAD_SHAPE, "answer buffers are too small").code;
    CHECK_ARG(distinct_out && (c->count == 0 || columns_out) && (pos.empty() || siblings_out),
              "null output");
    Ctx *ctx = c->ctx;
    TRY_HIP(hipSetDevice(ctx->device));
    memcpy(distinct_out, d.data(), 8 * d.size());
    // one upload of both index lists and the readbacks through the context's pinned block, so
    // every copy is asynchronous and the call waits once
    const size_t nd = d.size(), np = pos.size();
    const size_t up_bytes = 8 * (nd + np), col_bytes = 8 * nd * c->count, sib_bytes = 32 * np;
    uint8_t *pin;
    TRY(pinned_host(ctx, up_bytes + col_bytes + sib_bytes, &pin));
    memcpy(pin, d.data(), 8 * nd);
    if (np) memcpy(pin + 8 * nd, pos.data(), 8 * np);
    DevBuf<uint64_t> idx, cols;
    DevBuf<uint32_t> sib;
    TRY(idx.alloc(ctx, nd + np));
    TRY_HIP(hipMemcpyAsync(idx.ptr, pin, up_bytes, hipMemcpyHostToDevice, ctx->stream));
    if (c->count) {
        TRY(cols.alloc(ctx, nd * c->count));
        TRY(gather_columns(c->codewords.ptr, c->n, c->count, idx.ptr, nd, cols.ptr, ctx->stream));
        TRY_HIP(hipMemcpyAsync(pin + up_bytes, cols.ptr, col_bytes, hipMemcpyDeviceToHost, ctx->stream));
    }
    if (np) {
        TRY(sib.alloc(ctx, 8 * np));
        TRY(gather_digests(c->nodes.ptr, idx.ptr + nd, np, sib.ptr, ctx->stream));
        TRY_HIP(hipMemcpyAsync(pin + up_bytes + col_bytes, sib.ptr, sib_bytes, hipMemcpyDeviceToHost,
                               ctx->stream));
    }
    TRY_HIP(tsr::stream_sync(ctx->stream));
    if (col_bytes) memcpy(columns_out, pin + up_bytes, col_bytes);
    if (sib_bytes) memcpy(siblings_out, pin + up_bytes + col_bytes, sib_bytes);
    return TSR_OK;
}

tsr_status tsr_row_coefficients(tsr_ctx *ctx, uint32_t row_vars, const uint64_t *rows, uint64_t N,
                                const uint64_t *row_points, const uint64_t *gammas,
                                const uint32_t *group_sizes, uint32_t groups, const uint64_t *xi_point,
                                uint64_t *out) {
    CHECK_ARG(ctx && out && (N == 0 || rows) && (groups == 0 || group_sizes) && (row_vars == 0 || xi_point),
              "null argument");
    CHECK_SHAPE(row_vars <= 64, "RowVars: a row index has at most 64 bits");
    TRY_HIP(hipSetDevice(ctx->device));
    std::vector<uint32_t> first(groups + 1, 0);
    for (uint32_t g = 0; g < groups; g++) first[g + 1] = first[g] + group_sizes[g];
    const uint64_t claims = first[groups];
    CHECK_ARG(claims == 0 || (gammas && (row_vars == 0 || row_points)), "null claims");
    if (N == 0) return TSR_OK;
    hipStream_t s = ctx->stream;
    DevBuf<uint64_t> d_rows, d_points, d_gammas, d_xi;
    DevBuf<uint32_t> d_first;
    DevBuf<gl::fp2> d_out;
    TRY(d_rows.alloc(ctx, N));
    TRY(d_points.alloc(ctx, std::max<uint64_t>(1, 2 * claims * row_vars)));
    TRY(d_gammas.alloc(ctx, std::max<uint64_t>(1, 2 * claims)));
    TRY(d_xi.alloc(ctx, std::max<uint32_t>(1, 2 * row_vars)));
    TRY(d_first.alloc(ctx, groups + 1));
    TRY_HIP(hipMemcpyAsync(d_rows.ptr, rows, 8 * N, hipMemcpyDefault, s));
    if (claims && row_vars) TRY_HIP(hipMemcpyAsync(d_points.ptr, row_points, 16 * claims * row_vars, hipMemcpyDefault, s));
    if (claims) TRY_HIP(hipMemcpyAsync(d_gammas.ptr, gammas, 16 * claims, hipMemcpyDefault, s));
    if (row_vars) TRY_HIP(hipMemcpyAsync(d_xi.ptr, xi_point, 16 * row_vars, hipMemcpyDefault, s));
    TRY_HIP(hipMemcpyAsync(d_first.ptr, first.data(), 4 * (groups + 1), hipMemcpyHostToDevice, s));
    gl::fp2 *dst = (gl::fp2 *)out;
    if (!is_device_ptr(out)) {
        TRY(d_out.alloc(ctx, (uint64_t)(groups + 1) * N));
        dst = d_out.ptr;
    }
    TRY(row_coefficients(ctx, row_vars, d_rows.ptr, N, (const gl::fp2 *)d_points.ptr,
                         (const gl::fp2 *)d_gammas.ptr, d_first.ptr, groups, (const gl::fp2 *)d_xi.ptr,
                         dst, s));
    if (!is_device_ptr(out))
        TRY_HIP(hipMemcpyAsync(out, dst, 16 * (uint64_t)(groups + 1) * N, hipMemcpyDeviceToHost, s));
    TRY_HIP(tsr::stream_sync(s));
    return TSR_OK;
}

tsr_status tsr_commitment_open(const tsr_commitment *c, const uint64_t *coefficients, uint32_t count,
                               uint64_t *u_out, uint64_t *mask_out) {
    CHECK_ARG(c && (count == 0 || (coefficients && u_out && mask_out)), "null argument");
    CHECK_SHAPE(c->secret, "open combines the rows of a masked (advice) commitment, which keeps them");
    Ctx *ctx = c->ctx;
    TRY_HIP(hipSetDevice(ctx->device));
    if (count == 0) return TSR_OK;
    hipStream_t s = ctx->stream;
    const uint64_t N = c->count, k = 1ull << c->params.log2_k, P = c->params.mask_coeffs;
    DevBuf<gl::fp2> d_coef, d_u, d_m;
    const gl::fp2 *coef = (const gl::fp2 *)coefficients;
    if (!is_device_ptr(coefficients)) {
        TRY(d_coef.alloc(ctx, std::max<uint64_t>(1, (uint64_t)count * N)));
        if (N) TRY_HIP(hipMemcpyAsync(d_coef.ptr, coefficients, 16 * (uint64_t)count * N, hipMemcpyHostToDevice, s));
        coef = d_coef.ptr;
    }
    TRY(d_u.alloc(ctx, (uint64_t)count * k));
    TRY(d_m.alloc(ctx, (uint64_t)count * P));
    TRY(combine_rows(ctx, coef, count, c->rows.ptr, N, k, d_u.ptr, s));
    TRY(combine_rows(ctx, coef, count, c->masks.ptr, N, P, d_m.ptr, s));
    TRY_HIP(hipMemcpyAsync(u_out, d_u.ptr, 16 * (uint64_t)count * k, hipMemcpyDefault, s));
    TRY_HIP(hipMemcpyAsync(mask_out, d_m.ptr, 16 * (uint64_t)count * P, hipMemcpyDefault, s));
    TRY_HIP(tsr::stream_sync(s));
    return TSR_OK;
}

void tsr_commitment_free(tsr_commitment *c) {
    if (!c) return;
    (void)hipSetDevice(c->ctx->device);
    if (c->secret) {
        zero_words(c->rows.ptr, c->rows.len, c->ctx->stream);
        zero_words(c->masks.ptr, c->masks.len, c->ctx->stream);
        zero_words(c->codewords.ptr, c->codewords.len, c->ctx->stream);
    }
    (void)tsr::stream_sync(c->ctx->stream);
    delete c;
}

}  // extern "C"
