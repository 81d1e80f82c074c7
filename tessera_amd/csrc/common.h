// Host-side plumbing shared by the library's translation units: status values, the context
// (one GPU, one stream, a memory budget), device buffers and error reporting.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <chrono>
#include <map>
#include <mutex>
#include <memory>
#include <string>
#include <tuple>
#include <vector>
#include "../../include/tessera_mi355x.h"

namespace tsr {

void set_error(const std::string &msg);

struct Status {
    tsr_status code = TSR_OK;
    static Status ok() { return Status{}; }
    static Status fail(tsr_status c, const std::string &msg) {
        set_error(msg);
        Status s;
        s.code = c;
        return s;
    }
    explicit operator bool() const { return code != TSR_OK; }  // true on failure
};

Status hip_fail(hipError_t e, const char *what);
Status check_launch(const char *what);

// Waits for a stream's work: polls for up to 200 us, then blocks (hipStreamSynchronize). The
// waits of small proofs are short, and a blocking wait returns tens of microseconds after the
// GPU is done (256 x 2^14 commit + 14-variable sumcheck: 0.58 -> 0.55 ms); a long wait does
// not keep a host core busy.
inline hipError_t stream_sync(hipStream_t s) {
    const auto t0 = std::chrono::steady_clock::now();
    for (;;) {
        const hipError_t e = hipStreamQuery(s);
        if (e != hipErrorNotReady) return e;
        if (std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(200))
            return hipStreamSynchronize(s);
        __builtin_ia32_pause();
    }
}

#define TSR_HIP(expr)                                         \
    do {                                                      \
        hipError_t e__ = (expr);                              \
        if (e__ != hipSuccess) return hip_fail(e__, #expr);   \
    } while (0)

struct Ctx;
struct Encoder;

// Device memory accounted against the context's budget. Freed blocks go to the context's pool
// (exact-size reuse on the same stream, so no synchronization is needed) until trimmed.
Status dev_alloc(Ctx *ctx, uint64_t bytes, void **out);
void dev_free(Ctx *ctx, void *p, uint64_t bytes);
void pool_trim(Ctx *ctx);

template <typename T>
struct DevBuf {
    Ctx *ctx = nullptr;
    T *ptr = nullptr;
    uint64_t len = 0;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    DevBuf(DevBuf &&o) noexcept : ctx(o.ctx), ptr(o.ptr), len(o.len) {
        o.ptr = nullptr;
        o.len = 0;
    }
    DevBuf &operator=(DevBuf &&o) noexcept {
        if (this != &o) {
            release();
            ctx = o.ctx;
            ptr = o.ptr;
            len = o.len;
            o.ptr = nullptr;
            o.len = 0;
        }
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr) dev_free(ctx, ptr, len * sizeof(T));
        ptr = nullptr;
        len = 0;
    }
    Status alloc(Ctx *c, uint64_t n) {
        release();
        ctx = c;
        if (n == 0) n = 1;
        void *p = nullptr;
        Status st = dev_alloc(c, n * sizeof(T), &p);
        if (st) return st;
        ptr = (T *)p;
        len = n;
        return Status::ok();
    }
    // grow-only reuse
    Status reserve(Ctx *c, uint64_t n) {
        if (ptr && len >= n && ctx == c) return Status::ok();
        return alloc(c, n);
    }
    Status upload(Ctx *c, const std::vector<T> &v);
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    uint64_t budget = 0;  // 0: none
    uint64_t in_use = 0;  // live allocations
    uint64_t pooled = 0;  // freed blocks kept for reuse
    std::multimap<uint64_t, void *> pool;
    std::mutex pool_mu;  // dev_alloc / dev_free may come from several threads (served proving)
    std::map<void *, uint64_t> user_allocs;  // tsr_alloc pointers and the bytes charged for them
    // batched commitments: the tuned (tile_log, max pass bits) per (log k, rate, log2 rows),
    // so that a later batch of the same shapes skips the tuner
    std::map<std::tuple<uint32_t, uint32_t, uint32_t>, std::pair<uint32_t, uint32_t>> batch_tunings;
    std::map<std::tuple<uint32_t, uint32_t, uint64_t, uint32_t>, std::shared_ptr<Encoder>> encoders;
    uint32_t keys[19][8];  // BLAKE3 context keys of the reference's 19 domains
    bool profile = false;
    uint32_t round_blocks = 0;  // workgroups of a sumcheck round launch at most (0: the default)
    struct Timed {
        const char *name;
        hipEvent_t start, stop;
    };
    std::vector<Timed> timed;          // recorded launches (profiling)
    std::vector<hipEvent_t> spare;     // event pool
    DevBuf<uint64_t> scratch;      // reusable per-call scratch (encode temporaries)
    DevBuf<uint64_t> staging_in;   // host-input staging
    DevBuf<uint64_t> staging_out;  // host-output staging
    // Pinned host memory for the small per-call uploads and readbacks of the sumcheck rounds
    // (grown on demand, freed with the context; a context's calls run one at a time).
    void *pinned = nullptr;
    size_t pinned_cap = 0;
};

// At least `bytes` of the context's pinned host buffer (hipHostMalloc, grown when short).
Status pinned_host(Ctx *ctx, size_t bytes, uint8_t **out);

template <typename T>
Status DevBuf<T>::upload(Ctx *c, const std::vector<T> &v) {
    Status st = alloc(c, v.size());
    if (st) return st;
    TSR_HIP(hipMemcpyAsync(ptr, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice,
                           c->stream));
    TSR_HIP(stream_sync(c->stream));
    return Status::ok();
}

// Records an event pair around a launch on `stream` when the context is profiling.
struct KernelTimer {
    Ctx *ctx;
    const char *name;
    hipStream_t stream;
    hipEvent_t start = nullptr;
    KernelTimer(Ctx *c, const char *n, hipStream_t s);
    ~KernelTimer();
};

// Whether p is device memory of the current device (else it is treated as host memory).
bool is_device_ptr(const void *p);

// Domain indices (crates/tessera-hash/src/domain.rs:41-81, declaration order).
enum Domain : int {
    DOM_ENVELOPE = 0, DOM_COMMITMENT_ID, DOM_COLUMN_INIT, DOM_COLUMN_STEP, DOM_MERKLE_NODE,
    DOM_COLUMN_MERGE, DOM_WEIGHTS_ROOT, DOM_TRANSCRIPT_INIT, DOM_TRANSCRIPT_ABSORB,
    DOM_TRANSCRIPT_SQUEEZE, DOM_TRANSCRIPT_RATCHET, DOM_POW, DOM_COIN_COMMIT, DOM_COIN_TOSS,
    DOM_TABLE, DOM_AUDIT_SAMPLE, DOM_AUDIT_SELECT, DOM_CLAIM_LIST, DOM_STATEMENT, DOM_COUNT
};
const char *domain_context(int d);
void derive_domain_keys(uint32_t keys[DOM_COUNT][8]);
// The BLAKE3 context key of a context string shorter than 64 bytes.
void derive_context_key(const char *context, uint32_t key[8]);

}  // namespace tsr
