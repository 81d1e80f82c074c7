// Served proving with continuous batching: requests of different lengths are admitted as worker
// slots and device-memory budget free up, and the work of admitted requests overlaps on the GPU.
//
// No reference counterpart (the reference proves one statement per call and has no serving
// loop). Each request's pipeline only uses the library's own entry points, so its result is
// the same alone and in any mix:
//   1. decode: the synthetic integer transformer (model.hip) generates the tokens greedily and
//      writes the witness, Goldilocks elements, WIT_STRIDE per token;
//   2. commit: the witness as rows of k = 2^13 values (zero-padded), rate 1/4, B = 32, unmasked
//      (tsr_commit_unmasked: encode, column chains, Merkle root);
//   3. transcript: SessionContext("tessera-protocol/1", NonInteractive, "test-small/1", zero
//      statement digest, session id = the request id LE), absorb the root under "serve/root",
//      squeeze n challenges "serve/r/{i}";
//   4. sumcheck: stage/0, degree 3, Sum_x eq(r, x) A(x) B(x) over layer 0's `a` and `m`
//      activations of every token (zero-padded to 2^n entries), on the device.
// Workers: `slots` threads, each with its own context (HIP stream), share the model weights.
// Admission is strict FIFO: the oldest waiting request starts when a slot is free and its
// device-memory footprint fits beside the footprints of the requests in flight.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <stdint.h>
#include <string.h>
#include <string>
#include <thread>
#include <vector>
#include "common.h"
#include "field.h"
#include "model.h"
#include "sumcheck.h"
#include "transcript.h"

namespace tsr {

constexpr uint32_t SERVE_LOG_K = 13;

__global__ void k_gather_base(const uint64_t *__restrict__ wit, uint32_t tokens, uint64_t offset,
                              uint64_t len, uint64_t *__restrict__ out) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= len) return;
    const uint64_t t = i / D_MODEL, c = i % D_MODEL;
    out[i] = t < tokens ? wit[t * WIT_STRIDE + offset + c] : 0;
}

__global__ void k_fill_zero(uint64_t *__restrict__ p, uint64_t n) {
    uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

static double now_s(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

static uint32_t ceil_log2(uint64_t x) {
    uint32_t h = 0;
    while ((1ull << h) < x) h++;
    return h;
}

// The device memory one request of `total` tokens needs (its admission footprint).
uint64_t request_footprint(uint32_t total) {
    const uint64_t k = 1ull << SERVE_LOG_K, wit = (uint64_t)total * WIT_STRIDE;
    const uint64_t rows = (wit + k - 1) / k, n = 4 * k;
    const uint64_t table = 1ull << ceil_log2((uint64_t)total * D_MODEL);
    return 8 * rows * k           // witness rows
           + 8 * rows * n         // codewords
           + 8 * rows * k         // encode scratch
           + 32 * 2 * n           // tree
           + (uint64_t)total * N_LAYERS * 2 * D_MODEL  // KV cache
           + 3 * 16 * table * 2;  // sumcheck tables and working buffers
}

struct Job {
    tsr_request req;
    std::vector<uint32_t> prompt;
    double t_submit = 0;
};

}  // namespace tsr

struct tsr_server {
    std::chrono::steady_clock::time_point t0;
    int device = 0;
    uint32_t slots = 1;
    uint64_t budget = 0, in_flight_bytes = 0;
    tsr_ctx *owner = nullptr;  // holds the weights
    int8_t *weights = nullptr;
    std::vector<tsr_ctx *> ctxs;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::mutex prove_mu;  // tsr_server_prove calls share the owner context: one at a time
    std::condition_variable cv_work, cv_done;
    std::deque<tsr::Job> queue;
    std::deque<tsr_result> done;
    bool stop = false;
    uint32_t active = 0, max_active = 0;
};

namespace tsr {

// The whole pipeline of one request on one context; fills `out` (not the timestamps).
static Status prove_request(Ctx *ctx, const int8_t *weights, const tsr_request &req,
                            tsr_result &out) {
    hipStream_t s = ctx->stream;
    const uint32_t total = req.n_prompt + req.n_output;
    const uint64_t k = 1ull << SERVE_LOG_K, wit_len = (uint64_t)total * WIT_STRIDE;
    const uint64_t rows = (wit_len + k - 1) / k;
    Status st;
    DevBuf<uint32_t> d_prompt, d_tokens;
    DevBuf<int8_t> kv;
    DevBuf<uint64_t> wit;
    if ((st = d_prompt.alloc(ctx, req.n_prompt)) || (st = d_tokens.alloc(ctx, total)) ||
        (st = kv.alloc(ctx, (uint64_t)N_LAYERS * total * 2 * D_MODEL)) ||
        (st = wit.alloc(ctx, rows * k)))
        return st;
    TSR_HIP(hipMemcpyAsync(d_prompt.ptr, req.prompt, 4ull * req.n_prompt, hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)((rows * k - wit_len + 255) / 256 + 1)), dim3(256),
                       0, s, wit.ptr + wit_len, rows * k - wit_len);
    if ((st = model_decode(ctx, weights, d_prompt.ptr, req.n_prompt, total, d_tokens.ptr, kv.ptr,
                           wit.ptr, s)))
        return st;

    // commit the witness rows
    tsr_commitment_params params{};
    params.log2_k = SERVE_LOG_K;
    params.log2_inv_rate = 2;
    params.mask_coeffs = 0;
    params.row_block = 32;
    memset(params.commitment_id, 0x5e, 32);
    tsr_commitment *c = nullptr;
    tsr_status cs = tsr_commit_unmasked((tsr_ctx *)ctx, &params, wit.ptr, rows, &c);
    if (cs != TSR_OK) return Status::fail(cs, std::string("serve commit: ") + tsr_last_error());
    tsr_commitment_root(c, out.root);
    tsr_commitment_free(c);

    // the transcript: the request's own session
    std::vector<uint8_t> tce;
    auto ascii = [&](const char *str) {
        uint64_t n = strlen(str);
        for (int i = 0; i < 8; i++) tce.push_back((uint8_t)(n >> (8 * i)));
        tce.insert(tce.end(), str, str + n);
    };
    ascii("tessera-protocol/1");
    tce.push_back(1);  // Mode::NonInteractive
    tce.push_back(0);
    ascii("test-small/1");
    tce.insert(tce.end(), 32, 0);
    for (int i = 0; i < 16; i++) tce.push_back(i < 8 ? (uint8_t)(req.id >> (8 * i)) : 0);
    Transcript tr;
    tr.init(ctx, tce.data(), tce.size());
    tr.absorb("serve/root", out.root, 32);
    const uint32_t n = ceil_log2((uint64_t)total * D_MODEL);
    std::vector<gl::fp2> r(n);
    for (uint32_t i = 0; i < n; i++) r[i] = tr.challenge_fp2(("serve/r/" + std::to_string(i)).c_str(), nullptr);

    // tables: eq(r, .), layer 0's a and m of every token
    const uint64_t len = 1ull << n;
    DevBuf<uint64_t> A, B;
    DevBuf<gl::fp2> eq, d_r;
    if ((st = A.alloc(ctx, len)) || (st = B.alloc(ctx, len)) || (st = eq.alloc(ctx, len)) ||
        (st = d_r.alloc(ctx, std::max(n, 1u))))
        return st;
    if (n) TSR_HIP(hipMemcpyAsync(d_r.ptr, r.data(), 16 * n, hipMemcpyHostToDevice, s));
    if ((st = eq_table(ctx, d_r.ptr, n, eq.ptr, s))) return st;
    hipLaunchKernelGGL(k_gather_base, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, wit.ptr,
                       total, WIT_A0, len, A.ptr);
    hipLaunchKernelGGL(k_gather_base, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, wit.ptr,
                       total, WIT_M0, len, B.ptr);
    if ((st = check_launch("serve tables"))) return st;
    tsr_table tabs[3] = {{(const uint64_t *)eq.ptr, 1, 1}, {A.ptr, 0, 0}, {B.ptr, 0, 0}};
    tsr_instance inst{n, 3, tabs};
    SumcheckOut so;
    std::vector<DevBuf<uint64_t>> uploads;
    if ((st = prove_batch(ctx, tr, "stage/0", 3, &inst, 1, so, uploads))) return st;

    memcpy(out.transcript_state, tr.state, 32);
    out.transcript_counter = tr.counter;
    out.witness_rows = rows;
    out.sumcheck_vars = n;
    out.n_tokens = total;
    std::vector<uint32_t> toks(total);
    TSR_HIP(hipMemcpyAsync(toks.data(), d_tokens.ptr, 4ull * total, hipMemcpyDeviceToHost, s));
    TSR_HIP(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < total && i < TSR_MAX_TOKENS; i++) out.tokens[i] = toks[i];
    out.final_a[0] = so.finals.size() > 0 ? so.finals[0].c0 : 0;
    out.final_a[1] = so.finals.size() > 0 ? so.finals[0].c1 : 0;
    return Status::ok();
}

// A request the pipeline can run: 1..512 tokens, the prompt at least one, every prompt token in
// the vocabulary (the decoder reads its embedding row without a further check).
static Status check_request(const tsr_request *req) {
    if (!req || !req->prompt) return Status::fail(TSR_BAD_ARG, "null request or prompt");
    const uint64_t total = (uint64_t)req->n_prompt + req->n_output;
    if (req->n_prompt == 0 || total > MAX_TOKENS)
        return Status::fail(TSR_BAD_SHAPE, "a request holds 1 to 512 tokens, the prompt at least one");
    for (uint32_t i = 0; i < req->n_prompt; i++)
        if (req->prompt[i] >= (uint32_t)VOCAB)
            return Status::fail(TSR_BAD_SHAPE, "a prompt token outside the vocabulary");
    return Status::ok();
}

static void worker(tsr_server *srv, uint32_t slot) {
    Ctx *ctx = (Ctx *)srv->ctxs[slot];
    (void)hipSetDevice(srv->device);
    for (;;) {
        Job job;
        uint64_t foot = 0;
        double t_admit = 0;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_work.wait(lk, [&] {
                if (srv->stop) return true;
                if (srv->queue.empty()) return false;
                const Job &front = srv->queue.front();
                uint64_t f = request_footprint(front.req.n_prompt + front.req.n_output);
                return srv->budget == 0 || srv->in_flight_bytes + f <= srv->budget ||
                       srv->active == 0;
            });
            if (srv->stop && srv->queue.empty()) return;
            job = std::move(srv->queue.front());
            srv->queue.pop_front();
            foot = request_footprint(job.req.n_prompt + job.req.n_output);
            srv->in_flight_bytes += foot;
            srv->active++;
            srv->max_active = std::max(srv->max_active, srv->active);
            t_admit = now_s(srv->t0);
        }
        tsr_result res{};
        res.id = job.req.id;
        res.t_submit = job.t_submit;
        res.t_admit = t_admit;
        res.slot = slot;
        job.req.prompt = job.prompt.data();
        Status st = prove_request(ctx, srv->weights, job.req, res);
        res.status = st.code;
        res.t_done = now_s(srv->t0);
        {
            std::lock_guard<std::mutex> lk(srv->mu);
            srv->in_flight_bytes -= foot;
            srv->active--;
            srv->done.push_back(res);
        }
        srv->cv_done.notify_all();
        srv->cv_work.notify_all();
    }
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_server_create(int device, const tsr_server_config *cfg, tsr_server **out) {
    if (!cfg || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (cfg->slots == 0 || cfg->slots > 64)
        return Status::fail(TSR_BAD_SHAPE, "slots must be 1..64").code;
    struct Cleanup {  // releases what a failed create made (no workers run yet)
        tsr_server *s;
        ~Cleanup() {
            if (!s) return;
            for (auto *c : s->ctxs) tsr_ctx_destroy(c);
            if (s->weights) tsr_free(s->owner, s->weights);
            if (s->owner) tsr_ctx_destroy(s->owner);
        }
    };
    std::unique_ptr<tsr_server> srv(new tsr_server());
    Cleanup guard{srv.get()};
    srv->t0 = std::chrono::steady_clock::now();
    srv->device = device;
    srv->slots = cfg->slots;
    srv->budget = cfg->hbm_budget;
    tsr_status s = tsr_ctx_create(device, &srv->owner);
    if (s) return s;
    void *w = nullptr;
    if ((s = tsr_alloc(srv->owner, MODEL_BYTES, &w))) return s;
    srv->weights = (int8_t *)w;
    Ctx *oc = (Ctx *)srv->owner;
    Status st = model_weights(oc, srv->weights, cfg->model_seed, oc->stream);
    if (st) return st.code;
    if ((s = tsr_ctx_synchronize(srv->owner))) return s;
    for (uint32_t i = 0; i < cfg->slots; i++) {
        tsr_ctx *c = nullptr;
        if ((s = tsr_ctx_create(device, &c))) return s;
        srv->ctxs.push_back(c);
    }
    for (uint32_t i = 0; i < cfg->slots; i++) srv->workers.emplace_back(worker, srv.get(), i);
    guard.s = nullptr;
    *out = srv.release();
    return TSR_OK;
}

tsr_status tsr_server_submit(tsr_server *srv, const tsr_request *req) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    if (Status st = check_request(req)) return st.code;
    Job job;
    job.req = *req;
    job.prompt.assign(req->prompt, req->prompt + req->n_prompt);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        job.t_submit = now_s(srv->t0);
        srv->queue.push_back(std::move(job));
    }
    srv->cv_work.notify_all();
    return TSR_OK;
}

tsr_status tsr_server_wait(tsr_server *srv, tsr_result *out, uint32_t timeout_ms) {
    if (!srv || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    std::unique_lock<std::mutex> lk(srv->mu);
    if (!srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [&] { return !srv->done.empty(); }))
        return Status::fail(TSR_DECLINED, "no result within the timeout").code;
    *out = srv->done.front();
    srv->done.pop_front();
    return TSR_OK;
}

tsr_status tsr_server_stats(tsr_server *srv, uint32_t *max_active, uint64_t *in_flight_bytes) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (max_active) *max_active = srv->max_active;
    if (in_flight_bytes) *in_flight_bytes = srv->in_flight_bytes;
    return TSR_OK;
}

tsr_status tsr_server_prove(tsr_server *srv, const tsr_request *req, tsr_result *out) {
    if (!srv || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (Status st = check_request(req)) return st.code;
    std::lock_guard<std::mutex> lk(srv->prove_mu);
    *out = tsr_result{};
    out->id = req->id;
    Ctx *ctx = (Ctx *)srv->owner;
    hipError_t e = hipSetDevice(srv->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    out->t_submit = out->t_admit = now_s(srv->t0);
    Status st = prove_request(ctx, srv->weights, *req, *out);
    out->t_done = now_s(srv->t0);
    out->status = st.code;
    return st.code;
}

uint64_t tsr_server_footprint(uint32_t tokens) { return request_footprint(tokens); }

double tsr_server_clock(tsr_server *srv) { return srv ? now_s(srv->t0) : 0.0; }

void tsr_server_destroy(tsr_server *srv) {
    if (!srv) return;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->stop = true;
    }
    srv->cv_work.notify_all();
    for (auto &w : srv->workers) w.join();
    for (auto *c : srv->ctxs) tsr_ctx_destroy(c);
    if (srv->weights) tsr_free(srv->owner, srv->weights);
    tsr_ctx_destroy(srv->owner);
    delete srv;
}

}  // extern "C"
