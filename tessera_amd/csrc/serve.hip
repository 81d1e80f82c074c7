// Served proving with continuous batching: requests of different lengths are admitted as slots
// and device-memory budget free up, decode together in shared launches, and are proved on
// overlapping streams.
//
// No reference counterpart (the reference proves one statement per call and has no serving
// loop). Each request's pipeline only uses the library's own kernels and entry points, so its
// result is the same alone and in any mix:
//   1. decode: the synthetic integer transformer (model.hip) generates the tokens greedily and
//      writes the witness, Goldilocks elements, WIT_STRIDE per token;
//   2. commit: the witness as rows of k = 2^13 values (zero-padded), rate 1/4, B = 32, unmasked
//      (tsr_commit_unmasked: encode, column chains, Merkle root);
//   3. transcript: SessionContext("tessera-protocol/1", NonInteractive, "test-small/1", zero
//      statement digest, session id = the request id LE), absorb the root under "serve/root",
//      squeeze n challenges "serve/r/{i}";
//   4. sumcheck: stage/0, degree 3, Sum_x eq(r, x) A(x) B(x) over layer 0's `a` and `m`
//      activations of every token (zero-padded to 2^n entries), on the device.
//
// Engine: one decoder thread owns a context (stream). Each iteration it admits waiting requests
// (strict FIFO: the oldest starts when fewer than `slots` requests are in flight and its
// footprint fits the budget beside those in flight), then launches one decode over every
// decoding request, each advancing by up to `chunk` positions (one workgroup per request), and
// waits for it. A request whose decode is complete goes to the prover threads (`provers`, each
// with its own context and stream), which commit and prove it while decoding continues. A
// request's buffers come from the decoder's pool when it is admitted and return to it when its
// proof is done; the KV cache returns as soon as its decode ends.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <map>
#include <mutex>
#include <set>
#include <stdint.h>
#include <string.h>
#include <string>
#include <thread>
#include <vector>
#include "commit.h"
#include "common.h"
#include "field.h"
#include "model.h"
#include "ntt.h"
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

static uint64_t witness_rows(uint32_t total) {
    const uint64_t k = 1ull << SERVE_LOG_K;
    return ((uint64_t)total * WIT_STRIDE + k - 1) / k;
}

// The device memory one request of `total` tokens needs (its admission footprint).
uint64_t request_footprint(uint32_t total) {
    const uint64_t k = 1ull << SERVE_LOG_K, rows = witness_rows(total), n = 4 * k;
    const uint64_t table = 1ull << ceil_log2((uint64_t)total * D_MODEL);
    return 8 * rows * k           // witness rows
           + 8 * rows * n         // codewords
           + 8 * rows * k         // encode scratch
           + 32 * 2 * n           // tree
           + 4 * kv_words(total)  // KV cache
           + 3 * 16 * table * 2;  // sumcheck tables and working buffers
}

// A request check: 1..512 tokens, the prompt at least one, every prompt token in the vocabulary
// (the decoder reads its embedding row without a further check).
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

// One request's device buffers, from the pool of the context that admitted it.
struct ReqBufs {
    DevBuf<uint32_t> prompt, tokens;
    DevBuf<int32_t> kv;
    DevBuf<uint64_t> wit;
    uint32_t total = 0;
    uint64_t rows = 0;

    // Allocates, uploads the prompt and zeroes the witness padding, on `ctx`'s stream.
    Status init(Ctx *ctx, const tsr_request &req) {
        total = req.n_prompt + req.n_output;
        rows = witness_rows(total);
        const uint64_t k = 1ull << SERVE_LOG_K, wit_len = (uint64_t)total * WIT_STRIDE;
        Status st;
        if ((st = prompt.alloc(ctx, req.n_prompt)) || (st = tokens.alloc(ctx, total)) ||
            (st = kv.alloc(ctx, kv_words(total))) || (st = wit.alloc(ctx, rows * k)))
            return st;
        TSR_HIP(hipMemcpyAsync(prompt.ptr, req.prompt, 4ull * req.n_prompt, hipMemcpyHostToDevice,
                               ctx->stream));
        const uint64_t pad = rows * k - wit_len;
        hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)(pad / 256 + 1)), dim3(256), 0, ctx->stream,
                           wit.ptr + wit_len, pad);
        return check_launch("serve buffers");
    }

    DecodeTask task(const tsr_request &req, uint32_t start, uint32_t end) const {
        return DecodeTask{prompt.ptr, tokens.ptr, kv.ptr, wit.ptr, req.n_prompt, total, start, end};
    }
};

// The witness commitment's parameters: rows of 2^13, rate 1/4, B = 32, unmasked.
constexpr uint32_t SERVE_LOG_R = 2, SERVE_BLOCK = 32;
constexpr uint8_t SERVE_COMMITMENT_ID_BYTE = 0x5e;  // commitment_id = 32 bytes of 0x5e

static tsr_commitment_params serve_params() {
    tsr_commitment_params params{};
    params.log2_k = SERVE_LOG_K;
    params.log2_inv_rate = SERVE_LOG_R;
    params.mask_coeffs = 0;
    params.row_block = SERVE_BLOCK;
    memset(params.commitment_id, SERVE_COMMITMENT_ID_BYTE, 32);
    return params;
}

// A request's witness commitment, freed with tsr_commitment_free.
struct CommitmentFree {
    void operator()(tsr_commitment *c) const { tsr_commitment_free(c); }
};
using Commitment = std::unique_ptr<tsr_commitment, CommitmentFree>;

// Step 2 of one decoded request: tsr_commit_unmasked of its witness rows.
static Status commit_one(Ctx *ctx, const ReqBufs &b, Commitment &out) {
    const tsr_commitment_params params = serve_params();
    tsr_commitment *c = nullptr;
    tsr_status cs = tsr_commit_unmasked((tsr_ctx *)ctx, &params, b.wit.ptr, b.rows, &c);
    if (cs != TSR_OK) return Status::fail(cs, std::string("serve commit: ") + tsr_last_error());
    out.reset(c);
    return Status::ok();
}

// Step 2 of several decoded requests together, each root equal to its commit_one root. Every
// witness row of the group goes through one launch per NTT pass (per-row pointer tables: every
// witness has rows of 2^13, so one encoder serves all), one launch chains every column of every
// request, one launch per Merkle level hashes that level of every tree, and the roots come back
// in one copy. The small per-request commitments thus share their launches instead of paying
// a launch sequence and a synchronization each.
static Status commit_group(Ctx *ctx, const std::vector<const ReqBufs *> &reqs, std::vector<Commitment> &out) {
    hipStream_t s = ctx->stream;
    const uint32_t count = (uint32_t)reqs.size();
    const uint64_t k = 1ull << SERVE_LOG_K, n = k << SERVE_LOG_R;
    const uint32_t levels = SERVE_LOG_K + SERVE_LOG_R;
    Status st;
    Encoder *enc = nullptr;
    if ((st = get_encoder(ctx, SERVE_LOG_K, SERVE_LOG_R, 0, &enc))) return st;
    std::vector<DevBuf<uint64_t>> cw(count), tmp(count);
    std::vector<DevBuf<uint32_t>> nodes(count);
    std::vector<MatDesc> mats(count);
    std::vector<uint64_t> src_p, tmp_p, cw_p;
    std::vector<BlockRef> refs;  // the chains' blocks, then each Merkle level's
    std::vector<uint32_t> first(levels + 2, 0);
    for (uint32_t i = 0; i < count; i++) {
        const ReqBufs &b = *reqs[i];
        if ((st = cw[i].alloc(ctx, b.rows * n)) || (st = tmp[i].alloc(ctx, b.rows * k)) ||
            (st = nodes[i].alloc(ctx, 8 * (2 * n - 1))))
            return st;
        MatDesc &m = mats[i];
        m.cw = cw[i].ptr;
        m.n = n;
        m.rows = b.rows;
        m.block = SERVE_BLOCK;
        memset(m.id, SERVE_COMMITMENT_ID_BYTE, 32);
        m.nodes = nodes[i].ptr;
        for (uint64_t r = 0; r < b.rows; r++) {
            src_p.push_back((uint64_t)(b.wit.ptr + r * k));
            tmp_p.push_back((uint64_t)(tmp[i].ptr + r * k));
            cw_p.push_back((uint64_t)(cw[i].ptr + r * n));
        }
    }
    for (uint32_t l = 0; l <= levels; l++) {  // l = 0: chains over n columns; l >= 1: n >> l parents
        first[l] = (uint32_t)refs.size();
        for (uint32_t i = 0; i < count; i++)
            for (uint64_t f = 0; f < (n >> l); f += 256) refs.push_back(BlockRef{i, 0, f});
    }
    first[levels + 1] = (uint32_t)refs.size();
    const uint64_t rows = src_p.size();
    // one upload: [mats | src | tmp | cw row tables | block refs], then the roots come back
    size_t at = 0;
    auto place = [&](size_t bytes) {
        at = (at + 255) & ~(size_t)255;
        const size_t o = at;
        at += bytes;
        return o;
    };
    const size_t o_mats = place(sizeof(MatDesc) * count), o_src = place(8 * rows),
                 o_tmp = place(8 * rows), o_cw = place(8 * rows), o_refs = place(sizeof(BlockRef) * refs.size());
    const size_t up_end = at, o_roots = place(32 * count);
    DevBuf<uint8_t> blob;
    if ((st = blob.alloc(ctx, at))) return st;
    uint8_t *host;
    if ((st = pinned_host(ctx, at, &host))) return st;
    memcpy(host + o_mats, mats.data(), sizeof(MatDesc) * count);
    memcpy(host + o_src, src_p.data(), 8 * rows);
    memcpy(host + o_tmp, tmp_p.data(), 8 * rows);
    memcpy(host + o_cw, cw_p.data(), 8 * rows);
    memcpy(host + o_refs, refs.data(), sizeof(BlockRef) * refs.size());
    TSR_HIP(hipMemcpyAsync(blob.ptr, host, up_end, hipMemcpyHostToDevice, s));
    const MatDesc *d_mats = (const MatDesc *)(blob.ptr + o_mats);
    const BlockRef *d_refs = (const BlockRef *)(blob.ptr + o_refs);
    if ((st = enc->encode_rowmap((const uint64_t *const *)(blob.ptr + o_src), (uint64_t *const *)(blob.ptr + o_tmp),
                                 (uint64_t *const *)(blob.ptr + o_cw), rows, s)))
        return st;
    launch_chain_leaves_multi(ctx, d_mats, d_refs + first[0], first[1] - first[0], s);
    for (uint32_t l = 1; l <= levels; l++)
        launch_merkle_level_multi(ctx, d_mats, d_refs + first[l], first[l + 1] - first[l], l, s);
    launch_gather_roots(d_mats, count, (uint32_t *)(blob.ptr + o_roots), s);
    if ((st = check_launch("serve group commit"))) return st;
    // the readback reuses the pinned block: the stream has finished the upload before it
    TSR_HIP(hipMemcpyAsync(host, blob.ptr + o_roots, 32 * count, hipMemcpyDeviceToHost, s));
    TSR_HIP(tsr::stream_sync(s));
    // each request's codewords and tree become its commitment, as tsr_commit_unmasked's
    out.clear();
    for (uint32_t i = 0; i < count; i++) {
        Commitment c(new tsr_commitment());
        c->ctx = ctx;
        c->params = serve_params();
        c->count = reqs[i]->rows;
        c->n = n;
        c->codewords = std::move(cw[i]);
        c->nodes = std::move(nodes[i]);
        memcpy(c->root, host + 32 * (size_t)i, 32);
        out.push_back(std::move(c));
    }
    return Status::ok();
}

// Steps 3-4 of a decoded request whose witness root is out.root, on `ctx` (the witness and
// tokens may belong to another context's pool: they are only read here). Fills `out` except the
// timestamps.
static Status prove_committed(Ctx *ctx, const tsr_request &req, const ReqBufs &b, tsr_result &out) {
    hipStream_t s = ctx->stream;
    const uint32_t total = b.total;
    Status st;

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

    // tables: eq(r, .) given by its point (never built), layer 0's a and m of every token
    const uint64_t len = 1ull << n;
    DevBuf<uint64_t> A, B;
    if ((st = A.alloc(ctx, len)) || (st = B.alloc(ctx, len))) return st;
    hipLaunchKernelGGL(k_gather_base, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, b.wit.ptr,
                       total, WIT_A0, len, A.ptr);
    hipLaunchKernelGGL(k_gather_base, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, s, b.wit.ptr,
                       total, WIT_M0, len, B.ptr);
    if ((st = check_launch("serve tables"))) return st;
    tsr_table tabs[3] = {{(const uint64_t *)r.data(), TSR_TABLE_EQ, 1}, {A.ptr, 0, 0}, {B.ptr, 0, 0}};
    tsr_instance inst{n, 3, tabs};
    SumcheckOut so;
    std::vector<DevBuf<uint64_t>> uploads;
    if ((st = prove_batch(ctx, tr, "stage/0", 3, &inst, 1, so, uploads))) return st;

    memcpy(out.transcript_state, tr.state, 32);
    out.transcript_counter = tr.counter;
    out.witness_rows = b.rows;
    out.sumcheck_vars = n;
    out.n_tokens = total;
    std::vector<uint32_t> toks(total);
    TSR_HIP(hipMemcpyAsync(toks.data(), b.tokens.ptr, 4ull * total, hipMemcpyDeviceToHost, s));
    TSR_HIP(tsr::stream_sync(s));
    for (uint32_t i = 0; i < total && i < TSR_MAX_TOKENS; i++) out.tokens[i] = toks[i];
    out.final_a[0] = so.finals.size() > 0 ? so.finals[0].c0 : 0;
    out.final_a[1] = so.finals.size() > 0 ? so.finals[0].c1 : 0;
    return Status::ok();
}

// Steps 2-4 of a decoded request on its own; its commitment into `com`.
static Status prove_decoded(Ctx *ctx, const tsr_request &req, const ReqBufs &b, tsr_result &out,
                            Commitment &com) {
    Status st = commit_one(ctx, b, com);
    if (st) return st;
    memcpy(out.root, com->root, 32);
    return prove_committed(ctx, req, b, out);
}

// A proved request's witness commitment (codewords and tree) kept on the device for openings
// until it is released; `bytes` is what it holds of the admission budget.
struct Resident {
    Commitment c;
    uint64_t bytes = 0;
};

// Where a request is, for tsr_server_cancel.
enum JobState { JOB_QUEUED, JOB_DECODING, JOB_DECODED, JOB_PROVING };

struct Job {
    tsr_request req;
    std::vector<uint32_t> prompt;
    double t_submit = 0, t_admit = 0;
    uint64_t foot = 0;
    uint32_t pos = 0;  // decoded positions
    ReqBufs bufs;
    Status status;
};

}  // namespace tsr

struct tsr_server {
    std::chrono::steady_clock::time_point t0;
    int device = 0;
    uint32_t slots = 1, provers = 1, chunk = 32, group = 1;
    uint64_t budget = 0, in_flight_bytes = 0;
    tsr_ctx *owner = nullptr;  // holds the weights; runs tsr_server_prove
    tsr_ctx *dec = nullptr;    // the decoder's context: its pool holds the requests' buffers
    int8_t *weights = nullptr;
    std::vector<tsr_ctx *> prover_ctxs;
    std::thread decoder;
    std::vector<std::thread> workers;
    std::mutex mu;
    std::mutex prove_mu;  // tsr_server_prove calls share the owner context: one at a time
    std::condition_variable cv_work, cv_prove, cv_done;
    std::deque<std::unique_ptr<tsr::Job>> queue;    // waiting for admission
    std::deque<std::unique_ptr<tsr::Job>> decoded;  // waiting for a prover
    std::deque<tsr_result> done;
    std::map<uint64_t, int> state;  // JobState of every request in flight, by id
    std::set<uint64_t> cancel;      // decoding requests to end after the current decode launch
    bool keep = false;              // resident commitments (tsr_server_open, tsr_server_release)
    std::map<uint64_t, std::shared_ptr<tsr::Resident>> resident;
    uint64_t resident_bytes = 0;
    bool stop = false, decoder_done = false;
    uint32_t in_flight = 0, max_active = 0;
};

namespace tsr {

// Ends a job: its result goes to the done queue and its budget and slot are released. A
// commitment kept resident holds its own bytes of the budget instead, until released.
static void finish(tsr_server *srv, Job &job, const tsr_result *filled, Commitment keep = nullptr) {
    tsr_result res{};
    if (filled) res = *filled;
    res.id = job.req.id;
    res.status = job.status.code;
    res.t_submit = job.t_submit;
    res.t_admit = job.t_admit;
    res.t_done = now_s(srv->t0);
    job.bufs = ReqBufs();  // buffers back to the decoder's pool
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->in_flight_bytes -= job.foot;
        srv->in_flight--;
        if (keep && res.status == TSR_OK) {
            auto r = std::make_shared<Resident>();
            r->bytes = 8 * keep->codewords.len + 4 * keep->nodes.len;
            // openings run on the server's own context (tsr_server_open); the buffers go back to
            // the prover's pool when released
            keep->ctx = (Ctx *)srv->owner;
            r->c = std::move(keep);
            srv->resident_bytes += r->bytes;
            srv->resident[res.id] = std::move(r);
        }
        srv->state.erase(job.req.id);
        srv->cancel.erase(job.req.id);
        srv->done.push_back(res);
    }
    srv->cv_done.notify_all();
    srv->cv_work.notify_all();
}

static void decoder_loop(tsr_server *srv) {
    Ctx *ctx = (Ctx *)srv->dec;
    (void)hipSetDevice(srv->device);
    std::vector<std::unique_ptr<Job>> active;
    std::vector<DecodeTask> host_tasks;
    DevBuf<DecodeTask> dev_tasks;
    const Status tasks_st = dev_tasks.alloc(ctx, srv->slots);
    for (;;) {
        std::vector<std::unique_ptr<Job>> admitted;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            auto admissible = [&] {
                if (srv->queue.empty() || srv->in_flight >= srv->slots) return false;
                const uint64_t f = srv->queue.front()->foot;
                return srv->budget == 0 || srv->in_flight_bytes + srv->resident_bytes + f <= srv->budget ||
                       srv->in_flight == 0;
            };
            if (active.empty())
                srv->cv_work.wait(lk, [&] { return admissible() || (srv->stop && srv->queue.empty()); });
            while (admissible()) {
                std::unique_ptr<Job> job = std::move(srv->queue.front());
                srv->queue.pop_front();
                srv->in_flight_bytes += job->foot;
                srv->in_flight++;
                srv->max_active = std::max(srv->max_active, srv->in_flight);
                job->t_admit = now_s(srv->t0);
                srv->state[job->req.id] = JOB_DECODING;
                admitted.push_back(std::move(job));
            }
            if (active.empty() && admitted.empty() && srv->stop && srv->queue.empty()) break;
        }
        for (auto &job : admitted) {
            job->req.prompt = job->prompt.data();
            job->status = tasks_st ? tasks_st : job->bufs.init(ctx, job->req);
            if (job->status) {
                finish(srv, *job, nullptr);
                continue;
            }
            active.push_back(std::move(job));
        }
        if (active.empty()) continue;
        host_tasks.clear();
        for (auto &job : active) {
            const uint32_t end = std::min(job->pos + srv->chunk, job->bufs.total);
            host_tasks.push_back(job->bufs.task(job->req, job->pos, end));
        }
        Status ls;
        hipError_t e = hipMemcpyAsync(dev_tasks.ptr, host_tasks.data(), sizeof(DecodeTask) * host_tasks.size(),
                                      hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) ls = hip_fail(e, "decode tasks");
        if (!ls) ls = model_decode(ctx, srv->weights, dev_tasks.ptr, (uint32_t)host_tasks.size(), ctx->stream);
        if (!ls && (e = tsr::stream_sync(ctx->stream)) != hipSuccess) ls = hip_fail(e, "decode");
        std::vector<std::unique_ptr<Job>> still;
        for (size_t i = 0; i < active.size(); i++) {
            std::unique_ptr<Job> &job = active[i];
            job->pos = host_tasks[i].end;
            bool cancelled;
            {
                std::lock_guard<std::mutex> lk(srv->mu);
                cancelled = srv->cancel.count(job->req.id) != 0;
            }
            if (ls || cancelled) {
                job->status = ls ? ls : Status::fail(TSR_CANCELLED, "cancelled");
                finish(srv, *job, nullptr);
            } else if (job->pos == job->bufs.total) {
                job->bufs.kv = DevBuf<int32_t>();  // the KV cache is done with
                {
                    std::lock_guard<std::mutex> lk(srv->mu);
                    srv->state[job->req.id] = JOB_DECODED;
                    srv->decoded.push_back(std::move(job));
                }
                srv->cv_prove.notify_one();
            } else {
                still.push_back(std::move(job));
            }
        }
        active.swap(still);
    }
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        srv->decoder_done = true;
    }
    srv->cv_prove.notify_all();
}

// A prover takes the decoded requests waiting (up to srv->group, oldest first), commits them
// together (commit_group) and proves them one after another on its stream. A group whose
// commitment fails (for memory, say) commits its requests one at a time instead. The default
// group is 1: served witnesses are about 1.4 rows of 2^13 per token, so their commitments keep
// the GPU busy on their own, and a group's requests would hold their slots until the whole
// group's commitment is done (measured: capacity 2 240 proofs/s one by one, 2 050 in groups of
// up to 16, the same latency at 400 requests/s within the runs' spread).
static void prover_loop(tsr_server *srv, uint32_t slot) {
    Ctx *ctx = (Ctx *)srv->prover_ctxs[slot];
    (void)hipSetDevice(srv->device);
    for (;;) {
        std::vector<std::unique_ptr<Job>> jobs;
        {
            std::unique_lock<std::mutex> lk(srv->mu);
            srv->cv_prove.wait(lk, [&] { return !srv->decoded.empty() || srv->decoder_done; });
            if (srv->decoded.empty()) return;
            while (!srv->decoded.empty() && jobs.size() < srv->group) {
                jobs.push_back(std::move(srv->decoded.front()));
                srv->decoded.pop_front();
                srv->state[jobs.back()->req.id] = JOB_PROVING;
            }
        }
        const uint32_t g = (uint32_t)jobs.size();
        std::vector<tsr_result> res(g);
        std::vector<Commitment> coms;
        bool grouped = false;
        if (g > 1) {
            std::vector<const ReqBufs *> reqs;
            for (auto &job : jobs) reqs.push_back(&job->bufs);
            grouped = !commit_group(ctx, reqs, coms);
            (void)tsr::stream_sync(ctx->stream);  // (after an error) nothing still runs
            if (!grouped) coms.clear();
        }
        coms.resize(g);
        for (uint32_t i = 0; i < g; i++) {
            Job &job = *jobs[i];
            res[i].slot = slot;
            res[i].group = grouped ? g : 1;
            if (grouped) {
                memcpy(res[i].root, coms[i]->root, 32);
                job.status = prove_committed(ctx, job.req, job.bufs, res[i]);
            } else {
                job.status = prove_decoded(ctx, job.req, job.bufs, res[i], coms[i]);
            }
            (void)tsr::stream_sync(ctx->stream);  // (after an error) nothing still reads the buffers
            finish(srv, job, &res[i], srv->keep ? std::move(coms[i]) : Commitment());
            coms[i].reset();
        }
    }
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_server_create(int device, const tsr_server_config *cfg, tsr_server **out) {
    if (!cfg || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (cfg->slots == 0 || cfg->slots > 1024)
        return Status::fail(TSR_BAD_SHAPE, "slots must be 1..1024").code;
    if (cfg->provers > 64) return Status::fail(TSR_BAD_SHAPE, "provers must be 0..64").code;
    if (cfg->chunk > MAX_TOKENS) return Status::fail(TSR_BAD_SHAPE, "chunk must be 0..512").code;
    if (cfg->group > 1024) return Status::fail(TSR_BAD_SHAPE, "group must be 0..1024").code;
    struct Cleanup {  // releases what a failed create made (no threads run yet)
        tsr_server *s;
        ~Cleanup() {
            if (!s) return;
            for (auto *c : s->prover_ctxs) tsr_ctx_destroy(c);
            if (s->dec) tsr_ctx_destroy(s->dec);
            if (s->weights) tsr_free(s->owner, s->weights);
            if (s->owner) tsr_ctx_destroy(s->owner);
        }
    };
    std::unique_ptr<tsr_server> srv(new tsr_server());
    Cleanup guard{srv.get()};
    srv->t0 = std::chrono::steady_clock::now();
    srv->device = device;
    srv->slots = cfg->slots;
    srv->provers = cfg->provers ? cfg->provers : 4;
    srv->chunk = cfg->chunk ? cfg->chunk : 32;
    srv->group = cfg->group ? cfg->group : 1;
    srv->keep = cfg->resident != 0;
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
    if ((s = tsr_ctx_create(device, &srv->dec))) return s;
    for (uint32_t i = 0; i < srv->provers; i++) {
        tsr_ctx *c = nullptr;
        if ((s = tsr_ctx_create(device, &c))) return s;
        srv->prover_ctxs.push_back(c);
    }
    srv->decoder = std::thread(decoder_loop, srv.get());
    for (uint32_t i = 0; i < srv->provers; i++) srv->workers.emplace_back(prover_loop, srv.get(), i);
    guard.s = nullptr;
    *out = srv.release();
    return TSR_OK;
}

tsr_status tsr_server_submit(tsr_server *srv, const tsr_request *req) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    if (Status st = check_request(req)) return st.code;
    std::unique_ptr<Job> job(new Job());
    job->req = *req;
    job->prompt.assign(req->prompt, req->prompt + req->n_prompt);
    job->foot = request_footprint(req->n_prompt + req->n_output);
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        if (srv->state.count(req->id))
            return Status::fail(TSR_BAD_ARG, "a request with id " + std::to_string(req->id) + " is in flight").code;
        if (srv->resident.count(req->id))
            return Status::fail(TSR_BAD_ARG, "request " + std::to_string(req->id) +
                                                 "'s commitment is resident: release it first").code;
        job->t_submit = now_s(srv->t0);
        srv->state[req->id] = JOB_QUEUED;
        srv->queue.push_back(std::move(job));
    }
    srv->cv_work.notify_all();
    return TSR_OK;
}

tsr_status tsr_server_cancel(tsr_server *srv, uint64_t id) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    std::unique_ptr<Job> job;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->state.find(id);
        if (it == srv->state.end())
            return Status::fail(TSR_DECLINED, "no request " + std::to_string(id) + " in flight").code;
        switch (it->second) {
            case JOB_QUEUED: {
                // never admitted: no slot, budget or buffer to give back
                for (auto q = srv->queue.begin(); q != srv->queue.end(); ++q) {
                    if ((*q)->req.id != id) continue;
                    tsr_result res{};
                    res.id = id;
                    res.status = TSR_CANCELLED;
                    res.t_submit = (*q)->t_submit;
                    res.t_done = now_s(srv->t0);
                    srv->queue.erase(q);
                    srv->state.erase(it);
                    srv->done.push_back(res);
                    break;
                }
                srv->cv_done.notify_all();
                return TSR_OK;
            }
            case JOB_DECODING:
                srv->cancel.insert(id);  // the decoder ends it after its current launch
                return TSR_OK;
            case JOB_DECODED:
                for (auto q = srv->decoded.begin(); q != srv->decoded.end(); ++q) {
                    if ((*q)->req.id != id) continue;
                    job = std::move(*q);
                    srv->decoded.erase(q);
                    break;
                }
                break;
            default:
                return Status::fail(TSR_DECLINED, "request " + std::to_string(id) + " is being proved").code;
        }
    }
    if (!job) return Status::fail(TSR_DECLINED, "request " + std::to_string(id) + " is not waiting").code;
    job->status = Status::fail(TSR_CANCELLED, "cancelled");
    finish(srv, *job, nullptr);
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
    ReqBufs b;
    Status st = b.init(ctx, *req);
    DevBuf<DecodeTask> task;
    if (!st) st = task.alloc(ctx, 1);
    if (!st) {
        const DecodeTask t = b.task(*req, 0, b.total);
        e = hipMemcpyAsync(task.ptr, &t, sizeof t, hipMemcpyHostToDevice, ctx->stream);
        if (e != hipSuccess) st = hip_fail(e, "decode task");
    }
    if (!st) st = model_decode(ctx, srv->weights, task.ptr, 1, ctx->stream);
    if (!st) {
        b.kv = DevBuf<int32_t>();
        Commitment com;
        st = prove_decoded(ctx, *req, b, *out, com);
    }
    out->t_done = now_s(srv->t0);
    out->status = st.code;
    return st.code;
}

tsr_status tsr_server_open(tsr_server *srv, uint64_t id, const uint64_t *queries, uint64_t q,
                           uint64_t *distinct_out, uint64_t distinct_cap, uint64_t *columns_out,
                           uint8_t *siblings_out, uint64_t siblings_cap, uint64_t *n_distinct,
                           uint64_t *n_siblings) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    std::shared_ptr<Resident> r;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->resident.find(id);
        if (it == srv->resident.end())
            return Status::fail(TSR_DECLINED, "no resident commitment for request " + std::to_string(id)).code;
        r = it->second;  // kept alive through the call even if released meanwhile
    }
    // on the server's own context, which tsr_server_prove calls share
    std::lock_guard<std::mutex> lk(srv->prove_mu);
    hipError_t e = hipSetDevice(srv->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    return tsr_commitment_answer(r->c.get(), queries, q, distinct_out, distinct_cap, columns_out,
                                 siblings_out, siblings_cap, n_distinct, n_siblings);
}

tsr_status tsr_server_release(tsr_server *srv, uint64_t id) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    std::shared_ptr<Resident> r;
    {
        std::lock_guard<std::mutex> lk(srv->mu);
        auto it = srv->resident.find(id);
        if (it == srv->resident.end())
            return Status::fail(TSR_DECLINED, "no resident commitment for request " + std::to_string(id)).code;
        r = std::move(it->second);
        srv->resident.erase(it);
        srv->resident_bytes -= r->bytes;
    }
    srv->cv_work.notify_all();  // its bytes may let a waiting request in
    r.reset();                  // the last holder frees the buffers
    return TSR_OK;
}

tsr_status tsr_server_resident(tsr_server *srv, uint32_t *count, uint64_t *bytes) {
    if (!srv) return Status::fail(TSR_BAD_ARG, "null server").code;
    std::lock_guard<std::mutex> lk(srv->mu);
    if (count) *count = (uint32_t)srv->resident.size();
    if (bytes) *bytes = srv->resident_bytes;
    return TSR_OK;
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
    if (srv->decoder.joinable()) srv->decoder.join();
    for (auto &w : srv->workers) w.join();
    srv->resident.clear();  // back to the provers' pools before the contexts go
    for (auto *c : srv->prover_ctxs) tsr_ctx_destroy(c);
    tsr_ctx_destroy(srv->dec);
    if (srv->weights) tsr_free(srv->owner, srv->weights);
    tsr_ctx_destroy(srv->owner);
    delete srv;
}

}  // extern "C"
