// Batched commitments: many unmasked matrices of different shapes encoded and Merkle-committed
// together, each with exactly the root ReferenceCommitment::commit_unmasked gives it
// (crates/tessera-pcs/src/commitment.rs:343-446).
//
// A batch is planned once: matrices are grouped by (log k, rate) and each group's rows are
// encoded by one launch per NTT pass through per-row pointer tables, whatever matrix a row
// belongs to; one launch chains every column of every matrix; one launch per Merkle level
// hashes that level of every matrix still that tall. The tile size of each group's passes is
// chosen by timing the candidates on the batch's own rows (the autotuner), and the whole launch
// sequence is captured once into a hipGraph that every later run replays.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <map>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "commit.h"
#include "common.h"
#include "ntt.h"

struct tsr_batch {
    tsr::Ctx *ctx = nullptr;
    uint32_t count = 0;
    uint32_t flags = 0;
    std::vector<tsr_commit_desc> descs;
    // per matrix
    std::vector<tsr::DevBuf<uint64_t>> rows_up, codewords, scratch;
    std::vector<tsr::DevBuf<uint32_t>> nodes;
    std::vector<tsr::MatDesc> hmats;
    tsr::DevBuf<tsr::MatDesc> mats;
    // per (log k, rate) group
    struct Group {
        uint32_t log_k, log_r, tile_log, max_nb = 0;
        uint64_t rows;
        double ms = 0;  // the chosen plan's time when tuned
        tsr::DevBuf<uint64_t> src, tmp, cw;  // row pointer tables
        tsr::Encoder *enc = nullptr;
        hipStream_t side = nullptr;  // the group's branch (parallel batches)
        hipEvent_t done = nullptr;
    };
    hipEvent_t fork = nullptr;
    std::vector<std::unique_ptr<Group>> groups;
    tsr::DevBuf<tsr::BlockRef> chain_refs;
    uint32_t chain_blocks = 0;
    std::vector<tsr::DevBuf<tsr::BlockRef>> level_refs;
    std::vector<uint32_t> level_blocks;
    tsr::DevBuf<uint32_t> roots;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<double> tune_ms;  // per group, per candidate (flattened) when autotuned
};

namespace tsr {

static Status upload_ptrs(Ctx *ctx, DevBuf<uint64_t> &buf, const std::vector<uint64_t> &v) {
    return buf.upload(ctx, v);
}

// Enqueues the whole batch on the stream (no synchronization). The shape groups' encodes are
// independent: unless the batch is serial, each runs on a branch of its own (a side stream
// forked from and joined back to `s`, so a captured graph holds them as parallel branches), and
// small-k groups overlap the large ones. Chains and Merkle levels follow on `s`.
static Status enqueue(tsr_batch *b, hipStream_t s) {
    const bool par = !(b->flags & TSR_BATCH_SERIAL) && b->groups.size() > 1;
    if (par) TSR_HIP(hipEventRecord(b->fork, s));
    for (auto &g : b->groups) {
        hipStream_t gs = par ? g->side : s;
        if (par) TSR_HIP(hipStreamWaitEvent(gs, b->fork, 0));
        Status st = g->enc->encode_rowmap((const uint64_t *const *)g->src.ptr,
                                          (uint64_t *const *)g->tmp.ptr,
                                          (uint64_t *const *)g->cw.ptr, g->rows, gs);
        if (st) return st;
        if (par) TSR_HIP(hipEventRecord(g->done, gs));
    }
    if (par)
        for (auto &g : b->groups) TSR_HIP(hipStreamWaitEvent(s, g->done, 0));
    launch_chain_leaves_multi(b->ctx, b->mats.ptr, b->chain_refs.ptr, b->chain_blocks, s);
    for (size_t l = 0; l < b->level_refs.size(); l++)
        launch_merkle_level_multi(b->ctx, b->mats.ptr, b->level_refs[l].ptr, b->level_blocks[l],
                                  (uint32_t)(l + 1), s);
    launch_gather_roots(b->mats.ptr, b->count, b->roots.ptr, s);
    return check_launch("commit batch");
}

static Status time_group(tsr_batch *b, tsr_batch::Group &g, double *ms) {
    hipStream_t s = b->ctx->stream;
    hipEvent_t e0, e1;
    TSR_HIP(hipEventCreate(&e0));
    TSR_HIP(hipEventCreate(&e1));
    Status st = g.enc->encode_rowmap((const uint64_t *const *)g.src.ptr, (uint64_t *const *)g.tmp.ptr,
                                     (uint64_t *const *)g.cw.ptr, g.rows, s);  // warm-up
    TSR_HIP(hipEventRecord(e0, s));
    for (int i = 0; i < 2 && !st; i++)
        st = g.enc->encode_rowmap((const uint64_t *const *)g.src.ptr, (uint64_t *const *)g.tmp.ptr,
                                  (uint64_t *const *)g.cw.ptr, g.rows, s);
    TSR_HIP(hipEventRecord(e1, s));
    TSR_HIP(hipEventSynchronize(e1));
    float t = 0;
    TSR_HIP(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = t / 2;
    return st;
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_commit_batch_create(tsr_ctx *ctx_, const tsr_commit_desc *descs, uint32_t count,
                                   uint32_t flags, tsr_batch **out) {
    Ctx *ctx = (Ctx *)ctx_;
    if (!ctx || !out || (count && !descs)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (count == 0) return Status::fail(TSR_BAD_SHAPE, "an empty batch").code;
    for (uint32_t i = 0; i < count; i++) {
        const tsr_commit_desc &d = descs[i];
        if (d.log2_inv_rate == 0)
            return Status::fail(TSR_BAD_SHAPE, "InverseRate: matrix " + std::to_string(i)).code;
        if ((uint64_t)d.log2_k + d.log2_inv_rate > 32)
            return Status::fail(TSR_BAD_SHAPE, "TooLong: matrix " + std::to_string(i)).code;
        if (d.row_block == 0 || d.row_block > 124)
            return Status::fail(d.row_block ? TSR_DECLINED : TSR_BAD_SHAPE,
                                "row block of matrix " + std::to_string(i)).code;
        if (d.count && !d.rows) return Status::fail(TSR_BAD_ARG, "null rows").code;
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    hipStream_t s = ctx->stream;
    std::unique_ptr<tsr_batch> b(new tsr_batch());
    b->ctx = ctx;
    b->count = count;
    b->flags = flags;
    b->descs.assign(descs, descs + count);
    b->rows_up.resize(count);
    b->codewords.resize(count);
    b->scratch.resize(count);
    b->nodes.resize(count);
    b->hmats.resize(count);
    Status st;
    // per matrix buffers
    std::map<std::pair<uint32_t, uint32_t>, std::vector<uint32_t>> by_shape;
    for (uint32_t i = 0; i < count; i++) {
        const tsr_commit_desc &d = descs[i];
        const uint64_t k = 1ull << d.log2_k, n = k << d.log2_inv_rate;
        const uint64_t *rows = d.rows;
        if (d.count && !is_device_ptr(d.rows)) {
            if ((st = b->rows_up[i].alloc(ctx, d.count * k))) return st.code;
            e = hipMemcpyAsync(b->rows_up[i].ptr, d.rows, 8 * d.count * k, hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return hip_fail(e, "upload").code;
            rows = b->rows_up[i].ptr;
        }
        b->descs[i].rows = rows;
        if ((st = b->codewords[i].alloc(ctx, d.count * n)) || (st = b->scratch[i].alloc(ctx, d.count * k)) ||
            (st = b->nodes[i].alloc(ctx, 8 * (2 * n - 1))))
            return st.code;
        MatDesc &m = b->hmats[i];
        m.cw = b->codewords[i].ptr;
        m.n = n;
        m.rows = d.count;
        m.block = (uint32_t)d.row_block;
        memcpy(m.id, d.commitment_id, 32);
        m.nodes = b->nodes[i].ptr;
        if (d.count) by_shape[{d.log2_k, d.log2_inv_rate}].push_back(i);
    }
    {
        std::vector<MatDesc> v = b->hmats;
        if ((st = b->mats.alloc(ctx, count))) return st.code;
        e = hipMemcpyAsync(b->mats.ptr, v.data(), sizeof(MatDesc) * count, hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "upload").code;
    }
    // groups and their row tables
    for (auto &kv : by_shape) {
        std::unique_ptr<tsr_batch::Group> g(new tsr_batch::Group());
        g->log_k = kv.first.first;
        g->log_r = kv.first.second;
        const uint64_t k = 1ull << g->log_k, n = k << g->log_r;
        std::vector<uint64_t> src, tmp, cw;
        for (uint32_t i : kv.second) {
            for (uint64_t r = 0; r < descs[i].count; r++) {
                src.push_back((uint64_t)(b->descs[i].rows + r * k));
                tmp.push_back((uint64_t)(b->scratch[i].ptr + r * k));
                cw.push_back((uint64_t)(b->codewords[i].ptr + r * n));
            }
        }
        g->rows = src.size();
        if ((st = upload_ptrs(ctx, g->src, src)) || (st = upload_ptrs(ctx, g->tmp, tmp)) ||
            (st = upload_ptrs(ctx, g->cw, cw)))
            return st.code;
        g->tile_log = 13;
        if ((st = get_encoder(ctx, g->log_k, g->log_r, 0, &g->enc, 13))) return st.code;
        if (flags & TSR_BATCH_PLAN) {
            g->tile_log = (flags >> 8) & 0xFF;
            g->max_nb = (flags >> 16) & 0xFF;
            if (g->tile_log != 12 && g->tile_log != 13)
                return Status::fail(TSR_BAD_ARG, "a plan's tile_log is 12 or 13").code;
            if ((st = get_encoder(ctx, g->log_k, g->log_r, 0, &g->enc, g->tile_log, g->max_nb))) return st.code;
        } else if (flags & TSR_BATCH_AUTOTUNE) {
            // plans: the LDS tile (2^13 or 2^12 elements) and the most stage bits of a strided
            // pass (the pass split), timed on the group's own rows; the choice is kept in the
            // context for later batches of the same shape and height
            const auto key = std::make_tuple(g->log_k, g->log_r, (uint32_t)(63 - __builtin_clzll(g->rows)));
            auto hit = ctx->batch_tunings.find(key);
            if (hit != ctx->batch_tunings.end()) {
                g->tile_log = hit->second.first;
                g->max_nb = hit->second.second;
            } else {
                double best = 1e30;
                const std::pair<uint32_t, uint32_t> plans[] = {{13, 0}, {13, 7}, {12, 0}, {12, 6}};
                for (const auto &pl : plans) {
                    Encoder *enc = nullptr;
                    if ((st = get_encoder(ctx, g->log_k, g->log_r, 0, &enc, pl.first, pl.second))) return st.code;
                    g->enc = enc;
                    double ms = 0;
                    if ((st = time_group(b.get(), *g, &ms))) return st.code;
                    b->tune_ms.push_back(ms);
                    if (ms < best) {
                        best = ms;
                        g->tile_log = pl.first;
                        g->max_nb = pl.second;
                        g->ms = ms;
                    }
                }
                ctx->batch_tunings[key] = {g->tile_log, g->max_nb};
            }
            if ((st = get_encoder(ctx, g->log_k, g->log_r, 0, &g->enc, g->tile_log, g->max_nb))) return st.code;
        }
        if (!(flags & TSR_BATCH_SERIAL)) {
            e = hipStreamCreateWithFlags(&g->side, hipStreamNonBlocking);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&g->done, hipEventDisableTiming);
            if (e != hipSuccess) return hip_fail(e, "branch stream").code;
        }
        b->groups.push_back(std::move(g));
    }
    // chain blocks over every column of every matrix
    {
        std::vector<BlockRef> refs;
        for (uint32_t i = 0; i < count; i++)
            for (uint64_t f = 0; f < b->hmats[i].n; f += 256) refs.push_back(BlockRef{i, 0, f});
        b->chain_blocks = (uint32_t)refs.size();
        if ((st = b->chain_refs.alloc(ctx, refs.size()))) return st.code;
        e = hipMemcpyAsync(b->chain_refs.ptr, refs.data(), sizeof(BlockRef) * refs.size(),
                           hipMemcpyHostToDevice, s);
        if (e != hipSuccess) return hip_fail(e, "upload").code;
    }
    // Merkle level blocks: level l of every matrix with n >= 2^l
    uint32_t max_levels = 0;
    for (uint32_t i = 0; i < count; i++)
        max_levels = std::max<uint32_t>(max_levels, (uint32_t)__builtin_ctzll(b->hmats[i].n));
    b->level_refs.resize(max_levels);
    b->level_blocks.resize(max_levels);
    for (uint32_t l = 1; l <= max_levels; l++) {
        std::vector<BlockRef> refs;
        for (uint32_t i = 0; i < count; i++) {
            const uint64_t cnt = b->hmats[i].n >> l;
            for (uint64_t f = 0; f < cnt; f += 256) refs.push_back(BlockRef{i, 0, f});
        }
        b->level_blocks[l - 1] = (uint32_t)refs.size();
        if ((st = b->level_refs[l - 1].alloc(ctx, std::max<size_t>(1, refs.size())))) return st.code;
        if (!refs.empty()) {
            e = hipMemcpyAsync(b->level_refs[l - 1].ptr, refs.data(), sizeof(BlockRef) * refs.size(),
                               hipMemcpyHostToDevice, s);
            if (e != hipSuccess) return hip_fail(e, "upload").code;
        }
    }
    if ((st = b->roots.alloc(ctx, 8 * count))) return st.code;
    if (!(flags & TSR_BATCH_SERIAL)) {
        e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
        if (e != hipSuccess) return hip_fail(e, "fork event").code;
    }
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    *out = b.release();
    return TSR_OK;
}

tsr_status tsr_commit_batch_run(tsr_batch *b) {
    if (!b) return Status::fail(TSR_BAD_ARG, "null batch").code;
    Ctx *ctx = b->ctx;
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    hipStream_t s = ctx->stream;
    if (!(b->flags & TSR_BATCH_GRAPH)) {
        KernelTimer timer(ctx, "commit_batch", s);
        Status st = enqueue(b, s);
        if (st) return st.code;
    } else {
        if (!b->exec) {
            // capture the launch sequence once (the profiler's event records stay out of it)
            const bool prof = ctx->profile;
            ctx->profile = false;
            e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
            if (e != hipSuccess) {
                ctx->profile = prof;
                return hip_fail(e, "hipStreamBeginCapture").code;
            }
            Status st = enqueue(b, s);
            hipGraph_t g = nullptr;
            hipError_t e2 = hipStreamEndCapture(s, &g);
            ctx->profile = prof;
            if (st) {
                if (g) (void)hipGraphDestroy(g);
                return st.code;
            }
            if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture").code;
            b->graph = g;
            e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
            if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate").code;
        }
        KernelTimer timer(ctx, "commit_batch", s);
        e = hipGraphLaunch(b->exec, s);
        if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch").code;
    }
    e = tsr::stream_sync(s);
    if (e != hipSuccess) return hip_fail(e, "sync").code;
    return TSR_OK;
}

tsr_status tsr_commit_batch_roots(const tsr_batch *b, uint8_t *roots) {
    if (!b || !roots) return Status::fail(TSR_BAD_ARG, "null argument").code;
    hipError_t e = hipSetDevice(b->ctx->device);
    if (e == hipSuccess) e = hipMemcpyAsync(roots, b->roots.ptr, 32 * b->count, hipMemcpyDeviceToHost, b->ctx->stream);
    if (e == hipSuccess) e = tsr::stream_sync(b->ctx->stream);
    if (e != hipSuccess) return hip_fail(e, "roots").code;
    return TSR_OK;
}

tsr_status tsr_commit_batch_tuning(const tsr_batch *b, uint32_t *plans, uint32_t cap,
                                   uint32_t *n_groups) {
    if (!b) return Status::fail(TSR_BAD_ARG, "null batch").code;
    if (n_groups) *n_groups = (uint32_t)b->groups.size();
    for (uint32_t g = 0; g < b->groups.size() && g < cap && plans; g++) {
        const tsr_batch::Group &gr = *b->groups[g];
        plans[5 * g] = gr.log_k;
        plans[5 * g + 1] = gr.log_r;
        plans[5 * g + 2] = gr.tile_log;
        plans[5 * g + 3] = gr.max_nb;
        plans[5 * g + 4] = (uint32_t)(gr.ms * 1000.0);  // the tuned plan's encode time (us)
    }
    return TSR_OK;
}

void tsr_commit_batch_free(tsr_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    (void)tsr::stream_sync(b->ctx->stream);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto &g : b->groups) {
        if (g->side) {
            (void)tsr::stream_sync(g->side);
            (void)hipStreamDestroy(g->side);
        }
        if (g->done) (void)hipEventDestroy(g->done);
    }
    if (b->fork) (void)hipEventDestroy(b->fork);
    delete b;
}

}  // extern "C"
