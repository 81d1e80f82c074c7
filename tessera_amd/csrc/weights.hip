// The weight commitment (RFC-0040) on gfx950: WeightLayout::new and WeightCommitment::commit
// (crates/tessera-pcs/src/weights/layout.rs:186-260,466-560; weights/commit.rs:61-135), the
// descriptor of crates/tessera-pcs/src/descriptor.rs:118-160 and the public artifact of
// weights/public.rs:22-48 sealed as an envelope (crates/tessera-codec/src/envelope.rs).
//
// A weight type T of out x in INT8 entries per layer pads out, in and its layers to powers of two,
// has the message width k_T = max(in_pad, min(k_W, out_pad in_pad)), a_T = log2(k_T / in_pad),
// h_T = log2 out_pad - a_T, and puts W_l[o][i] at row l 2^h + o / 2^a, column (o mod 2^a) in_pad
// + i. Its committed rows are the rows that hold an output, m = l ceil(out / 2^a) + o_hi in
// increasing row order. The rows are assembled on the device from the INT8 entries (one byte per
// weight crosses PCIe, the field elements exist only in HBM), negative weights as p + w, and
// committed unmasked (rate 2^-r, chains of B rows, Merkle root) by the commit path. ChannelParams,
// a few thousand entries, is laid out on the host.
#include <hip/hip_runtime.h>
#include <algorithm>
#include <memory>
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "common.h"
#include "field.h"
#include "transcript.h"

struct tsr_weight_commitment {
    tsr::Ctx *ctx = nullptr;
    tsr_commitment *matrices[10] = {nullptr};
    uint32_t log2_k[10] = {0}, row_vars[10] = {0};
    uint64_t rows[10] = {0};
    uint8_t public_tce[334];
    uint8_t weights_root[32];
};

namespace tsr {

static const char *const MATRIX_NAMES[10] = {"Embed", "Q",    "K",  "V",      "O",
                                             "Gate",  "Up",   "Down", "LmHead", "ChannelParams"};

static bool per_layer(uint32_t t) { return t >= 1 && t <= 7; }

static uint32_t log2_pad(uint64_t x) {
    uint32_t l = 0;
    while ((1ull << l) < x) l++;
    return l;
}

// Committed rows of weight matrix row m: l = m / per_layer, o_hi = m % per_layer; column
// c = (o_lo, i). Entries past out or in are padding (0).
__global__ void k_weight_rows(const int8_t *__restrict__ w, uint32_t out, uint32_t in,
                              uint32_t log2_in_pad, uint32_t a, uint64_t per_layer,
                              uint32_t log2_k, uint64_t total, uint64_t *__restrict__ dst) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += stride) {
        const uint64_t m = e >> log2_k, c = e & ((1ull << log2_k) - 1);
        const uint64_t l = m / per_layer, oh = m - l * per_layer;
        const uint64_t ol = c >> log2_in_pad, i = c & ((1ull << log2_in_pad) - 1);
        const uint64_t o = (oh << a) + ol;
        uint64_t v = 0;
        if (o < out && i < in) {
            const int64_t x = w[(l * out + o) * in + i];
            v = x < 0 ? gl::P - (uint64_t)(-x) : (uint64_t)x;
        }
        dst[e] = v;
    }
}

struct MatrixShape {
    uint32_t log2_k = 0, row_vars = 0, a = 0, h = 0, layers = 1;
    uint64_t rows = 0;
};

static void put_u32(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
static void put_u64(std::vector<uint8_t> &b, uint64_t v) {
    for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
static void put_layout(std::vector<uint8_t> &b, const tsr_weight_layout_params &p) {
    b.push_back(p.log2_k);
    b.push_back(p.log2_inv_rate);
    put_u32(b, p.mask_coeffs);
    put_u32(b, p.blinding_rows);
}

static std::string weight_name(uint32_t t, uint32_t layer) {
    return std::string(t < 10 ? MATRIX_NAMES[t] : "?") + " layer " + std::to_string(layer);
}

static Status check_code(const tsr_weight_layout_params &p, uint32_t log2_k) {
    if (p.log2_inv_rate == 0) return Status::fail(TSR_BAD_SHAPE, "InverseRate: log2_inv_rate is 0");
    if (log2_k + p.log2_inv_rate > 32)
        return Status::fail(TSR_BAD_SHAPE, "TooLong: log_len " + std::to_string(log2_k + p.log2_inv_rate));
    return Status::ok();
}

static bool has_weight(uint32_t layers, uint32_t t, uint32_t layer) {
    if (t >= 9) return false;
    return per_layer(t) ? layer < layers : layer == 0;
}

Status weight_commit(Ctx *ctx, const tsr_weight_layout_params &params, uint32_t layers,
                     const tsr_weight_shape *shapes, const uint32_t *cw_pairs, uint32_t ncw,
                     uint32_t row_block, uint32_t epoch, const tsr_weight_tensor *tensors,
                     uint32_t ntensors, const tsr_channel_params *channels, uint32_t nchannels,
                     tsr_weight_commitment *wc) {
    // ---- WeightLayout::new (layout.rs:186-260): the nine weight types, then ChannelParams ----
    MatrixShape ms[10];
    for (uint32_t t = 0; t < 9; t++) {
        const tsr_weight_shape &s = shapes[t];
        if (s.out == 0 || s.inputs == 0) return Status::fail(TSR_BAD_SHAPE, std::string("EmptyWeight: ") + MATRIX_NAMES[t]);
        const uint32_t lo = log2_pad(s.out), li = log2_pad(s.inputs);
        const uint32_t log2_k = std::max(li, std::min<uint32_t>(params.log2_k, lo + li));
        Status st = check_code(params, log2_k);
        if (st) return st;
        ms[t].layers = per_layer(t) ? layers : 1;
        ms[t].log2_k = log2_k;
        ms[t].a = log2_k - li;
        ms[t].h = lo - ms[t].a;
        ms[t].row_vars = log2_pad(ms[t].layers) + ms[t].h;
        ms[t].rows = (uint64_t)ms[t].layers * (((uint64_t)s.out + (1ull << ms[t].a) - 1) >> ms[t].a);
    }
    std::vector<std::pair<uint32_t, uint32_t>> cws;
    for (uint32_t i = 0; i < ncw; i++) {
        const uint32_t t = cw_pairs[2 * i], l = cw_pairs[2 * i + 1];
        if (!has_weight(layers, t, l))
            return Status::fail(TSR_BAD_SHAPE, "WeightPlace: ChannelParams of " + weight_name(t, l));
        cws.push_back({t, l});
    }
    std::sort(cws.begin(), cws.end());
    for (size_t i = 1; i < cws.size(); i++)
        if (cws[i] == cws[i - 1])
            return Status::fail(TSR_BAD_SHAPE, "WeightTwice: ChannelParams of " + weight_name(cws[i].first, cws[i].second));
    uint32_t log2_c = 0;
    for (auto &p : cws) log2_c = std::max(log2_c, log2_pad(shapes[p.first].out));
    const uint32_t log2_lpad = log2_pad(layers);
    const uint32_t log2_total = 4 + log2_lpad + log2_c;  // 16 channel types
    const uint32_t log2_kcp = std::min<uint32_t>(params.log2_k, log2_total);
    Status st = check_code(params, log2_kcp);
    if (st) return st;
    ms[9].log2_k = log2_kcp;
    ms[9].row_vars = log2_total - log2_kcp;
    if (ms[9].row_vars > 64) return Status::fail(TSR_BAD_SHAPE, "RowVars");
    // committed rows of ChannelParams: every row that holds a live position
    std::vector<uint64_t> cp_rows;
    for (auto &p : cws) {
        const uint64_t f0 = (((uint64_t)p.first << log2_lpad) + p.second) << log2_c;
        const uint64_t f1 = f0 + shapes[p.first].out - 1;
        for (uint64_t r = f0 >> log2_kcp; r <= (f1 >> log2_kcp); r++) cp_rows.push_back(r);
    }
    std::sort(cp_rows.begin(), cp_rows.end());
    cp_rows.erase(std::unique(cp_rows.begin(), cp_rows.end()), cp_rows.end());
    ms[9].rows = cp_rows.size();

    // ---- WeightLayout::check (layout.rs:466-560): tensors, then channel parameters ----
    std::vector<std::pair<std::pair<uint32_t, uint32_t>, const tsr_weight_tensor *>> tens;
    for (uint32_t i = 0; i < ntensors; i++) {
        const tsr_weight_tensor &w = tensors[i];
        if (!has_weight(layers, w.matrix, w.layer))
            return Status::fail(TSR_BAD_SHAPE, "WeightPlace: " + weight_name(w.matrix, w.layer));
        const uint64_t expected = (uint64_t)shapes[w.matrix].out * shapes[w.matrix].inputs;
        if (w.len != expected)
            return Status::fail(TSR_BAD_SHAPE, "WeightLength: " + weight_name(w.matrix, w.layer) + " holds " +
                                                   std::to_string(w.len) + " entries, not " + std::to_string(expected));
        if (!w.data) return Status::fail(TSR_BAD_ARG, "null weight data");
        tens.push_back({{w.matrix, w.layer}, &w});
    }
    std::sort(tens.begin(), tens.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (size_t i = 1; i < tens.size(); i++)
        if (tens[i].first == tens[i - 1].first)
            return Status::fail(TSR_BAD_SHAPE, "WeightTwice: " + weight_name(tens[i].first.first, tens[i].first.second));
    {
        size_t at = 0;
        for (uint32_t t = 0; t < 9; t++)
            for (uint32_t l = 0; l < ms[t].layers; l++, at++)
                if (at >= tens.size() || tens[at].first != std::make_pair(t, l))
                    return Status::fail(TSR_BAD_SHAPE, "WeightMissing: " + weight_name(t, l));
    }
    std::vector<std::pair<std::pair<uint32_t, uint32_t>, const tsr_channel_params *>> chs;
    for (uint32_t i = 0; i < nchannels; i++) {
        const tsr_channel_params &c = channels[i];
        if (!std::binary_search(cws.begin(), cws.end(), std::make_pair(c.matrix, c.layer)))
            return Status::fail(TSR_BAD_SHAPE, "WeightPlace: ChannelParams of " + weight_name(c.matrix, c.layer));
        if (c.count != shapes[c.matrix].out)
            return Status::fail(TSR_BAD_SHAPE, "WeightLength: ChannelParams of " + weight_name(c.matrix, c.layer));
        if (!c.values || c.kind > 1) return Status::fail(TSR_BAD_ARG, "null channel values or unknown kind");
        chs.push_back({{c.matrix, c.layer}, &c});
    }
    std::sort(chs.begin(), chs.end(), [](const auto &x, const auto &y) { return x.first < y.first; });
    for (size_t i = 1; i < chs.size(); i++)
        if (chs[i].first == chs[i - 1].first)
            return Status::fail(TSR_BAD_SHAPE, "WeightTwice: ChannelParams of " + weight_name(chs[i].first.first, chs[i].first.second));
    for (size_t i = 0; i < cws.size(); i++)
        if (i >= chs.size() || chs[i].first != cws[i])
            return Status::fail(TSR_BAD_SHAPE, "WeightMissing: ChannelParams of " + weight_name(cws[i].first, cws[i].second));

    // ---- WeightCommitment::commit (commit.rs:61-135) ----
    hipStream_t s = ctx->stream;
    uint8_t roots[10][32];
    for (uint32_t t = 0; t < 10; t++) {
        const uint64_t k = 1ull << ms[t].log2_k, m = ms[t].rows;
        tsr_commitment_params cp{};
        cp.log2_k = ms[t].log2_k;
        cp.log2_inv_rate = params.log2_inv_rate;
        cp.mask_coeffs = 0;
        cp.row_block = row_block;
        tsr_weight_matrix_commitment_id(&params, t, (uint8_t)ms[t].log2_k, row_block, m, epoch, nullptr,
                                        cp.commitment_id);
        DevBuf<uint64_t> rows;
        if (m && (st = rows.alloc(ctx, m * k))) return st;
        if (t < 9 && m) {
            // the type's layers contiguously on the device, one byte per weight
            const uint64_t per = (uint64_t)shapes[t].out * shapes[t].inputs;
            DevBuf<uint64_t> w8;
            if ((st = w8.alloc(ctx, (ms[t].layers * per + 7) / 8))) return st;
            for (uint32_t l = 0; l < ms[t].layers; l++) {
                const tsr_weight_tensor *src = nullptr;
                for (auto &x : tens)
                    if (x.first == std::make_pair(t, l)) src = x.second;
                TSR_HIP(hipMemcpyAsync((int8_t *)w8.ptr + l * per, src->data, per, hipMemcpyDefault, s));
            }
            const uint64_t total = m * k;
            const unsigned blocks = (unsigned)std::min<uint64_t>((total + 255) / 256, 1u << 20);
            KernelTimer timer(ctx, "k_weight_rows", s);
            hipLaunchKernelGGL(k_weight_rows, dim3(blocks), dim3(256), 0, s, (const int8_t *)w8.ptr,
                               shapes[t].out, shapes[t].inputs, log2_pad(shapes[t].inputs), ms[t].a,
                               m / ms[t].layers, ms[t].log2_k, total, rows.ptr);
            if ((st = check_launch("weight rows"))) return st;
            TSR_HIP(tsr::stream_sync(s));  // w8 is freed at the end of this scope
        } else if (m) {
            std::vector<uint64_t> host(m * k, 0);
            for (auto &x : chs) {
                const uint64_t f0 = (((uint64_t)x.first.first << log2_lpad) + x.first.second) << log2_c;
                const tsr_channel_params &c = *x.second;
                for (uint64_t j = 0; j < c.count; j++) {
                    const uint64_t f = f0 + j, r = f >> log2_kcp;
                    const uint64_t row = std::lower_bound(cp_rows.begin(), cp_rows.end(), r) - cp_rows.begin();
                    host[row * k + (f & (k - 1))] =
                        c.kind == 0 ? ((const uint8_t *)c.values)[j] : ((const uint16_t *)c.values)[j];
                }
            }
            TSR_HIP(hipMemcpyAsync(rows.ptr, host.data(), 8 * m * k, hipMemcpyHostToDevice, s));
            TSR_HIP(tsr::stream_sync(s));
        }
        tsr_commitment *c = nullptr;
        tsr_status rc = tsr_commit_unmasked((tsr_ctx *)ctx, &cp, m ? rows.ptr : nullptr, m, &c);
        if (rc != TSR_OK) {  // the reason is already in the thread's last error
            Status f;
            f.code = rc;
            return f;
        }
        wc->matrices[t] = c;
        wc->log2_k[t] = ms[t].log2_k;
        wc->row_vars[t] = ms[t].row_vars;
        wc->rows[t] = m;
        if ((rc = tsr_commitment_root(c, roots[t])) != TSR_OK) {
            Status f;
            f.code = rc;
            return f;
        }
    }
    // ---- WeightCommitmentPublic (public.rs:22-48) and R_W ----
    std::vector<uint8_t> pub;
    put_layout(pub, params);
    for (uint32_t t = 0; t < 10; t++) pub.insert(pub.end(), roots[t], roots[t] + 32);
    put_u32(pub, epoch);
    memcpy(wc->public_tce, pub.data(), sizeof wc->public_tce);
    hash_domain(ctx, DOM_WEIGHTS_ROOT, pub.data(), pub.size(), wc->weights_root);
    return Status::ok();
}

// Envelope (RFC-0003): b"TSRA", u16 1, u16 schema, u16 version, u16 0, u64 length, payload,
// H("envelope", the preceding bytes).
static void seal(const Ctx *ctx, uint16_t schema, uint16_t version, const uint8_t *payload,
                 uint64_t len, uint8_t *out) {
    std::vector<uint8_t> b = {'T', 'S', 'R', 'A', 1, 0};
    b.push_back((uint8_t)schema);
    b.push_back((uint8_t)(schema >> 8));
    b.push_back((uint8_t)version);
    b.push_back((uint8_t)(version >> 8));
    b.push_back(0);
    b.push_back(0);
    put_u64(b, len);
    b.insert(b.end(), payload, payload + len);
    memcpy(out, b.data(), b.size());
    hash_domain(ctx, DOM_ENVELOPE, b.data(), b.size(), out + b.size());
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_weight_commit(tsr_ctx *ctx, const tsr_weight_layout_params *params, uint32_t layers,
                             const tsr_weight_shape shapes[9], const uint32_t *channel_weights,
                             uint32_t n_channel_weights, uint32_t row_block, uint32_t epoch,
                             const tsr_weight_tensor *tensors, uint32_t ntensors,
                             const tsr_channel_params *channels, uint32_t nchannels,
                             tsr_weight_commitment **out) {
    if (!ctx || !params || !shapes || !out || (n_channel_weights && !channel_weights) ||
        (ntensors && !tensors) || (nchannels && !channels))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (row_block == 0) return Status::fail(TSR_BAD_SHAPE, "the row block B is at least 1").code;
    Ctx *c = (Ctx *)ctx;
    hipError_t e = hipSetDevice(c->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice").code;
    std::unique_ptr<tsr_weight_commitment> wc(new tsr_weight_commitment());
    wc->ctx = c;
    Status st = weight_commit(c, *params, layers, shapes, channel_weights, n_channel_weights, row_block,
                              epoch, tensors, ntensors, channels, nchannels, wc.get());
    if (st) {
        for (auto *m : wc->matrices) tsr_commitment_free(m);
        return st.code;
    }
    *out = wc.release();
    return TSR_OK;
}

tsr_status tsr_weight_commitment_public(const tsr_weight_commitment *w, uint8_t out[334]) {
    if (!w || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    memcpy(out, w->public_tce, 334);
    return TSR_OK;
}

tsr_status tsr_weight_commitment_weights_root(const tsr_weight_commitment *w, uint8_t out[32]) {
    if (!w || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    memcpy(out, w->weights_root, 32);
    return TSR_OK;
}

tsr_status tsr_weight_commitment_artifact(const tsr_weight_commitment *w, uint8_t out[386]) {
    if (!w || !out) return Status::fail(TSR_BAD_ARG, "null argument").code;
    seal(w->ctx, 0x0301, 1, w->public_tce, 334, out);
    return TSR_OK;
}

tsr_status tsr_weight_commitment_matrix(const tsr_weight_commitment *w, uint32_t matrix,
                                        const tsr_commitment **commitment, uint32_t *log2_k,
                                        uint32_t *row_vars, uint64_t *rows) {
    if (!w) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (matrix >= 10) return Status::fail(TSR_BAD_SHAPE, "a WeightMatrix is 0 (Embed) to 9 (ChannelParams)").code;
    if (commitment) *commitment = w->matrices[matrix];
    if (log2_k) *log2_k = w->log2_k[matrix];
    if (row_vars) *row_vars = w->row_vars[matrix];
    if (rows) *rows = w->rows[matrix];
    return TSR_OK;
}

void tsr_weight_commitment_free(tsr_weight_commitment *w) {
    if (!w) return;
    for (auto *m : w->matrices) tsr_commitment_free(m);
    delete w;
}

tsr_status tsr_envelope_seal(tsr_ctx *ctx, uint16_t schema_id, uint16_t schema_version,
                             const uint8_t *payload, uint64_t len, uint8_t *out, uint64_t cap,
                             uint64_t *out_len) {
    if (!ctx || (len && !payload)) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (out_len) *out_len = len + 52;
    if (!out || cap < len + 52) return Status::fail(TSR_BAD_SHAPE, "the envelope needs payload + 52 bytes").code;
    seal((const Ctx *)ctx, schema_id, schema_version, payload, len, out);
    return TSR_OK;
}

}  // extern "C"
