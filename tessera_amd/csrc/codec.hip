// Wire formats of the proving path, on the host: the canonical encoding (TCE, RFC-0003;
// crates/tessera-codec/src/tce.rs:306-441) of SumcheckProof (crates/tessera-sumcheck/src/
// proof.rs:9-24: rounds as Vec<Vec<Fp2>>, then final_evals as Vec<Fp2>) with its decoder's checks
// in the RFC's order and offsets, and the commitment descriptors (crates/tessera-pcs/src/
// descriptor.rs:10-160) whose hash under "commitment-id" starts every column chain.
//
// TCE: integers little-endian at fixed width; Vec<T> = u64 length, then the items; a struct its
// fields in order; an enum a u16 tag then its fields; Fp a canonical u64 (< p); Fp2 (c0, c1).
// Decoding checks, front to back: the input against max_total_bytes (offset 0); each composite
// value's depth before its bytes (LIMIT at its first byte); each length against max_vec_len
// (LIMIT), then length x min_size against the bytes left (TRUNCATED), at the length's first
// byte; each Fp below p (NONCANONICAL at the coefficient's first byte); leftover bytes
// (NONCANONICAL at the first of them).
#include <stdint.h>
#include <string.h>
#include <string>
#include <vector>
#include "common.h"
#include "field.h"
#include "transcript.h"

namespace tsr {

enum CodecError : int {
    CODEC_OK = 0,
    CODEC_MAGIC = 1,
    CODEC_VERSION = 2,
    CODEC_NONCANONICAL = 3,
    CODEC_TRUNCATED = 4,
    CODEC_DIGEST = 5,
    CODEC_LIMIT = 6,
};

static const char *codec_name(int e) {
    static const char *const names[] = {"OK", "ERR-CODEC-MAGIC", "ERR-CODEC-VERSION",
                                        "ERR-CODEC-NONCANONICAL", "ERR-CODEC-TRUNCATED",
                                        "ERR-CODEC-DIGEST", "ERR-CODEC-LIMIT"};
    return e >= 0 && e <= 6 ? names[e] : "?";
}

struct Reader {
    const uint8_t *p;
    uint64_t len, at = 0;
    uint64_t max_vec_len;
    uint32_t max_depth, depth = 0;
    int err = CODEC_OK;
    uint64_t err_at = 0;

    bool fail(int e, uint64_t off) {
        if (err == CODEC_OK) {
            err = e;
            err_at = off;
        }
        return false;
    }
    bool enter() {
        if (depth + 1 > max_depth) return fail(CODEC_LIMIT, at);
        depth++;
        return true;
    }
    void leave() { depth--; }
    bool u64(uint64_t &v) {
        if (len - at < 8) return fail(CODEC_TRUNCATED, at);
        v = 0;
        for (int i = 0; i < 8; i++) v |= (uint64_t)p[at + i] << (8 * i);
        at += 8;
        return true;
    }
    // a sequence's length, checked against the limit and the bytes left
    bool seq(uint64_t &n, uint64_t min_size) {
        const uint64_t start = at;
        if (!u64(n)) return false;
        if (n > max_vec_len) return fail(CODEC_LIMIT, start);
        if (n > (len - at) / min_size) return fail(CODEC_TRUNCATED, start);
        return true;
    }
    bool fp(uint64_t &v) {
        const uint64_t start = at;
        if (!u64(v)) return false;
        if (v >= gl::P) return fail(CODEC_NONCANONICAL, start);
        return true;
    }
    bool fp2(uint64_t out[2]) {
        if (!enter()) return false;
        if (!fp(out[0]) || !fp(out[1])) return false;
        leave();
        return true;
    }
    bool u8(uint8_t &v) {
        if (len - at < 1) return fail(CODEC_TRUNCATED, at);
        v = p[at++];
        return true;
    }
    bool u32(uint32_t &v) {
        if (len - at < 4) return fail(CODEC_TRUNCATED, at);
        v = 0;
        for (int i = 0; i < 4; i++) v |= (uint32_t)p[at + i] << (8 * i);
        at += 4;
        return true;
    }
    // a WeightMatrix id inside a by-matrix sequence: below 10 and above the previous entry's
    // (crates/tessera-protocol/src/matrix.rs, RFC-0036 'Matrix ids')
    bool matrix_id(uint8_t &v, int previous) {
        const uint64_t start = at;
        if (!u8(v)) return false;
        if (v >= 10 || (int)v <= previous) return fail(CODEC_NONCANONICAL, start);
        return true;
    }
    // Vec<Vec<Fp>>: the columns of an answer, each column's symbols after its length
    bool columns(std::vector<std::vector<uint64_t>> &out) {
        if (!enter()) return false;
        uint64_t n;
        if (!seq(n, 8)) return false;
        out.resize(n);
        for (uint64_t c = 0; c < n; c++) {
            if (!enter()) return false;
            uint64_t m;
            if (!seq(m, 8)) return false;
            out[c].resize(m);
            for (uint64_t i = 0; i < m; i++)
                if (!fp(out[c][i])) return false;
            leave();
        }
        leave();
        return true;
    }
    // MultiProof { siblings: Vec<Digest> }, a Digest a [u8; 32]
    bool multiproof(std::vector<uint8_t> &sib) {
        if (!enter() || !enter()) return false;
        uint64_t n;
        if (!seq(n, 32)) return false;
        sib.resize(32 * n);
        for (uint64_t i = 0; i < n; i++) {
            if (!enter()) return false;
            memcpy(&sib[32 * i], p + at, 32);  // seq checked that n digests fit
            at += 32;
            leave();
        }
        leave();
        leave();
        return true;
    }
    bool vec_u32(std::vector<uint32_t> &out) {
        if (!enter()) return false;
        uint64_t n;
        if (!seq(n, 4)) return false;
        out.resize(n);
        for (uint64_t i = 0; i < n; i++)
            if (!u32(out[i])) return false;
        leave();
        return true;
    }
    bool vec_fp2(std::vector<uint64_t> &out) {
        if (!enter()) return false;
        uint64_t n;
        if (!seq(n, 16)) return false;
        out.resize(2 * n);
        for (uint64_t i = 0; i < n; i++)
            if (!fp2(&out[2 * i])) return false;
        leave();
        return true;
    }
};

static void put64(std::vector<uint8_t> &b, uint64_t v) {
    for (int i = 0; i < 8; i++) b.push_back((uint8_t)(v >> (8 * i)));
}
static void put32(std::vector<uint8_t> &b, uint32_t v) {
    for (int i = 0; i < 4; i++) b.push_back((uint8_t)(v >> (8 * i)));
}

// Columns (Vec<Vec<Fp>>, column after column) then a MultiProof, as Round3Response and
// MatrixColumns carry them.
static Status put_columns_proof(std::vector<uint8_t> &b, const uint64_t *values, uint64_t columns,
                                uint64_t symbols, const uint8_t *proof, uint64_t siblings) {
    if ((columns && symbols && !values) || (siblings && !proof)) return Status::fail(TSR_BAD_ARG, "null argument");
    put64(b, columns);
    for (uint64_t c = 0; c < columns; c++) {
        put64(b, symbols);
        for (uint64_t i = 0; i < symbols; i++) {
            const uint64_t v = values[c * symbols + i];
            if (v >= gl::P) return Status::fail(TSR_BAD_ARG, "a symbol is not canonical");
            put64(b, v);
        }
    }
    put64(b, siblings);
    b.insert(b.end(), proof, proof + 32 * siblings);
    return Status::ok();
}

static tsr_status emit(const std::vector<uint8_t> &b, uint8_t *out, uint64_t cap, uint64_t *out_len) {
    if (out_len) *out_len = b.size();
    if (!out || cap < b.size())
        return Status::fail(TSR_BAD_SHAPE, "the message needs " + std::to_string(b.size()) + " bytes").code;
    memcpy(out, b.data(), b.size());
    return TSR_OK;
}

static tsr_status decoded(Reader &r, bool ok, int *codec_error, uint64_t *error_offset) {
    if (codec_error) *codec_error = r.err;
    if (error_offset) *error_offset = r.err_at;
    if (!ok)
        return Status::fail(TSR_BAD_SHAPE, std::string(codec_name(r.err)) + " at offset " + std::to_string(r.err_at)).code;
    return TSR_OK;
}

}  // namespace tsr

using namespace tsr;

extern "C" {

tsr_status tsr_sumcheck_proof_encode(const uint64_t *rounds, uint32_t n_rounds, uint32_t degree,
                                     const uint64_t *finals, uint64_t n_finals, uint8_t *out,
                                     uint64_t cap, uint64_t *out_len) {
    if ((n_rounds && degree && !rounds) || (n_finals && !finals))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    const uint64_t need = 8 + (uint64_t)n_rounds * (8 + 16ull * degree) + 8 + 16 * n_finals;
    if (out_len) *out_len = need;
    if (!out || cap < need) return Status::fail(TSR_BAD_SHAPE, "the proof needs " + std::to_string(need) + " bytes").code;
    for (uint64_t i = 0; i < 2ull * n_rounds * degree; i++)
        if (rounds[i] >= gl::P) return Status::fail(TSR_BAD_ARG, "a round value is not canonical").code;
    for (uint64_t i = 0; i < 2 * n_finals; i++)
        if (finals[i] >= gl::P) return Status::fail(TSR_BAD_ARG, "a final evaluation is not canonical").code;
    std::vector<uint8_t> b;
    b.reserve(need);
    put64(b, n_rounds);
    for (uint32_t j = 0; j < n_rounds; j++) {
        put64(b, degree);
        for (uint32_t z = 0; z < 2 * degree; z++) put64(b, rounds[(uint64_t)j * 2 * degree + z]);
    }
    put64(b, n_finals);
    for (uint64_t i = 0; i < 2 * n_finals; i++) put64(b, finals[i]);
    memcpy(out, b.data(), b.size());
    return TSR_OK;
}

tsr_status tsr_sumcheck_proof_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                                     uint32_t *n_rounds, uint32_t *round_lens, uint32_t round_lens_cap,
                                     uint64_t *values, uint64_t values_cap, uint64_t *n_values,
                                     uint64_t *finals, uint64_t finals_cap, uint64_t *n_finals,
                                     int *codec_error, uint64_t *error_offset) {
    if ((len && !in) || !limits) return Status::fail(TSR_BAD_ARG, "null argument").code;
    Reader r{in, len};
    r.max_vec_len = limits->max_vec_len;
    r.max_depth = limits->max_depth > 256 ? 256 : limits->max_depth;
    std::vector<std::vector<uint64_t>> rounds;
    std::vector<uint64_t> fin;
    bool ok = true;
    if (len > limits->max_total_bytes) {
        ok = r.fail(CODEC_LIMIT, 0);
    } else if ((ok = r.enter())) {  // SumcheckProof
        uint64_t n = 0;
        ok = r.enter() && r.seq(n, 8);  // rounds: Vec<Vec<Fp2>>
        for (uint64_t j = 0; ok && j < n; j++) {
            rounds.emplace_back();
            ok = r.vec_fp2(rounds.back());
        }
        if (ok) r.leave();
        ok = ok && r.vec_fp2(fin);  // final_evals
        if (ok) r.leave();
        if (ok && r.at != len) ok = r.fail(CODEC_NONCANONICAL, r.at);
    }
    if (codec_error) *codec_error = r.err;
    if (error_offset) *error_offset = r.err_at;
    if (!ok)
        return Status::fail(TSR_BAD_SHAPE, std::string(codec_name(r.err)) + " at offset " + std::to_string(r.err_at)).code;
    uint64_t total = 0;
    for (auto &v : rounds) total += v.size() / 2;
    if (n_rounds) *n_rounds = (uint32_t)rounds.size();
    if (n_values) *n_values = total;
    if (n_finals) *n_finals = fin.size() / 2;
    if (rounds.size() > round_lens_cap || total > values_cap || fin.size() / 2 > finals_cap ||
        (rounds.size() && !round_lens) || (total && !values) || (fin.size() && !finals))
        return Status::fail(TSR_BAD_SHAPE, "output capacity too small").code;
    uint64_t at = 0;
    for (size_t j = 0; j < rounds.size(); j++) {
        round_lens[j] = (uint32_t)(rounds[j].size() / 2);
        memcpy(values + 2 * at, rounds[j].data(), 8 * rounds[j].size());
        at += rounds[j].size() / 2;
    }
    if (!fin.empty()) memcpy(finals, fin.data(), 8 * fin.size());
    return TSR_OK;
}

// Round3Response (crates/tessera-protocol/src/round3.rs:100-175): advice_columns, advice_proof,
// weight_columns (Vec<MatrixColumns { matrix: u8, columns, proof }>).
tsr_status tsr_round3_encode(const uint64_t *advice_values, uint64_t advice_columns,
                             uint64_t advice_symbols, const uint8_t *advice_proof,
                             uint64_t advice_siblings, const tsr_matrix_columns *weights,
                             uint32_t n_weights, uint8_t *out, uint64_t cap, uint64_t *out_len) {
    if (n_weights && !weights) return Status::fail(TSR_BAD_ARG, "null argument").code;
    std::vector<uint8_t> b;
    Status st = put_columns_proof(b, advice_values, advice_columns, advice_symbols, advice_proof, advice_siblings);
    if (st) return st.code;
    put64(b, n_weights);
    for (uint32_t i = 0; i < n_weights; i++) {
        const tsr_matrix_columns &w = weights[i];
        if (w.matrix >= 10 || (i && w.matrix <= weights[i - 1].matrix))
            return Status::fail(TSR_BAD_ARG, "matrix ids name a WeightMatrix (0..9) and strictly increase").code;
        b.push_back(w.matrix);
        if ((st = put_columns_proof(b, w.values, w.columns, w.symbols, w.proof, w.siblings))) return st.code;
    }
    return emit(b, out, cap, out_len);
}

tsr_status tsr_round3_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                             uint64_t *advice_columns, uint32_t *n_matrices, int *codec_error,
                             uint64_t *error_offset) {
    if ((len && !in) || !limits) return Status::fail(TSR_BAD_ARG, "null argument").code;
    Reader r{in, len};
    r.max_vec_len = limits->max_vec_len;
    r.max_depth = limits->max_depth > 256 ? 256 : limits->max_depth;
    std::vector<std::vector<uint64_t>> cols;
    std::vector<uint8_t> sib;
    uint64_t nw = 0;
    bool ok = true;
    if (len > limits->max_total_bytes) {
        ok = r.fail(CODEC_LIMIT, 0);
    } else if ((ok = r.enter())) {
        ok = r.columns(cols) && r.multiproof(sib);
        // weight_columns: MatrixColumns of at least u8 + two lengths
        ok = ok && r.enter() && r.seq(nw, 1 + 8 + 8);
        int prev = -1;
        for (uint64_t i = 0; ok && i < nw; i++) {
            std::vector<std::vector<uint64_t>> wc;
            std::vector<uint8_t> ws;
            uint8_t id = 0;
            ok = r.enter() && r.matrix_id(id, prev) && r.columns(wc) && r.multiproof(ws);
            if (ok) r.leave();
            prev = id;
        }
        if (ok) r.leave();
        if (ok) r.leave();
        if (ok && r.at != len) ok = r.fail(CODEC_NONCANONICAL, r.at);
    }
    if (ok) {
        if (advice_columns) *advice_columns = cols.size();
        if (n_matrices) *n_matrices = (uint32_t)nw;
    }
    return decoded(r, ok, codec_error, error_offset);
}

// ColumnQueries (round3.rs:19-95): advice: Vec<u32>, weights: Vec<MatrixQueries { matrix: u8,
// indices: Vec<u32> }>.
tsr_status tsr_column_queries_encode(const uint32_t *advice, uint64_t n_advice, const uint8_t *matrices,
                                     const uint64_t *counts, const uint32_t *indices, uint32_t n_matrices,
                                     uint8_t *out, uint64_t cap, uint64_t *out_len) {
    if ((n_advice && !advice) || (n_matrices && (!matrices || !counts)))
        return Status::fail(TSR_BAD_ARG, "null argument").code;
    std::vector<uint8_t> b;
    put64(b, n_advice);
    for (uint64_t i = 0; i < n_advice; i++) put32(b, advice[i]);
    put64(b, n_matrices);
    uint64_t at = 0;
    for (uint32_t m = 0; m < n_matrices; m++) {
        if (matrices[m] >= 10 || (m && matrices[m] <= matrices[m - 1]))
            return Status::fail(TSR_BAD_ARG, "matrix ids name a WeightMatrix (0..9) and strictly increase").code;
        if (counts[m] && !indices) return Status::fail(TSR_BAD_ARG, "null indices").code;
        b.push_back(matrices[m]);
        put64(b, counts[m]);
        for (uint64_t i = 0; i < counts[m]; i++) put32(b, indices[at + i]);
        at += counts[m];
    }
    return emit(b, out, cap, out_len);
}

tsr_status tsr_column_queries_decode(const uint8_t *in, uint64_t len, const tsr_decode_limits *limits,
                                     uint32_t *advice, uint64_t advice_cap, uint64_t *n_advice,
                                     uint8_t *matrices, uint64_t *counts, uint32_t matrices_cap,
                                     uint32_t *n_matrices, uint32_t *indices, uint64_t indices_cap,
                                     uint64_t *n_indices, int *codec_error, uint64_t *error_offset) {
    if ((len && !in) || !limits) return Status::fail(TSR_BAD_ARG, "null argument").code;
    Reader r{in, len};
    r.max_vec_len = limits->max_vec_len;
    r.max_depth = limits->max_depth > 256 ? 256 : limits->max_depth;
    std::vector<uint32_t> adv, idx;
    std::vector<uint8_t> ids;
    std::vector<uint64_t> cnt;
    bool ok = true;
    if (len > limits->max_total_bytes) {
        ok = r.fail(CODEC_LIMIT, 0);
    } else if ((ok = r.enter())) {
        uint64_t nw = 0;
        ok = r.vec_u32(adv) && r.enter() && r.seq(nw, 1 + 8);
        int prev = -1;
        for (uint64_t i = 0; ok && i < nw; i++) {
            uint8_t id = 0;
            std::vector<uint32_t> q;
            ok = r.enter() && r.matrix_id(id, prev) && r.vec_u32(q);
            if (!ok) break;
            r.leave();
            prev = id;
            ids.push_back(id);
            cnt.push_back(q.size());
            idx.insert(idx.end(), q.begin(), q.end());
        }
        if (ok) r.leave();
        if (ok) r.leave();
        if (ok && r.at != len) ok = r.fail(CODEC_NONCANONICAL, r.at);
    }
    tsr_status s = decoded(r, ok, codec_error, error_offset);
    if (s) return s;
    if (n_advice) *n_advice = adv.size();
    if (n_matrices) *n_matrices = (uint32_t)ids.size();
    if (n_indices) *n_indices = idx.size();
    if (adv.size() > advice_cap || ids.size() > matrices_cap || idx.size() > indices_cap ||
        (adv.size() && !advice) || (ids.size() && (!matrices || !counts)) || (idx.size() && !indices))
        return Status::fail(TSR_BAD_SHAPE, "output capacity too small").code;
    if (!adv.empty()) memcpy(advice, adv.data(), 4 * adv.size());
    if (!ids.empty()) {
        memcpy(matrices, ids.data(), ids.size());
        memcpy(counts, cnt.data(), 8 * cnt.size());
    }
    if (!idx.empty()) memcpy(indices, idx.data(), 4 * idx.size());
    return TSR_OK;
}

tsr_status tsr_advice_commitment_id(const tsr_advice_descriptor *d, uint8_t tce_out[60],
                                    uint8_t commitment_id[32]) {
    if (!d || !commitment_id) return Status::fail(TSR_BAD_ARG, "null argument").code;
    std::vector<uint8_t> b = {0, 0};  // CommitmentDescriptor::Advice
    b.insert(b.end(), d->statement_digest, d->statement_digest + 32);
    put32(b, d->unit);
    b.push_back(d->log2_k);
    b.push_back(d->log2_inv_rate);
    put32(b, d->mask_coeffs);
    put32(b, d->row_block);
    put32(b, d->segments);
    put64(b, d->rows);
    if (tce_out) memcpy(tce_out, b.data(), 60);
    hash_domain(nullptr, DOM_COMMITMENT_ID, b.data(), b.size(), commitment_id);
    return TSR_OK;
}

tsr_status tsr_weight_matrix_commitment_id(const tsr_weight_layout_params *layout, uint32_t matrix,
                                           uint8_t log2_k, uint32_t row_block, uint64_t rows,
                                           uint32_t epoch, uint8_t tce_out[31],
                                           uint8_t commitment_id[32]) {
    if (!layout || !commitment_id) return Status::fail(TSR_BAD_ARG, "null argument").code;
    if (matrix >= 10) return Status::fail(TSR_BAD_SHAPE, "a WeightMatrix is 0 (Embed) to 9 (ChannelParams)").code;
    std::vector<uint8_t> b = {1, 0};  // CommitmentDescriptor::WeightMatrix
    b.push_back(layout->log2_k);
    b.push_back(layout->log2_inv_rate);
    put32(b, layout->mask_coeffs);
    put32(b, layout->blinding_rows);
    b.push_back((uint8_t)matrix);
    b.push_back(0);
    b.push_back(log2_k);
    put32(b, row_block);
    put64(b, rows);
    put32(b, epoch);
    if (tce_out) memcpy(tce_out, b.data(), 31);
    hash_domain(nullptr, DOM_COMMITMENT_ID, b.data(), b.size(), commitment_id);
    return TSR_OK;
}

}  // extern "C"
